// sn_autotune.cpp -- picks the sweep schedule (tile edge targets, fine
// colouring, sweep kernel) for a set of replicas by timing candidate schedules
// on the current device, and caches the winner in a text file.
//
// Uses only the public C ABI: every candidate is a real context
// (sn_create), loaded with a random lattice and timed over whole sweeps
// (sn_mc_sweeps + sn_sync), so the measured rate is the rate the caller gets.
// Candidates the kernels cannot run (LDS budget, untileable lattice) fail in
// sn_create and are skipped.
//
// Cache: one line per configuration, "key<TAB>t0 t1 t2 lanes colouring graph
// kernel block rate".  The key is the device architecture, DipoleCutOff, whether
// the species table gives mixed or uniform lengths (mixed lengths stage a fourth
// LDS array, which changes the best tile), and the multiset of replica sizes.
// Path: $SN_TUNE_CACHE, else $HOME/.cache/starrynight_amd/tune.txt; no cache
// when neither is set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/starrynight_amd.h"

namespace {

std::string cache_path()
{
    if (const char *p = std::getenv("SN_TUNE_CACHE")) return p;
    if (const char *h = std::getenv("HOME")) return std::string(h) + "/.cache/starrynight_amd/tune.txt";
    return "";
}

void make_parent_dirs(const std::string &path)
{
    for (size_t i = 1; i < path.size(); i++)
        if (path[i] == '/') ::mkdir(path.substr(0, i).c_str(), 0755);
}

std::string config_key(const sn_params *p, int n, bool mixed)
{
    hipDeviceProp_t prop;
    int dev = 0;
    std::string name = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        name = prop.gcnArchName;
    // replicas by geometry, with counts (order does not change the schedule's cost much)
    std::map<std::string, int> geo;
    for (int i = 0; i < n; i++) {
        std::ostringstream g;
        g << p[i].X << "x" << p[i].Y << "x" << p[i].Z;
        geo[g.str()]++;
    }
    std::ostringstream k;
    k << "v3|" << name << "|rc" << p[0].dipole_cutoff << (mixed ? "|mixed" : "|uniform");
    for (auto &kv : geo) k << "|" << kv.first << "*" << kv.second;
    return k.str();
}

bool cache_lookup(const std::string &key, sn_tuning *t, double *rate)
{
    const std::string path = cache_path();
    if (path.empty()) return false;
    std::ifstream f(path);
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos || line.compare(0, tab, key) != 0 || tab != key.size()) continue;
        std::istringstream v(line.substr(tab + 1));
        sn_tuning c{};
        double r = 0.0;
        if (v >> c.tile[0] >> c.tile[1] >> c.tile[2] >> c.lanes_per_site >> c.fine_colouring >> c.use_graph >>
            c.kernel >> c.block_threads >> r) {
            *t = c;       // the last entry for a key wins
            *rate = r;
            found = true;
        }
    }
    return found;
}

void cache_store(const std::string &key, const sn_tuning &t, double rate)
{
    const std::string path = cache_path();
    if (path.empty()) return;
    make_parent_dirs(path);
    std::ofstream f(path, std::ios::app);
    f << key << '\t' << t.tile[0] << ' ' << t.tile[1] << ' ' << t.tile[2] << ' ' << t.lanes_per_site << ' '
      << t.fine_colouring << ' ' << t.use_graph << ' ' << t.kernel << ' ' << t.block_threads << ' ' << rate
      << '\n';
}

// site-updates/s of one candidate, or < 0 when it cannot run
double time_candidate(const sn_params *p, int n, int nspec, const double *len, const double *prev,
                      const sn_tuning &t, double max_ms)
{
    sn_ctx *ctx = nullptr;
    if (sn_create(p, n, &t, &ctx) != SN_OK) {
        if (ctx) sn_destroy(ctx);
        return -1.0;
    }
    double sites = 0.0;
    for (int r = 0; r < n; r++) sites += (double)p[r].X * p[r].Y * p[r].Z;
    double rate = -1.0;
    bool ok = true;
    for (int r = 0; r < n && ok; r++) ok = sn_init_lattice_random(ctx, r, nspec, len, prev) == SN_OK;
    // warm up (graph capture, caches), then time whole sweeps until max_ms
    ok = ok && sn_mc_sweeps(ctx, 2) == SN_OK && sn_sync(ctx) == SN_OK;
    if (ok) {
        using clk = std::chrono::steady_clock;
        int done = 0;
        const auto t0 = clk::now();
        double ms = 0.0;
        int batch = 1;
        while (ok && (done < 4 || ms < max_ms) && done < 400) {
            ok = sn_mc_sweeps(ctx, batch) == SN_OK && sn_sync(ctx) == SN_OK;
            done += batch;
            ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
            if (ms < 0.25 * max_ms) batch = std::min(batch * 2, 64);
        }
        if (ok && ms > 0.0) rate = sites * done / (ms * 1e-3);
    }
    sn_destroy(ctx);
    return ok ? rate : -1.0;
}

}  // namespace

extern "C" int sn_autotune(const sn_params *params, int32_t n_replicas, int32_t n_species, const double *lengths,
                           const double *prevalence, int32_t budget_ms, int32_t use_cache, sn_tuning *best,
                           double *best_rate)
{
    if (!params || n_replicas < 1 || !best) return SN_ERR_INVALID;
    if (n_species < 0 || n_species > 10 || (n_species > 0 && (!lengths || !prevalence))) return SN_ERR_INVALID;
    // mixed: two species of non-zero prevalence with different lengths
    bool mixed = false;
    for (int i = 0; i < n_species; i++)
        for (int j = 0; j < i; j++)
            if (prevalence[i] > 0.0 && prevalence[j] > 0.0 && lengths[i] != lengths[j]) mixed = true;
    const std::string key = config_key(params, n_replicas, mixed);
    double rate = 0.0;
    if (use_cache && cache_lookup(key, best, &rate)) {
        if (best_rate) *best_rate = rate;
        return SN_OK;
    }
    const int rc = params[0].dipole_cutoff;
    const bool two_d = params[0].Z == 1;
    std::vector<std::vector<int>> tiles;
    if (two_d)
        tiles = {{8, 8, 1}, {12, 12, 1}, {16, 16, 1}, {16, 8, 1}, {24, 24, 1}};
    else
        tiles = {{8, 8, 8}, {8, 8, 6}, {8, 8, 10}, {6, 8, 8}, {6, 6, 6}, {8, 6, 6}, {4, 8, 8}, {8, 8, 4}, {4, 4, 8}};
    std::vector<int> colourings = (!two_d && rc == 3) ? std::vector<int>{3, 2, 1} : std::vector<int>{2, 1};
    std::vector<sn_tuning> cands;
    for (int lanes : {0, 8, 16}) {
        // whole-lattice sweeps (small replicas only; refused at creation otherwise)
        sn_tuning t{};
        t.kernel = 3;
        t.lanes_per_site = lanes;
        cands.push_back(t);
    }
    // (kernel, threads per tile): gather with 8 or 4 waves, field-update.  The
    // kernels vary fastest, then the tiles (most promising first), so a budget
    // that ends the search early has tried every kernel on the leading tiles.
    const int kb[3][2] = {{2, 512}, {2, 256}, {1, 0}};
    for (int c : colourings)
        for (auto &tl : tiles)
            for (const auto &v : kb) {
                sn_tuning t{};
                t.tile[0] = tl[0]; t.tile[1] = tl[1]; t.tile[2] = tl[2];
                t.fine_colouring = c;
                t.kernel = v[0];
                t.block_threads = v[1];
                cands.push_back(t);
            }
    const double per = std::max(5.0, (double)budget_ms / (double)cands.size() * 0.5);
    double best_r = -1.0;
    sn_tuning best_t{};
    const auto t0 = std::chrono::steady_clock::now();
    for (const sn_tuning &t : cands) {
        const double r = time_candidate(params, n_replicas, n_species, lengths, prevalence, t, per);
        if (r > best_r) { best_r = r; best_t = t; }
        const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (budget_ms > 0 && el > budget_ms && best_r > 0.0) break;
    }
    if (best_r <= 0.0) return SN_ERR_UNSUPPORTED;
    *best = best_t;
    if (best_rate) *best_rate = best_r;
    if (use_cache) cache_store(key, best_t, best_r);
    return SN_OK;
}

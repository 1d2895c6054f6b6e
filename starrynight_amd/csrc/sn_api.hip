// sn_api.hip -- C ABI (include/starrynight_amd.h) over the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/starrynight_amd.h"
#include "sn_aux_kernels.h"
#include "sn_dispatch.h"
#include "sn_schedule.h"

using namespace sn;

namespace {

struct RepHost {
    int X, Y, Z;
    long long N, base;
    int mixed;
    double L0;
};

}  // namespace

// Partial sums of sn_gradient_all_device_noalloc: a fixed device array, one
// copy per GPU, held until the process ends.  Host code never uses the name
// `scratch` as an address (it is a host-side stand-in): its device address is
// resolved once per device with hipGetSymbolAddress (resolve_scratch, called
// from sn_create with the context's device current) and kept below.
static __device__ __align__(256) double scratch[SN_NOALLOC_PARTIALS];
constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices];   // device address of `scratch` per device id
static std::mutex g_scratch_mu;

struct sn_ctx {
    int R = 0, rc = 0;
    bool two_d = false;
    std::vector<sn_params> params;
    std::vector<RepHost> reps;
    std::vector<ReplicaGeom> geom;
    Schedule sched;
    sn_tuning tuning{};
    int G = 16, NPL = 4, npairs = 0, noff = 0;
    int blk = 256;   // gather kernel's threads per workgroup
    int prop_chunk = kBlock;   // gather kernel's proposals per chunk (LDS slots)
    long long total_sites = 0;
    size_t lds_bytes = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    // device buffers
    double4 *d_lat = nullptr;
    RepDev *d_reps = nullptr;
    TileDev *d_tiles = nullptr;
    ShapeDev *d_shapes = nullptr;
    int *d_class_start = nullptr, *d_class_ri = nullptr, *d_chunk_col = nullptr;
    int *d_gsteps = nullptr, *d_gchunk = nullptr;   // gather kernel's step tables
    int max_gsteps = 0;
    int *d_field_steps = nullptr, *d_tile_coord = nullptr;
    int max_steps = 0;   // longest step table of a tile shape (field kernel)
    int max_tile_padded = 0;   // largest padded tile (field kernel's tile-order arrays)
    PairDev *d_pairs = nullptr, *d_offs = nullptr;
    RadialOff *d_radial = nullptr;   // radial-order offsets sorted by r^2 (built on first use)
    StencilOff *d_stencil = nullptr;  // potential / field offsets (rebuilt when the cutoff changes)
    int stencil_kind = -1, stencil_n = 0;  // -1 none, 0 potential, c > 0 field with cutoff c
    double *d_obs = nullptr;          // observable output staging
    size_t obs_cap = 0;
    int *d_radial_bins = nullptr;
    std::vector<int> radial_bins;
    unsigned *d_sweep = nullptr;
    unsigned long long *d_counters = nullptr;   // 2 per tile (accepted, rejected)
    std::vector<unsigned long long> replay_counts;  // 2 per replica, from sn_mc_replay
    bool use_field = true;
    double *d_partials = nullptr, *d_result = nullptr;
    size_t partial_cap = 0;
    // staging for the host-buffer entry points (site energies, replay, gradient
    // copies, species tables): grows as needed, freed with the context
    void *d_stage = nullptr;
    size_t stage_cap = 0;
    double *d_scratch = nullptr;   // this device's `scratch` array (not owned)
    int max_tile_sites = 0, max_cls = 0, max_chunks = 0;
    int gather_region = 0;   // largest staged region of the gather kernel (z rows padded, gather_zstride)
    int field_region = 0;   // largest staged region of the field kernel (z rows padded to odd)
    int *d_tile_list = nullptr;          // tile indices grouped by replica (tiled energy kernel)
    int *d_tile_seg = nullptr;           // R + 1 offsets into d_tile_list (device copy)
    double *h_result = nullptr;          // pinned host copy of d_result
    std::vector<int> rep_tile_off;       // R + 1 offsets into d_tile_list
    bool energy_tiled = false;           // energy / gradient through tile_energy_kernel
    // whole-lattice sweeps of small replicas (kernel 3): one workgroup per replica
    bool torus = false;
    Sublattice torus_lat{};
    std::vector<std::vector<int32_t>> torus_order, torus_cstart;   // per replica: sites by colour, colour starts
    int torus_ncol = 0;                                             // colours per sweep (shorter colourings padded)
    int torus_G = 0, torus_NPL = 0, torus_blk = 0, torus_ms = 0, torus_mc = 0, torus_pc = 0;
    int *d_torus_sites = nullptr, *d_torus_off = nullptr, *d_torus_cs = nullptr;
    unsigned long long *d_rep_counters = nullptr;   // 2 per replica (torus sweeps)
    bool can_sweep = false;
    std::string sweep_err;
    std::vector<unsigned long long> sweeps_done;
    // graphs of one sweep and of kGraphSweeps sweeps (built on first use)
    hipGraphExec_t graph_exec = nullptr, graph_exec_k = nullptr;
    hipGraph_t graph = nullptr, graph_k = nullptr;
    int graph_mixed = -1;
    std::string err;
};

static int fail(sn_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, SN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static bool any_mixed(const sn_ctx *c)
{
    for (const RepHost &r : c->reps)
        if (r.mixed) return true;
    return false;
}

static double beta_of(double T)
{
    // beta = 1/((float)T/300.0) for integer T (src/starrynight-main.c:215);
    // for a real-valued T the same ratio in double.
    if (T == 0.0) return INFINITY;
    if (T == std::floor(T) && std::fabs(T) < 16777216.0) return 1.0 / ((double)(float)T / 300.0);
    return 300.0 / T;
}

static RepDev make_repdev(const sn_params &p, const RepHost &h)
{
    RepDev d{};
    d.X = h.X; d.Y = h.Y; d.Z = h.Z; d.mixed = h.mixed;
    d.base = h.base;
    d.beta = beta_of(p.temperature);
    d.J = p.cage_strain; d.K = p.K; d.L0 = h.L0;
    d.E[0] = p.efield[0]; d.E[1] = p.efield[1]; d.E[2] = p.efield[2];
    d.seed = p.seed;
    d.constrain_to_x = p.constrain_to_x ? 1 : 0;
    return d;
}

static int upload_rep(sn_ctx *c, int r)
{
    RepDev d = make_repdev(c->params[r], c->reps[r]);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_reps + r, &d, sizeof(d), hipMemcpyHostToDevice));
    return SN_OK;
}

static void free_graph(sn_ctx *c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->graph_exec_k) (void)hipGraphExecDestroy(c->graph_exec_k);
    if (c->graph_k) (void)hipGraphDestroy(c->graph_k);
    c->graph_exec = c->graph_exec_k = nullptr;
    c->graph = c->graph_k = nullptr;
    c->graph_mixed = -1;
}

template <typename T>
static int upload(sn_ctx *c, T **dst, const std::vector<T> &v)
{
    size_t n = std::max<size_t>(v.size(), 1);
    HIPCHK(c, hipMalloc((void **)dst, n * sizeof(T)));
    if (!v.empty()) HIPCHK(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return SN_OK;
}

static int ensure_partials(sn_ctx *c, size_t n)
{
    if (n <= c->partial_cap) return SN_OK;
    if (c->d_partials) (void)hipFree(c->d_partials);
    c->d_partials = nullptr;
    HIPCHK(c, hipMalloc((void **)&c->d_partials, n * sizeof(double)));
    c->partial_cap = n;
    return SN_OK;
}

// The staging buffer, at least `bytes` long.  Its contents do not survive a
// call that grows it.
static int ensure_stage(sn_ctx *c, size_t bytes)
{
    if (bytes <= c->stage_cap) return SN_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));   // no queued work may still read the old buffer
    if (c->d_stage) (void)hipFree(c->d_stage);
    c->d_stage = nullptr;
    c->stage_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 1 << 16);
    HIPCHK(c, hipMalloc(&c->d_stage, cap));
    c->stage_cap = cap;
    return SN_OK;
}

// Byte offsets of consecutive 256-byte-aligned slices of the staging buffer.
struct StageLayout {
    size_t end = 0;
    size_t take(size_t bytes)
    {
        const size_t at = end;
        end += (bytes + 255) & ~(size_t)255;
        return at;
    }
};

template <typename T>
static T *stage_ptr(sn_ctx *c, size_t offset)
{
    return (T *)((char *)c->d_stage + offset);
}

// HIP events destroyed when the list goes out of scope (every return path).
struct EventList {
    std::vector<hipEvent_t> ev;
    ~EventList()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// Column sums of the nblk x width partials table into d_result (at least 256
// doubles), kReduceCols columns per workgroup.
static void launch_reduce(sn_ctx *c, int nblk, int width)
{
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((width + kReduceCols - 1) / kReduceCols), dim3(256), 0,
                       c->stream, c->d_partials, nblk, width, c->d_result, (const int *)nullptr);
}

// Largest tile + halo (sites) staged in LDS: 4 doubles per site (mixed lengths)
// plus the proposal slots and shape tables stay under the 160 KB of a CU.
constexpr int kMaxRegion = 4096;
// Replicas of at most this many sites sweep with the whole-lattice kernel
// when the tuning leaves the kernel to the library (and it fits in LDS).
// Measured against gather tiles, site-updates/s at DipoleCutOff 3: 14^3
// +40-44%, 16 x 16 x 12 +80%, 16^3 +0-6% (one or 64 replicas); 16^3 at
// DipoleCutOff 4 +21-25% (5-10% half vacant), at 3 half vacant -7-10%.
constexpr long long kAutoTorusSites = 4096;
// Largest tile: the field kernel gives each lane at most 2 items of 2 sites
// (3D) or 1 site (2D).
constexpr int kMaxTile3D = kFieldMaxTile3D, kMaxTile2D = kFieldMaxTile2D;

static bool plan(const std::vector<ReplicaGeom> &geom, int rc, const sn_tuning &t, Schedule &s,
                 int chunk = kBlock)
{
    // default edge targets: 8^3 in 3D (the 14^3 staged region at rc = 3);
    // 16 x 16 in 2D, where a tile's region stays small and its colour classes
    // grow from about 6 to 26 sites (1024^2 at rc = 3: 1.03e10 site-updates/s
    // against 5.5e9 with 8 x 8 tiles; rc = 4: 6.8e9 against 4.8e9)
    const bool two_d = geom[0].Z == 1;
    const int d = two_d ? 16 : 8;
    int target[3] = {t.tile[0] > 0 ? t.tile[0] : d, t.tile[1] > 0 ? t.tile[1] : d,
                     t.tile[2] > 0 ? t.tile[2] : 8};
    int kind = t.fine_colouring;
    if (kind == 0) kind = (!two_d && rc == 3) ? 3 : 2;
    return build_schedule(geom, rc, target, kind, chunk, kMaxRegion, two_d ? kMaxTile2D : kMaxTile3D, s);
}

// dynamic LDS bytes of the per-phase gather kernel
static size_t sweep_lds(const sn_ctx *c, bool mixed)
{
    size_t dbl = (size_t)((mixed ? 4 : 3) * c->gather_region + 4 * c->prop_chunk) * sizeof(double);
    size_t ints = (size_t)(c->max_tile_sites + c->max_cls + 1 + 2 * (c->max_chunks + 1) +
                           (c->max_gsteps + 1) * (c->blk / c->G + 1)) * sizeof(int);
    return dbl + ints;
}

// LDS doubles of the field kernel's shared area: the staged region (padded)
// first, then the same space as class-ordered compact arrays.
static int field_region_doubles(const sn_ctx *c) { return c->field_region + 2; }

static int field_phaseA(const sn_ctx *c, bool mixed)
{
    int region = (mixed ? 4 : 3) * field_region_doubles(c);
    int compact = (mixed ? 11 : 10) * c->max_tile_padded;
    return std::max(region, compact);
}

static size_t field_lds(const sn_ctx *c, bool mixed)
{
    return (size_t)field_phaseA(c, mixed) * sizeof(double) +
           (size_t)(c->max_tile_padded + c->max_steps * (kBlock / kFieldG + 1)) * sizeof(int);
}

extern "C" {

int sn_plan_schedule(const sn_params *params, int32_t n_replicas, const sn_tuning *tuning,
                     int32_t replica, int64_t *n_sites, int32_t *n_phases, int32_t *order,
                     int32_t *phase_start)
{
    if (!params || n_replicas < 1 || replica < 0 || replica >= n_replicas) return SN_ERR_INVALID;
    std::vector<ReplicaGeom> geom;
    for (int r = 0; r < n_replicas; r++) {
        const sn_params &p = params[r];
        if (p.dipole_cutoff != params[0].dipole_cutoff) return SN_ERR_INVALID;
        if (p.X < 1 || p.Y < 1 || p.Z < 1) return SN_ERR_INVALID;
        if ((p.Z == 1) != (params[0].Z == 1)) return SN_ERR_INVALID;
        if (p.dipole_cutoff < 1 || p.dipole_cutoff > 8) return SN_ERR_UNSUPPORTED;
        if (p.X <= p.dipole_cutoff || p.Y <= p.dipole_cutoff || (p.Z > 1 && p.Z <= p.dipole_cutoff))
            return SN_ERR_UNSUPPORTED;
        geom.push_back({p.X, p.Y, p.Z});
    }
    sn_tuning t{};
    if (tuning) t = *tuning;
    Schedule s;
    if (!plan(geom, params[0].dipole_cutoff, t, s)) return SN_ERR_UNSUPPORTED;
    std::vector<int32_t> o, ps;
    sweep_order(s, geom[replica], replica, o, ps);
    if (n_sites) *n_sites = (int64_t)o.size();
    if (n_phases) *n_phases = (int32_t)ps.size() - 1;
    if (order) std::memcpy(order, o.data(), o.size() * sizeof(int32_t));
    if (phase_start) std::memcpy(phase_start, ps.data(), ps.size() * sizeof(int32_t));
    return SN_OK;
}

const char *sn_status_string(int status)
{
    switch (status) {
    case SN_OK: return "ok";
    case SN_ERR_INVALID: return "invalid argument";
    case SN_ERR_HIP: return "HIP runtime error";
    case SN_ERR_NOMEM: return "out of memory";
    case SN_ERR_RANGE: return "index out of range";
    case SN_ERR_UNSUPPORTED: return "unsupported configuration";
    }
    return "unknown status";
}

const char *sn_last_error(const sn_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// Whole-lattice sweeps (kernel 3): every replica in LDS at once, coloured by
// a sublattice that holds the replica boxes (periodic_sublattice).  Refused
// when a replica is too large for LDS or no such colouring exists.
constexpr int kTorusMaxIndex = 128;

static size_t torus_lds(const sn_ctx *c, bool mixed)
{
    return (size_t)((mixed ? 4 : 3) * c->torus_ms + 4 * c->torus_pc) * sizeof(double) +
           (size_t)c->torus_ms * sizeof(int);
}

// lanes per site, pairs per lane and threads for npairs pairs and colours of
// at most mc sites: 1024 threads when a colour has enough sites to keep them
// busy.  The greedy colourings' colours of up to 64 sites (wide) take 8 lanes
// x 2 pairs in 512 threads (measured at 25 x 25: 4.3e7 site-updates/s against
// 4.1e7 with 16 lanes in 1024 threads).
static bool torus_shape(int np, int mc, bool wide, int &G, int &NPL, int &blk)
{
    blk = 256;
    if (np <= 16 && mc * 16 <= 512) { G = 16; NPL = 1; blk = mc * 16 > 256 ? 512 : 256; }
    else if (np <= 16 && wide && mc * 8 <= 512) { G = 8; NPL = 2; blk = 512; }
    else if (np <= 16) { G = 4; NPL = 4; }
    else if (np <= 32) { G = 8; NPL = 4; }
    else if (np <= 64) {
        if (mc * 16 >= 1024) { G = 32; NPL = 2; blk = 1024; }
        else { G = 16; NPL = 4; }
    }
    else if (np <= 128) { G = 64; NPL = 2; blk = 1024; }
    else return false;
    return true;
}

// Steps plus barriers of one sweep: each colour takes ceil(size / groups)
// steps and one barrier.
static long long torus_cost(const std::vector<std::vector<int32_t>> &cstart, int groups)
{
    long long cost = 0;
    for (const auto &st : cstart)
        for (size_t k = 0; k + 1 < st.size(); k++) cost += (st[k + 1] - st[k] + groups - 1) / groups + 1;
    return cost;
}

static int setup_torus(sn_ctx *c)
{
    const int np = c->npairs;
    for (const ReplicaGeom &g : c->geom)
        if (g.X > 1023 || g.Y > 1023 || g.Z > 1023)
            return fail(c, SN_ERR_UNSUPPORTED, "whole-lattice sweep: an axis is longer than 1023");
    for (int r = 0; r < c->R; r++) c->torus_ms = std::max<int>(c->torus_ms, (int)c->reps[r].N);
    // the x, y, z, length arrays and the site table alone (before colouring)
    if ((size_t)c->torus_ms * (4 * sizeof(double) + sizeof(int)) > 150 * 1024)
        return fail(c, SN_ERR_UNSUPPORTED, "whole-lattice sweep: a replica does not fit in LDS (" +
                                               std::to_string(c->torus_ms) + " sites)");
    // Two colourings: a sublattice holding every replica box (the same colours
    // for all replicas), and, at up to 16 pairs, a greedy colouring per replica
    // with colours of at most 64 sites (one step of 16 lanes x 64 sites).
    // The one with fewer steps plus barriers per sweep is kept.
    auto max_class = [](const std::vector<std::vector<int32_t>> &cst) {
        int mc = 1;
        for (const auto &st : cst)
            for (size_t k = 0; k + 1 < st.size(); k++) mc = std::max(mc, st[k + 1] - st[k]);
        return mc;
    };
    std::vector<std::vector<int32_t>> lo, ls, go, gs;
    long long lcost = -1, gcost = -1;
    if (periodic_sublattice(c->geom, c->rc, kTorusMaxIndex, c->torus_lat)) {
        lo.resize(c->R); ls.resize(c->R);
        for (int r = 0; r < c->R; r++) torus_classes(c->torus_lat, c->geom[r], lo[r], ls[r]);
        int G, NPL, blk;
        if (torus_shape(np, max_class(ls), false, G, NPL, blk)) lcost = torus_cost(ls, blk / G);
    }
    if (np <= 16) {
        go.resize(c->R); gs.resize(c->R);
        bool ok = true;
        for (int r = 0; r < c->R && ok; r++) {
            int same = -1;   // replicas of one box share its colouring
            for (int q = 0; q < r && same < 0; q++)
                if (c->geom[q].X == c->geom[r].X && c->geom[q].Y == c->geom[r].Y && c->geom[q].Z == c->geom[r].Z)
                    same = q;
            if (same >= 0) { go[r] = go[same]; gs[r] = gs[same]; continue; }
            greedy_torus_classes(c->geom[r], c->rc, 64, go[r], gs[r]);
            ok = valid_torus_classes(c->geom[r], c->rc, go[r], gs[r]);
        }
        int G, NPL, blk;
        if (ok && torus_shape(np, max_class(gs), true, G, NPL, blk)) gcost = torus_cost(gs, blk / G);
    }
    if (lcost < 0 && gcost < 0)
        return fail(c, SN_ERR_UNSUPPORTED, "whole-lattice sweep: no periodic colouring of these lattice sizes");
    const bool greedy = gcost >= 0 && (lcost < 0 || gcost < lcost);
    c->torus_order = greedy ? go : lo;
    c->torus_cstart = greedy ? gs : ls;
    // no empty colour before a replica's last site: the kernel's lookahead
    // takes an empty colour for the end of the replica's sites
    for (auto &st_ : c->torus_cstart) st_.erase(std::unique(st_.begin(), st_.end()), st_.end());
    c->torus_ncol = 0;
    for (const auto &st : c->torus_cstart) c->torus_ncol = std::max(c->torus_ncol, (int)st.size() - 1);
    std::vector<int> sites, off, cs;
    for (int r = 0; r < c->R; r++) {
        const ReplicaGeom &g = c->geom[r];
        off.push_back((int)sites.size());
        for (int32_t si : c->torus_order[r]) {   // packed coordinates x | y << 10 | z << 20
            const int z = si % g.Z, xy = si / g.Z, y = xy % g.Y, x = xy / g.Y;
            sites.push_back(x | (y << 10) | (z << 20));
        }
        std::vector<int32_t> &st_ = c->torus_cstart[r];
        while ((int)st_.size() < c->torus_ncol + 1) st_.push_back(st_.back());   // empty colours
        cs.insert(cs.end(), st_.begin(), st_.end());
    }
    c->torus_mc = max_class(c->torus_cstart);
    int G, NPL, blk;
    if (!torus_shape(np, c->torus_mc, greedy, G, NPL, blk))
        return fail(c, SN_ERR_UNSUPPORTED, "whole-lattice sweep: DipoleCutOff too large");
    if (c->tuning.lanes_per_site > 0) {
        // requested lanes per site: pairs per lane to cover the table, threads
        // for up to one step per colour
        G = c->tuning.lanes_per_site;
        NPL = 1;
        while (NPL * G < np) NPL <<= 1;
        blk = c->torus_mc * G > 512 ? 1024 : c->torus_mc * G > 256 ? 512 : 256;
        if (!pick_torus(G, NPL, false, blk) && blk < 1024 && pick_torus(G, NPL, false, 1024)) blk = 1024;
        if (!pick_torus(G, NPL, false, blk))   // fewer threads: several rounds per colour
            for (int b : {512, 256})
                if (pick_torus(G, NPL, false, b)) { blk = b; break; }
    }
    c->torus_G = G; c->torus_NPL = NPL; c->torus_blk = blk;
    // proposal slots: whole runs of colours per draw, as many as LDS leaves room
    // for (at most the largest replica, at least the largest colour)
    const long long room = (150LL * 1024 / (long long)sizeof(double) - 4LL * c->torus_ms - c->torus_ms / 2) / 4;
    c->torus_pc = (int)std::max<long long>(c->torus_mc, std::min<long long>(room, c->torus_ms));
    const size_t lds = torus_lds(c, true);
    if (lds > 150 * 1024)
        return fail(c, SN_ERR_UNSUPPORTED, "whole-lattice sweep: a replica does not fit in LDS (" +
                                               std::to_string(c->torus_ms) + " sites)");
    int st;
    for (int m = 0; m < 2; m++) {
        const void *f = (const void *)pick_torus(G, NPL, m == 1, blk);
        if (!f) return fail(c, SN_ERR_UNSUPPORTED, "whole-lattice sweep: no kernel for this cutoff");
        HIPCHK(c, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    if ((st = upload(c, &c->d_torus_sites, sites))) return st;
    if ((st = upload(c, &c->d_torus_off, off))) return st;
    if ((st = upload(c, &c->d_torus_cs, cs))) return st;
    std::vector<unsigned long long> z(2 * c->R, 0ull);
    if ((st = upload(c, &c->d_rep_counters, z))) return st;
    c->torus = true;
    return SN_OK;
}

// This device's address of `scratch`, resolved on the first context created
// on it (the device is current), then reused by every later context.
static int resolve_scratch(sn_ctx *c)
{
    if (c->device < 0 || c->device >= kMaxDevices) return fail(c, SN_ERR_UNSUPPORTED, "device id beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, SN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[c->device] = p;
    }
    c->d_scratch = (double *)g_scratch_dev[c->device];
    return SN_OK;
}

int sn_create(const sn_params *params, int32_t n_replicas, const sn_tuning *tuning, sn_ctx **out)
{
    if (!out) return SN_ERR_INVALID;
    *out = nullptr;
    if (!params || n_replicas < 1) return SN_ERR_INVALID;
    sn_ctx *c = new (std::nothrow) sn_ctx();
    if (!c) return SN_ERR_NOMEM;
    *out = c;
    c->R = n_replicas;
    c->params.assign(params, params + n_replicas);
    c->rc = params[0].dipole_cutoff;
    if (tuning) c->tuning = *tuning;
    for (int r = 0; r < n_replicas; r++) {
        const sn_params &p = params[r];
        if (p.dipole_cutoff != c->rc)
            return fail(c, SN_ERR_INVALID, "all replicas of a context must share DipoleCutOff");
        if (p.X < 1 || p.Y < 1 || p.Z < 1) return fail(c, SN_ERR_INVALID, "lattice size must be positive");
        if (p.dipole_cutoff < 1 || p.dipole_cutoff > 8)
            return fail(c, SN_ERR_UNSUPPORTED, "DipoleCutOff must be in 1..8");
        // every offset must wrap at most once (the reference's % wrap needs X > -dx too)
        if (p.X <= p.dipole_cutoff || p.Y <= p.dipole_cutoff || (p.Z > 1 && p.Z <= p.dipole_cutoff))
            return fail(c, SN_ERR_UNSUPPORTED, "each lattice axis (except Z == 1) must exceed DipoleCutOff");
        if ((p.Z == 1) != (params[0].Z == 1))
            return fail(c, SN_ERR_INVALID, "replicas of a context must all be 2D or all 3D");
        if (!(p.temperature >= 0.0)) return fail(c, SN_ERR_INVALID, "temperature must be >= 0");
        long long N = (long long)p.X * p.Y * p.Z;
        if (N > 0x7fffffffLL) return fail(c, SN_ERR_RANGE, "lattice larger than 2^31 sites");
        RepHost h{p.X, p.Y, p.Z, N, c->total_sites, 0, 1.0};
        c->reps.push_back(h);
        c->geom.push_back({p.X, p.Y, p.Z});
        c->total_sites += N;
    }
    c->two_d = params[0].Z == 1;
    // A lattice that cannot be tiled still supports the gather entry points
    // (site_energy, replay, energy, gradient); only sweeps are refused.
    c->can_sweep = plan(c->geom, c->rc, c->tuning, c->sched);
    // Tiles of at most 256 sites draw proposals in chunks of 128: the gather
    // kernel's proposal buffer halves, and three workgroups fit in a CU's LDS.
    if (c->can_sweep) {
        int mt = 0;
        for (const TileShape &sh_ : c->sched.shapes) mt = std::max(mt, sh_.tx * sh_.ty * sh_.tz);
        Schedule small;
        if (mt <= 256 && plan(c->geom, c->rc, c->tuning, small, 128)) {
            c->sched = small;
            c->prop_chunk = 128;
        }
    }
    if (!c->can_sweep) {
        c->sweep_err = c->sched.error;
        c->sched.shapes.clear();
        c->sched.tiles.clear();
        c->sched.coarse_start.assign(9, 0);
    }

    // neighbour table: +-r pairs for the tile sweep, full offsets for gathers
    std::vector<Offset> pairs = neighbour_pairs(c->rc, c->two_d);
    std::vector<Offset> offs = neighbour_offsets(c->rc, c->two_d);
    auto to_dev = [](const Offset &o) {
        PairDev p{};
        p.dx = o.dx; p.dy = o.dy; p.dz = o.dz; p.nn = (o.r2 == 1);
        double d = std::sqrt((double)o.r2);
        p.nx = o.dx / d; p.ny = o.dy / d; p.nz = o.dz / d;
        p.w = 1.0 / (d * d * d);
        return p;
    };
    std::vector<PairDev> hp, ho;
    for (const Offset &o : pairs) hp.push_back(to_dev(o));
    for (const Offset &o : offs) ho.push_back(to_dev(o));
    c->npairs = (int)hp.size();
    c->noff = (int)ho.size();

    // lanes per site and pairs per lane
    int max_class = 1;
    for (const TileShape &s : c->sched.shapes) max_class = std::max(max_class, s.max_class);
    int blk = c->tuning.block_threads > 0 ? c->tuning.block_threads : kBlock;
    if (blk != 256 && blk != 512) return fail(c, SN_ERR_INVALID, "block_threads must be 256 or 512");
    if (c->tuning.kernel < 0 || c->tuning.kernel > 3) return fail(c, SN_ERR_INVALID, "kernel must be 0..3");
    // kernel 0 picks by size (rates measured on MI355X): the whole-lattice
    // kernel for replicas of up to kAutoTorusSites sites (2-3x the gather
    // kernel at 12^3 and 25x25), else the per-phase gather kernel (1.2-2x the
    // field-update kernel from 16^3 to 128^3 at DipoleCutOff 2-4, batched or
    // not; 2.3-2.6x at DipoleCutOff 1 and 1.55x at 5, 64^3 and 128^3), else
    // the field-update kernel where no gather instance fits
    int kern = c->tuning.kernel;
    if (kern == 0 && blk != kBlock) kern = 2;
    if (blk != kBlock && kern != 2)
        return fail(c, SN_ERR_INVALID, "block_threads applies to the gather kernel (kernel = 2) only");
    // pairs per lane for G lanes a site: exactly enough where an instance
    // has that count, else the next power of two (padded slots carry zero
    // weight)
    auto pairs_per_lane = [&](int G, int threads) {
        const int need = (c->npairs + G - 1) / G;
        if (pick_sweep(G, need, false, threads)) return need;
        int n = 2;
        while (n < need) n <<= 1;
        return n;
    };
    // Lanes per site: the G whose groups hold a colour class in one step
    // (down to 4 lanes) with the fewest pair slots G x NPL, the larger G on
    // a tie.  Measured at 1024^2 (2D), site-updates/s: DipoleCutOff 3 (14
    // pairs) 16 x 1 5.8e9 against 32 x 2 4.7e9; DipoleCutOff 4 (24 pairs)
    // 16 x 2 4.8e9 against 32 x 2 4.3e9; DipoleCutOff 2 (6 pairs) 4 x 2
    // 6.8e9 against 16 x 1 6.7e9.  3D lattices keep 16 x 4 (61 pairs) and
    // 16 x 8 (128), and DipoleCutOff 5 (257 pairs) takes 64 x 5.
    auto lane_shape = [&](int threads, int &G, int &NPL) {
        G = c->tuning.lanes_per_site;
        if (G > 0) {
            NPL = pairs_per_lane(G, threads);
            return;
        }
        int best = 0, best_slots = 0;
        for (int g = 64; g >= 4; g >>= 1) {
            if (g > 4 && threads / g < max_class) continue;
            const int n = pairs_per_lane(g, threads);
            if (!pick_sweep(g, n, false, threads)) continue;
            if (best == 0 || g * n < best_slots) { best = g; best_slots = g * n; }
        }
        if (best == 0) {   // no instance: the shape the error message names
            best = 64;
            while (best > 4 && threads / best < max_class) best >>= 1;
        }
        G = best;
        NPL = pairs_per_lane(G, threads);
    };
    int G, NPL;
    lane_shape(blk, G, NPL);
    // The gather kernel runs 512 threads, unless the tuning fixes the block
    // or the lanes, where 256 threads at 4 lanes a site hold fewer sites than
    // a colour class (DipoleCutOff 1: 4 lanes x 1 pair, 1.06-1.09x the
    // 256-thread rate at 64^3 and 128^3).  DipoleCutOff 5 stays at 256
    // threads, 64 lanes x 5 pairs: 1.16x 512 threads of 64 lanes x 8 pairs,
    // 1.55x the field-update kernel at 128^3.
    // It also runs 512 threads where its staged region leaves room for one
    // 256-thread workgroup per CU (four waves) and 512 threads give a site
    // twice the lanes: 128^3 at DipoleCutOff 4 (16^3 region), 32 lanes x 4
    // pairs against 16 x 8, +4% (+17% half vacant).
    int region = 0;
    for (const TileShape &s : c->sched.shapes)
        region = std::max(region, (s.tx + 2 * c->rc) * (s.ty + 2 * c->rc) *
                                      gather_zstride(s.tz + (c->two_d ? 0 : 2 * c->rc)));
    const bool one_per_cu = (size_t)(3 * region + 4 * c->prop_chunk) * sizeof(double) > 80 * 1024;
    const bool wide = max_class > kBlock / 4;
    if (c->tuning.block_threads <= 0 && c->tuning.lanes_per_site <= 0 && (kern == 0 || kern == 2) &&
        (wide || one_per_cu)) {
        int G2, NPL2;
        lane_shape(512, G2, NPL2);
        if (pick_sweep(G2, NPL2, false, 512) && (wide || G2 > G)) { blk = 512; G = G2; NPL = NPL2; }
    }
    c->blk = blk;
    c->G = G; c->NPL = NPL;
    const bool field_ok = pick_field(c->rc, c->two_d, false) != nullptr;
    const bool gather_ok = pick_sweep(G, NPL, false, blk) != nullptr;
    bool auto_torus = false;
    if (kern == 0) {
        long long largest = 0;
        for (const RepHost &h : c->reps) largest = std::max(largest, h.N);
        auto_torus = largest <= kAutoTorusSites;
        kern = (gather_ok || !field_ok) ? 2 : 1;
    }
    if (c->can_sweep && (kern == 2 || !field_ok) && !gather_ok)
        return fail(c, SN_ERR_UNSUPPORTED, "no tile-sweep kernel for this lanes-per-site / cutoff (" +
                                               std::to_string(G) + " lanes x " + std::to_string(NPL) +
                                               " pairs per lane, " + std::to_string(blk) + " threads; " +
                                               std::to_string(c->npairs) + " pairs, colour classes of up to " +
                                               std::to_string(max_class) + " sites)");

    HIPCHK(c, hipGetDevice(&c->device));
    if (int se = resolve_scratch(c)) return se;
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(c, hipMalloc((void **)&c->d_lat, (size_t)c->total_sites * sizeof(double4)));
    HIPCHK(c, hipMemset(c->d_lat, 0, (size_t)c->total_sites * sizeof(double4)));
    std::vector<RepDev> rd;
    for (int r = 0; r < c->R; r++) rd.push_back(make_repdev(c->params[r], c->reps[r]));
    int st;
    if ((st = upload(c, &c->d_reps, rd))) return st;
    std::vector<TileDev> td;
    for (const Tile &t : c->sched.tiles) td.push_back({t.replica, t.x0, t.y0, t.z0, t.shape, 0, 0, 0});
    if ((st = upload(c, &c->d_tiles, td))) return st;
    std::vector<ShapeDev> sd;
    std::vector<int> cstart, cri, chunks, fsteps, tcoord, gsteps, gchunk;
    const int NGg = blk / G;   // lane groups of the gather kernel
    constexpr int NG = kBlock / kFieldG;   // lane groups of the field kernel
    const int hz = c->two_d ? 0 : c->rc;
    for (const TileShape &s : c->sched.shapes) {
        ShapeDev d{};
        d.tx = s.tx; d.ty = s.ty; d.tz = s.tz; d.ncol = s.ncol;
        d.nchunks = (int)s.chunk_start.size() - 1;
        d.class_off = (int)cstart.size();
        d.chunk_off = (int)chunks.size();
        d.sites_off = (int)cri.size();
        cstart.insert(cstart.end(), s.class_start.begin(), s.class_start.end());
        // by fine colour: region index (tile + halo) and tile-local index of
        // each site; by tile index: packed local coordinates
        const int RY = s.ty + 2 * c->rc, RZ = s.tz + 2 * hz, RZs = gather_zstride(RZ);
        for (size_t k = 0; k < s.class_sites.size(); k++) {
            int li = s.class_sites[k];
            int lz = li % s.tz, ly = (li / s.tz) % s.ty, lx = li / (s.tz * s.ty);
            cri.push_back(((lx + c->rc) * RY + (ly + c->rc)) * RZs + (lz + hz));
        }
        c->gather_region = std::max(c->gather_region, (s.tx + 2 * c->rc) * RY * RZs);
        // field kernel: tile-order arrays use padded strides (z in 3D, y in
        // 2D) so that a scatter's targets spread over the LDS banks
        auto pad = [](int t) { return (t % 2 == 0) ? t + 3 : t + 2; };
        d.tzp = c->two_d ? 1 : pad(s.tz);
        d.typ = c->two_d ? pad(s.ty) : s.ty;
        d.coord_off = (int)tcoord.size();
        const int npad = s.tx * d.typ * d.tzp;
        for (int i = 0; i < npad; i++) {
            const int lz = i % d.tzp, ly = (i / d.tzp) % d.typ, lx = i / (d.tzp * d.typ);
            tcoord.push_back(lz < s.tz && ly < s.ty ? (lx | (ly << 10) | (lz << 20)) : -1);
        }
        c->max_tile_padded = std::max(c->max_tile_padded, npad);
        auto padded = [&](int li) {
            const int lz = li % s.tz, ly = (li / s.tz) % s.ty, lx = li / (s.tz * s.ty);
            return (lx * d.typ + ly) * d.tzp + lz;
        };
        // field kernel steps: each colour in blocks of NG sites; per step a
        // barrier flag (last block of the colour) and NG padded tile indices (-1: none)
        d.step_off = (int)fsteps.size();
        d.nsteps = 0;
        for (int col = 0; col < s.ncol; col++) {
            const int a0 = s.class_start[col], a1 = s.class_start[col + 1];
            for (int b0 = a0; b0 < a1; b0 += NG) {
                fsteps.push_back(b0 + NG >= a1 ? 1 : 0);
                for (int g = 0; g < NG; g++) fsteps.push_back(b0 + g < a1 ? padded(s.class_sites[b0 + g]) : -1);
                d.nsteps++;
            }
        }
        c->max_steps = std::max(c->max_steps, d.nsteps);
        chunks.insert(chunks.end(), s.chunk_start.begin(), s.chunk_start.end());
        // gather kernel steps: per proposal chunk, each colour in blocks of NGg
        // sites; per step a barrier flag (last block of the colour) and NGg
        // entries (region index | slot in the chunk's proposals << 16, or -1)
        d.gstep_off = (int)gsteps.size();
        d.gchunk_off = (int)gchunk.size();
        int nst = 0;
        for (int ch = 0; ch < d.nchunks; ch++) {
            gchunk.push_back(nst);
            const int c0 = s.chunk_start[ch], c1 = s.chunk_start[ch + 1];
            const int s0 = s.class_start[c0];
            for (int col = c0; col < c1; col++) {
                const int a0 = s.class_start[col], a1 = s.class_start[col + 1];
                for (int b0 = a0; b0 < a1; b0 += NGg) {
                    gsteps.push_back(b0 + NGg >= a1 ? 1 : 0);
                    for (int g = 0; g < NGg; g++)
                        gsteps.push_back(b0 + g < a1 ? (cri[d.sites_off + b0 + g] | ((b0 + g - s0) << 16)) : -1);
                    nst++;
                }
            }
        }
        gchunk.push_back(nst);
        d.gsteps_n = nst;
        c->max_gsteps = std::max(c->max_gsteps, nst);
        c->max_tile_sites = std::max(c->max_tile_sites, s.tx * s.ty * s.tz);
        const int RZp = c->two_d ? RZ : (RZ | 1);
        c->field_region = std::max(c->field_region, (s.tx + 2 * c->rc) * RY * RZp);
        c->max_cls = std::max(c->max_cls, s.ncol);
        c->max_chunks = std::max(c->max_chunks, d.nchunks);
        sd.push_back(d);
    }
    if ((st = upload(c, &c->d_shapes, sd))) return st;
    if ((st = upload(c, &c->d_class_start, cstart))) return st;
    if ((st = upload(c, &c->d_class_ri, cri))) return st;
    if ((st = upload(c, &c->d_field_steps, fsteps))) return st;
    if ((st = upload(c, &c->d_tile_coord, tcoord))) return st;
    if ((st = upload(c, &c->d_chunk_col, chunks))) return st;
    if ((st = upload(c, &c->d_gsteps, gsteps))) return st;
    if ((st = upload(c, &c->d_gchunk, gchunk))) return st;
    if ((st = upload(c, &c->d_pairs, hp))) return st;
    if ((st = upload(c, &c->d_offs, ho))) return st;
    std::vector<unsigned> zeros_u(c->R, 0u);
    if ((st = upload(c, &c->d_sweep, zeros_u))) return st;
    std::vector<unsigned long long> zeros_c(2 * std::max<size_t>(c->sched.tiles.size(), 1), 0ull);
    if ((st = upload(c, &c->d_counters, zeros_c))) return st;
    const size_t nresult = std::max<size_t>(256, (size_t)c->R * kNPart);
    HIPCHK(c, hipMalloc((void **)&c->d_result, nresult * sizeof(double)));
    HIPCHK(c, hipHostMalloc((void **)&c->h_result, nresult * sizeof(double), hipHostMallocDefault));
    c->sweeps_done.assign(c->R, 0);
    c->replay_counts.assign(2 * c->R, 0ull);

    // kernel choice: the field-update kernel where it is instantiated
    c->use_field = kern != 2 && field_ok;
    if (kern == 1 && !c->use_field)
        return fail(c, SN_ERR_UNSUPPORTED, "field-update kernel needs DipoleCutOff in 1..5");
    size_t need_lds = c->use_field ? std::max(field_lds(c, false), field_lds(c, true))
                                   : sweep_lds(c, true);
    c->lds_bytes = need_lds;
    if (need_lds > 160 * 1024) {
        c->can_sweep = false;
        c->sweep_err = "tile + halo does not fit in LDS; use a smaller tile";
    }
    const bool tiles_ok = c->can_sweep;
    if (kern == 3) {
        if ((st = setup_torus(c))) return st;
        c->can_sweep = true;
    } else if (auto_torus) {
        // lattices the whole-lattice kernel cannot colour keep the tiles
        const std::string keep = c->err;
        const int ts = setup_torus(c);
        if (ts == SN_OK) c->can_sweep = true;
        else if (ts != SN_ERR_UNSUPPORTED) return ts;
        else c->err = keep;
    }
    if (!tiles_ok) return SN_OK;
    // the tiled energy kernel stages the field kernel's padded region
    const size_t energy_lds = (size_t)4 * field_region_doubles(c) * sizeof(double);
    if (pick_energy(c->rc, c->two_d, false) && energy_lds <= 150 * 1024) {
        std::vector<int> tl;
        c->rep_tile_off.assign(c->R + 1, 0);
        for (int r = 0; r < c->R; r++) {
            c->rep_tile_off[r] = (int)tl.size();
            for (size_t t = 0; t < c->sched.tiles.size(); t++)
                if (c->sched.tiles[t].replica == r) tl.push_back((int)t);
        }
        c->rep_tile_off[c->R] = (int)tl.size();
        if ((st = upload(c, &c->d_tile_list, tl))) return st;
        if ((st = upload(c, &c->d_tile_seg, c->rep_tile_off))) return st;
        for (int m = 0; m < 2; m++)
            HIPCHK(c, hipFuncSetAttribute((const void *)pick_energy(c->rc, c->two_d, m == 1),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)energy_lds));
        c->energy_tiled = true;
    }
    for (int m = 0; m < 2; m++) {
        const void *f = c->use_field
                            ? (const void *)pick_field(c->rc, c->two_d, m == 1)
                            : (const void *)pick_sweep(G, NPL, m == 1, c->blk, c->gather_region, c->prop_chunk);
        HIPCHK(c, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need_lds));
    }
    return SN_OK;
}

int sn_destroy(sn_ctx *c)
{
    if (!c) return SN_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_graph(c);
    void *bufs[] = {c->d_lat, c->d_reps, c->d_tiles, c->d_shapes, c->d_class_start, c->d_class_ri,
                    c->d_field_steps, c->d_tile_coord, c->d_chunk_col, c->d_gsteps, c->d_gchunk, c->d_pairs,
                    c->d_offs, c->d_sweep,
                    c->d_counters, c->d_partials, c->d_result, c->d_radial, c->d_radial_bins,
                    c->d_stencil, c->d_obs, c->d_tile_list, c->d_tile_seg, c->d_torus_sites,
                    c->d_torus_off, c->d_torus_cs, c->d_rep_counters, c->d_stage};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SN_OK;
}

static int check_rep(sn_ctx *c, int r)
{
    if (!c) return SN_ERR_INVALID;
    if (r < 0 || r >= c->R) return fail(c, SN_ERR_RANGE, "replica index out of range");
    return SN_OK;
}

int sn_set_lattice(sn_ctx *c, int32_t r, const double *xyzl)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!xyzl) return fail(c, SN_ERR_INVALID, "null lattice buffer");
    RepHost &h = c->reps[r];
    double L0 = xyzl[3];
    int mixed = 0;
    for (long long i = 0; i < h.N; i++) {
        const double *p = xyzl + 4 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || !std::isfinite(p[3]))
            return fail(c, SN_ERR_INVALID, "non-finite value in lattice");
        if (p[3] != L0) mixed = 1;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_lat + h.base, xyzl, h.N * sizeof(double4), hipMemcpyHostToDevice));
    h.mixed = mixed;
    h.L0 = mixed ? 0.0 : L0;
    return upload_rep(c, r);
}

int sn_set_lattice_device(sn_ctx *c, int32_t r, const void *xyzl_dev)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!xyzl_dev) return fail(c, SN_ERR_INVALID, "null lattice buffer");
    RepHost &h = c->reps[r];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the caller's writes to xyzl_dev are complete (its streams synchronised);
    // the source is checked before anything is copied, as sn_set_lattice does
    const int nblk = (int)std::min<long long>((h.N + 255) / 256, 1024);
    if ((st = ensure_partials(c, (size_t)nblk * 3))) return st;
    hipLaunchKernelGGL(lattice_scan_kernel, dim3(nblk), dim3(256), 0, c->stream, (const double4 *)xyzl_dev,
                       (long long)h.N, c->d_partials);
    HIPCHK(c, hipGetLastError());
    std::vector<double> part((size_t)nblk * 3);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->d_partials, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double bad = 0.0, lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < nblk; b++) {
        bad += part[3 * b];
        lo = std::min(lo, part[3 * b + 1]);
        hi = std::max(hi, part[3 * b + 2]);
    }
    if (bad > 0.0) return fail(c, SN_ERR_INVALID, "non-finite value in lattice");
    HIPCHK(c, hipMemcpyAsync(c->d_lat + h.base, xyzl_dev, h.N * sizeof(double4), hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    h.mixed = lo != hi;
    h.L0 = h.mixed ? 0.0 : lo;
    return upload_rep(c, r);
}

int sn_get_lattice(sn_ctx *c, int32_t r, double *xyzl)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!xyzl) return fail(c, SN_ERR_INVALID, "null lattice buffer");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(xyzl, c->d_lat + c->reps[r].base, c->reps[r].N * sizeof(double4),
                        hipMemcpyDeviceToHost));
    return SN_OK;
}

int sn_lattice_device_ptr(sn_ctx *c, int32_t r, void **p)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!p) return fail(c, SN_ERR_INVALID, "null output pointer");
    *p = (void *)(c->d_lat + c->reps[r].base);
    return SN_OK;
}

int sn_init_lattice_random(sn_ctx *c, int32_t r, int32_t n_species, const double *lengths,
                           const double *prevalence)
{
    int st = check_rep(c, r);
    if (st) return st;
    std::vector<double> len(1, 1.0), cum(1, 1.0);
    if (n_species > 0) {
        if (n_species > 10 || !lengths || !prevalence)
            return fail(c, SN_ERR_INVALID, "species table must have 1..10 entries");
        len.assign(lengths, lengths + n_species);
        cum.assign(n_species, 0.0);
        double acc = 0.0;
        for (int i = 0; i < n_species; i++) {
            if (!(prevalence[i] >= 0.0)) return fail(c, SN_ERR_INVALID, "negative prevalence");
            acc += prevalence[i];
            cum[i] = acc;
        }
        if (acc < 1.0 - 1e-6) return fail(c, SN_ERR_INVALID, "Prevalence must sum to 1");
    }
    RepHost &h = c->reps[r];
    // lengths that can occur: species with non-zero prevalence
    std::vector<double> used;
    for (size_t i = 0; i < len.size(); i++)
        if ((i == 0 ? cum[0] : cum[i] - cum[i - 1]) > 0.0) used.push_back(len[i]);
    if (used.empty()) used.push_back(len[0]);
    h.mixed = 0;
    for (double v : used)
        if (v != used[0]) h.mixed = 1;
    h.L0 = h.mixed ? 0.0 : used[0];
    if ((st = upload_rep(c, r))) return st;
    StageLayout lay;
    const size_t o_len = lay.take(len.size() * sizeof(double)), o_cum = lay.take(cum.size() * sizeof(double));
    if ((st = ensure_stage(c, lay.end))) return st;
    double *dl = stage_ptr<double>(c, o_len), *dc = stage_ptr<double>(c, o_cum);
    HIPCHK(c, hipMemcpyAsync(dl, len.data(), len.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dc, cum.data(), cum.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int nblk = (int)((h.N + 255) / 256);
    hipLaunchKernelGGL(init_random_kernel, dim3(nblk), dim3(256), 0, c->stream, c->d_lat, c->d_reps, r,
                       (int)len.size(), dl, dc);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SN_OK;
}

int sn_set_params(sn_ctx *c, int32_t r, double temperature, const double efield[3], double cage_strain)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!(temperature >= 0.0)) return fail(c, SN_ERR_INVALID, "temperature must be >= 0");
    sn_params &p = c->params[r];
    p.temperature = temperature;
    if (efield) { p.efield[0] = efield[0]; p.efield[1] = efield[1]; p.efield[2] = efield[2]; }
    p.cage_strain = cage_strain;
    return upload_rep(c, r);
}

int sn_site_energy_batch_device(sn_ctx *c, int32_t r, int64_t n, const int32_t *site,
                                const double *prop, double *dE)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (n < 0) return fail(c, SN_ERR_INVALID, "negative count");
    if (n == 0) return SN_OK;
    if (!site || !prop || !dE) return fail(c, SN_ERR_INVALID, "null buffer");
    long long nb = (n + 255) / 256;
    hipLaunchKernelGGL(site_energy_batch_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, c->d_lat,
                       c->d_reps, r, c->d_offs, c->noff, (long long)n, site, prop, dE);
    HIPCHK(c, hipGetLastError());
    return SN_OK;
}

int sn_site_energy_batch(sn_ctx *c, int32_t r, int64_t n, const int32_t *site, const double *prop,
                         double *dE)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (n < 0) return fail(c, SN_ERR_INVALID, "negative count");
    if (n == 0) return SN_OK;
    if (!site || !prop || !dE) return fail(c, SN_ERR_INVALID, "null buffer");
    for (int64_t k = 0; k < n; k++)
        if (site[k] < 0 || site[k] >= c->reps[r].N) return fail(c, SN_ERR_RANGE, "site index out of range");
    // inputs and outputs go through the context's staging buffer (no
    // allocation once it has grown to the largest batch)
    StageLayout lay;
    const size_t o_p = lay.take(n * 3 * sizeof(double)), o_e = lay.take(n * sizeof(double)),
                 o_s = lay.take(n * sizeof(int32_t));
    if ((st = ensure_stage(c, lay.end))) return st;
    double *dp = stage_ptr<double>(c, o_p), *de = stage_ptr<double>(c, o_e);
    int32_t *ds = stage_ptr<int32_t>(c, o_s);
    HIPCHK(c, hipMemcpyAsync(ds, site, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dp, prop, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((st = sn_site_energy_batch_device(c, r, n, ds, dp, de))) return st;
    HIPCHK(c, hipMemcpyAsync(dE, de, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SN_OK;
}

static SweepArgs sweep_args(sn_ctx *c)
{
    SweepArgs a;
    a.lat = c->d_lat; a.reps = c->d_reps; a.tiles = c->d_tiles; a.shapes = c->d_shapes;
    a.class_start = c->d_class_start; a.class_ri = c->d_class_ri; a.chunk_col = c->d_chunk_col;
    a.gsteps = c->d_gsteps; a.gchunk = c->d_gchunk;
    a.max_gsteps = c->max_gsteps; a.max_chunks = c->max_chunks;
    a.prop_stride = c->prop_chunk;
    a.pairs = c->d_pairs; a.sweep_ctr = c->d_sweep; a.counters = c->d_counters;
    a.npairs = c->npairs; a.rc = c->rc; a.two_d = c->two_d ? 1 : 0;
    a.max_region = c->gather_region;
    a.max_tile_sites = c->max_tile_sites;
    a.max_cls = c->max_cls;
    return a;
}

// Enqueue one sweep: one tile-sweep launch per non-empty coarse colour, then
// the sweep-counter bump.  events (optional) bracket each tile-sweep launch.
static int launch_coarse(sn_ctx *c, bool mixed, int first, int n)
{
    if (c->use_field) {
        FieldArgs fa;
        fa.s = sweep_args(c);
        fa.steps = c->d_field_steps;
        fa.tile_coord = c->d_tile_coord;
        fa.tile_counts = c->d_counters;
        fa.phaseA_doubles = field_phaseA(c, mixed);
        fa.region_doubles = field_region_doubles(c);
        fa.tile_doubles = c->max_tile_padded;
        hipLaunchKernelGGL(pick_field(c->rc, c->two_d, mixed), dim3(n), dim3(kBlock), field_lds(c, mixed),
                           c->stream, fa, first, n);
    } else {
        hipLaunchKernelGGL(pick_sweep(c->G, c->NPL, mixed, c->blk, c->gather_region, c->prop_chunk), dim3(n), dim3(c->blk),
                           sweep_lds(c, mixed),
                           c->stream, sweep_args(c), first, n);
    }
    HIPCHK(c, hipGetLastError());
    return SN_OK;
}

// Sweeps per whole-lattice launch (bounds one launch's duration).
constexpr int kTorusSweeps = 256;

// nsw whole-lattice sweeps of every replica in one launch, then the
// sweep-counter bump; events (optional) bracket the sweep kernel.
// Two new events, appended to *events as soon as each exists (so that the
// owner of the list destroys them whatever fails later).
static int event_pair(sn_ctx *c, std::vector<hipEvent_t> *events, hipEvent_t *e0, hipEvent_t *e1)
{
    HIPCHK(c, hipEventCreate(e0));
    events->push_back(*e0);
    HIPCHK(c, hipEventCreate(e1));
    events->push_back(*e1);
    return SN_OK;
}

static int launch_torus(sn_ctx *c, int nsw, std::vector<hipEvent_t> *events)
{
    const bool mixed = any_mixed(c);
    TorusArgs t;
    t.lat = c->d_lat; t.reps = c->d_reps; t.pairs = c->d_pairs; t.npairs = c->npairs;
    t.sites = c->d_torus_sites; t.site_off = c->d_torus_off; t.col_start = c->d_torus_cs;
    t.ncol = c->torus_ncol; t.sweep_ctr = c->d_sweep; t.counters = c->d_rep_counters;
    t.max_sites = c->torus_ms; t.prop_cap = c->torus_pc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (events && (event_pair(c, events, &e0, &e1) || hipEventRecord(e0, c->stream) != hipSuccess))
        return fail(c, SN_ERR_HIP, "timing events");
    hipLaunchKernelGGL(pick_torus(c->torus_G, c->torus_NPL, mixed, c->torus_blk), dim3(c->R), dim3(c->torus_blk),
                       torus_lds(c, mixed), c->stream, t, nsw);
    HIPCHK(c, hipGetLastError());
    if (events) HIPCHK(c, hipEventRecord(e1, c->stream));
    hipLaunchKernelGGL(bump_sweep_kernel, dim3((c->R + 255) / 256), dim3(256), 0, c->stream, c->d_sweep, c->R,
                       (unsigned)nsw);
    HIPCHK(c, hipGetLastError());
    return SN_OK;
}

static int enqueue_sweep(sn_ctx *c, std::vector<hipEvent_t> *events)
{
    const bool mixed = any_mixed(c);
    for (int cc = 0; cc < 8; cc++) {
        int first = c->sched.coarse_start[cc], n = c->sched.coarse_start[cc + 1] - first;
        if (n <= 0) continue;
        int st;
        if (events) {
            hipEvent_t e0, e1;
            if ((st = event_pair(c, events, &e0, &e1))) return st;
            HIPCHK(c, hipEventRecord(e0, c->stream));
            if ((st = launch_coarse(c, mixed, first, n))) return st;
            HIPCHK(c, hipEventRecord(e1, c->stream));
        } else if ((st = launch_coarse(c, mixed, first, n))) return st;
    }
    hipLaunchKernelGGL(bump_sweep_kernel, dim3((c->R + 255) / 256), dim3(256), 0, c->stream, c->d_sweep, c->R, 1u);
    HIPCHK(c, hipGetLastError());
    return SN_OK;
}

// Sweeps per replay of the multi-sweep graph: small lattices, whose sweeps
// take tens of microseconds, would otherwise be bound by one graph launch
// per sweep on the host.
constexpr int kGraphSweeps = 16;

// capture `sweeps` consecutive sweeps into *g / *exec
static int capture_sweeps(sn_ctx *c, int sweeps, hipGraph_t *graph, hipGraphExec_t *exec)
{
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int st = SN_OK;
    for (int i = 0; i < sweeps && !st; i++) st = enqueue_sweep(c, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st) { if (g) (void)hipGraphDestroy(g); return st; }
    if (e != hipSuccess) return fail(c, SN_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    *graph = g;
    HIPCHK(c, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return SN_OK;
}

static int build_graph(sn_ctx *c)
{
    free_graph(c);
    int st = capture_sweeps(c, 1, &c->graph, &c->graph_exec);
    if (st) return st;
    c->graph_mixed = any_mixed(c) ? 1 : 0;
    return SN_OK;
}

int sn_mc_sweeps(sn_ctx *c, int32_t nsweeps)
{
    if (!c) return SN_ERR_INVALID;
    if (!c->can_sweep) return fail(c, SN_ERR_UNSUPPORTED, "sweeps unavailable: " + c->sweep_err);
    if (nsweeps < 0) return fail(c, SN_ERR_INVALID, "negative sweep count");
    if (c->torus) {
        for (int left = nsweeps; left > 0; left -= kTorusSweeps) {
            int st = launch_torus(c, std::min(left, kTorusSweeps), nullptr);
            if (st) return st;
        }
        for (int r = 0; r < c->R; r++) c->sweeps_done[r] += (unsigned long long)nsweeps;
        return SN_OK;
    }
    const bool use_graph = c->tuning.use_graph >= 0;  // 0 (auto) and 1 capture a graph
    if (use_graph) {
        int want = any_mixed(c) ? 1 : 0;
        if (!c->graph_exec || c->graph_mixed != want) {
            int st = build_graph(c);
            if (st) return st;
        }
        int left = nsweeps;
        if (left >= kGraphSweeps) {
            if (!c->graph_exec_k) {
                int st = capture_sweeps(c, kGraphSweeps, &c->graph_k, &c->graph_exec_k);
                if (st) return st;
            }
            for (; left >= kGraphSweeps; left -= kGraphSweeps) HIPCHK(c, hipGraphLaunch(c->graph_exec_k, c->stream));
        }
        for (; left > 0; left--) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    } else {
        for (int i = 0; i < nsweeps; i++) {
            int st = enqueue_sweep(c, nullptr);
            if (st) return st;
        }
    }
    for (int r = 0; r < c->R; r++) c->sweeps_done[r] += (unsigned long long)nsweeps;
    return SN_OK;
}

int sn_sync(sn_ctx *c)
{
    if (!c) return SN_ERR_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SN_OK;
}

int sn_time_sweep_kernel(sn_ctx *c, int32_t nsweeps, double *mean_ms, int32_t *launches)
{
    if (!c || nsweeps < 1 || !mean_ms) return fail(c, SN_ERR_INVALID, "bad arguments");
    if (!c->can_sweep) return fail(c, SN_ERR_UNSUPPORTED, "sweeps unavailable: " + c->sweep_err);
    EventList evl;   // destroyed on every return path
    std::vector<hipEvent_t> &ev = evl.ev;
    int st = SN_OK;
    for (int i = 0; i < nsweeps && st == SN_OK; i++) st = c->torus ? launch_torus(c, 1, &ev) : enqueue_sweep(c, &ev);
    if (st) return st;
    for (int r = 0; r < c->R; r++) c->sweeps_done[r] += (unsigned long long)nsweeps;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        tot += ms;
    }
    int nl = (int)(ev.size() / 2);
    *mean_ms = nl ? tot / nl : 0.0;
    if (launches) *launches = nl;
    return SN_OK;
}

int sn_mc_replay(sn_ctx *c, int32_t r, int64_t n, const int32_t *site, const double *prop,
                 const double *u, const uint8_t *u_drawn, const uint8_t *ref_accept,
                 uint8_t *decision_out, uint64_t *mismatches, uint64_t *near_ties)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (n < 0) return fail(c, SN_ERR_INVALID, "negative count");
    if (n == 0) { if (mismatches) *mismatches = 0; if (near_ties) *near_ties = 0; return SN_OK; }
    if (!site || !prop || !u || !u_drawn || !ref_accept || !decision_out)
        return fail(c, SN_ERR_INVALID, "null buffer");
    for (int64_t k = 0; k < n; k++)
        if (site[k] < 0 || site[k] >= c->reps[r].N) return fail(c, SN_ERR_RANGE, "site index out of range");
    StageLayout lay;
    const size_t o_p = lay.take(n * 3 * sizeof(double)), o_u = lay.take(n * sizeof(double)),
                 o_st = lay.take(4 * sizeof(unsigned long long)), o_s = lay.take(n * sizeof(int32_t)),
                 o_d = lay.take(n), o_a = lay.take(n), o_o = lay.take(n);
    if ((st = ensure_stage(c, lay.end))) return st;
    double *dp = stage_ptr<double>(c, o_p), *du = stage_ptr<double>(c, o_u);
    unsigned long long *dstat = stage_ptr<unsigned long long>(c, o_st);
    int32_t *ds = stage_ptr<int32_t>(c, o_s);
    uint8_t *dd = stage_ptr<uint8_t>(c, o_d), *da = stage_ptr<uint8_t>(c, o_a), *dout = stage_ptr<uint8_t>(c, o_o);
    HIPCHK(c, hipMemcpyAsync(ds, site, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dp, prop, n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(du, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dd, u_drawn, n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(da, ref_accept, n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(replay_kernel, dim3(1), dim3(256), 0, c->stream, c->d_lat, c->d_reps, r, c->d_offs,
                       c->noff, (long long)n, ds, dp, du, dd, da, dout, dstat);
    HIPCHK(c, hipGetLastError());
    unsigned long long stats[4];
    HIPCHK(c, hipMemcpyAsync(decision_out, dout, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(stats, dstat, sizeof(stats), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mismatches) *mismatches = stats[0];
    if (near_ties) *near_ties = stats[1];
    // the replay's decisions count toward the replica's counters
    c->replay_counts[2 * r] += stats[2];
    c->replay_counts[2 * r + 1] += stats[3];
    return SN_OK;
}

int sn_counters(sn_ctx *c, int32_t r, uint64_t *accept, uint64_t *reject, uint64_t *attempted)
{
    int st = check_rep(c, r);
    if (st) return st;
    unsigned long long acc = c->replay_counts[2 * r], rej = c->replay_counts[2 * r + 1];
    const size_t nt = c->sched.tiles.size();
    if (nt) {
        // per-tile slots (each tile's workgroup owns its slot: no atomics)
        std::vector<unsigned long long> tc(2 * nt);
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(tc.data(), c->d_counters, tc.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost));
        for (size_t t = 0; t < nt; t++)
            if (c->sched.tiles[t].replica == r) { acc += tc[2 * t]; rej += tc[2 * t + 1]; }
    }
    if (c->torus) {
        unsigned long long rc_[2];
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(rc_, c->d_rep_counters + 2 * r, sizeof(rc_), hipMemcpyDeviceToHost));
        acc += rc_[0];
        rej += rc_[1];
    }
    if (accept) *accept = acc;
    if (reject) *reject = rej;
    if (attempted) *attempted = c->sweeps_done[r] * (unsigned long long)c->reps[r].N;
    return SN_OK;
}

int sn_reset_counters(sn_ctx *c)
{
    if (!c) return SN_ERR_INVALID;
    std::vector<unsigned long long> z(2 * std::max<size_t>(c->sched.tiles.size(), 1), 0ull);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_counters, z.data(), z.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    if (c->d_rep_counters) {
        std::vector<unsigned long long> zr(2 * c->R, 0ull);
        HIPCHK(c, hipMemcpy(c->d_rep_counters, zr.data(), zr.size() * sizeof(unsigned long long),
                            hipMemcpyHostToDevice));
    }
    std::fill(c->sweeps_done.begin(), c->sweeps_done.end(), 0ull);
    std::fill(c->replay_counts.begin(), c->replay_counts.end(), 0ull);
    return SN_OK;
}

int sn_get_sweep_index(sn_ctx *c, int32_t r, uint64_t *index)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!index) return fail(c, SN_ERR_INVALID, "null output");
    unsigned v = 0;
    HIPCHK(c, hipMemcpyAsync(&v, c->d_sweep + r, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *index = v;
    return SN_OK;
}

int sn_set_sweep_index(sn_ctx *c, int32_t r, uint64_t index)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (index > 0xFFFFFFFFull) return fail(c, SN_ERR_RANGE, "sweep index beyond 32 bits");
    const unsigned v = (unsigned)index;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_sweep + r, &v, sizeof(v), hipMemcpyHostToDevice));
    return SN_OK;
}

// Energy parts and parameter derivatives (kNPart sums per replica into
// d_result) and the gradient into dg when non-null.  Lattices with a sweep
// schedule and DipoleCutOff <= 5 take the tiled kernel (one workgroup per
// tile, LDS-staged stencil), the others the per-site gather kernel.
// r >= 0: replica r alone, dg holds its N x 3 rows.  r < 0: every replica,
// dg holds all sites' rows in lattice order (replica r from row base_r), and
// replica r's sums land at d_result + r * kNPart.
static int run_energy(sn_ctx *c, int r, double *dg, double *fixed_partials = nullptr,
                      size_t fixed_cap = 0)
{
    static_assert(kEnergyParts == kNPart, "energy kernels share the partial-sum layout");
    const bool all = r < 0;
    int st;
    // the partial sums: the context's growable buffer, or a caller-given fixed
    // one of fixed_cap doubles (checked, never grown)
    auto partials = [&](size_t n, double **out) -> int {
        if (fixed_partials) {
            if (n > fixed_cap)
                return fail(c, SN_ERR_RANGE, "partial sums need " + std::to_string(n) + " doubles, the fixed scratch holds " +
                                                 std::to_string(fixed_cap));
            *out = fixed_partials;
            return SN_OK;
        }
        const int e = ensure_partials(c, n);
        *out = c->d_partials;
        return e;
    };
    double *part = nullptr;
    if (c->energy_tiled) {
        const int t0 = all ? 0 : c->rep_tile_off[r], t1 = all ? c->rep_tile_off[c->R] : c->rep_tile_off[r + 1];
        const int nblk = t1 - t0;
        if ((st = partials((size_t)nblk * kNPart, &part))) return st;
        // one launch: the mixed-length variant serves uniform replicas too
        const bool mixed = all ? any_mixed(c) : c->reps[r].mixed != 0;
        EnergyArgs ea;
        ea.lat = c->d_lat; ea.reps = c->d_reps; ea.tiles = c->d_tiles; ea.shapes = c->d_shapes;
        ea.tile_list = c->d_tile_list + t0;
        ea.grad = dg; ea.grad_origin = all ? 0 : c->reps[r].base;
        ea.partials = part;
        ea.region_doubles = field_region_doubles(c);
        const size_t lds = (size_t)(mixed ? 4 : 3) * ea.region_doubles * sizeof(double);
        hipLaunchKernelGGL(pick_energy(c->rc, c->two_d, mixed), dim3(nblk), dim3(kBlock), lds, c->stream, ea);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(reduce_partials_kernel, dim3((kNPart + kReduceCols - 1) / kReduceCols, all ? c->R : 1),
                           dim3(256), 0, c->stream, part, nblk, kNPart, c->d_result,
                           all ? (const int *)c->d_tile_seg : (const int *)nullptr);
        HIPCHK(c, hipGetLastError());
        return SN_OK;
    }
    // per-site gather kernel, one replica at a time (the partials buffer is
    // reused in stream order)
    for (int q = all ? 0 : r; q < (all ? c->R : r + 1); q++) {
        const RepHost &h = c->reps[q];
        const int nblk = (int)((h.N + 255) / 256);
        if ((st = partials((size_t)nblk * kNPart, &part))) return st;
        double *g = dg ? dg + 3 * (all ? h.base : 0) : nullptr;
        hipLaunchKernelGGL(energy_grad_kernel, dim3(nblk), dim3(256), 0, c->stream, c->d_lat, c->d_reps, q,
                           c->d_offs, c->noff, g, part);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(reduce_partials_kernel, dim3((kNPart + kReduceCols - 1) / kReduceCols), dim3(256), 0,
                           c->stream, part, nblk, kNPart, c->d_result + (all ? q * kNPart : 0),
                           (const int *)nullptr);
        HIPCHK(c, hipGetLastError());
    }
    return SN_OK;
}

// n doubles of d_result into the pinned host copy, after the stream's work
static int fetch_result(sn_ctx *c, size_t n)
{
    HIPCHK(c, hipMemcpyAsync(c->h_result, c->d_result, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SN_OK;
}

// total energy from its four parts, compensated (Neumaier)
static double energy_total(const double *parts)
{
    double s = 0.0, cmp = 0.0;
    for (int i = 0; i < 4; i++) {
        const double t = s + parts[i];
        cmp += (std::fabs(s) >= std::fabs(parts[i])) ? (s - t) + parts[i] : (parts[i] - t) + s;
        s = t;
    }
    return s + cmp;
}

int sn_gradient(sn_ctx *c, int32_t r, double *grad_p, double dparams[6], double *energy)
{
    int st = check_rep(c, r);
    if (st) return st;
    const RepHost &h = c->reps[r];
    double *dg = nullptr;
    if (grad_p) {
        if ((st = ensure_stage(c, (size_t)h.N * 3 * sizeof(double)))) return st;
        dg = stage_ptr<double>(c, 0);
    }
    if ((st = run_energy(c, r, dg))) return st;
    if (grad_p)
        HIPCHK(c, hipMemcpyAsync(grad_p, dg, h.N * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if ((st = fetch_result(c, kNPart))) return st;
    if (dparams) for (int i = 0; i < 6; i++) dparams[i] = c->h_result[4 + i];
    if (energy) *energy = energy_total(c->h_result);
    return SN_OK;
}

int sn_gradient_device(sn_ctx *c, int32_t r, void *grad_dev, double dparams[6], double *energy)
{
    int st = check_rep(c, r);
    if (st) return st;
    if ((st = run_energy(c, r, (double *)grad_dev)) || (st = fetch_result(c, kNPart))) return st;
    if (dparams) for (int i = 0; i < 6; i++) dparams[i] = c->h_result[4 + i];
    if (energy) *energy = energy_total(c->h_result);
    return SN_OK;
}

int sn_gradient_all_device(sn_ctx *c, void *grad_dev, double *dparams, double *energy)
{
    if (!c) return SN_ERR_INVALID;
    int st;
    if ((st = run_energy(c, -1, (double *)grad_dev)) || (st = fetch_result(c, (size_t)c->R * kNPart))) return st;
    for (int r = 0; r < c->R; r++) {
        const double *res = c->h_result + (size_t)r * kNPart;
        if (dparams) for (int i = 0; i < 6; i++) dparams[6 * r + i] = res[4 + i];
        if (energy) energy[r] = energy_total(res);
    }
    return SN_OK;
}

int sn_gradient_all_device_noalloc(sn_ctx *c, void *grad_dev, double *dparams, double *energy)
{
    if (!c) return SN_ERR_INVALID;
    if (!c->d_scratch) return fail(c, SN_ERR_HIP, "no-allocation scratch not resolved");
    // the partials go to this device's fixed `scratch` array (every row the
    // reduction reads is written first, so nothing needs clearing); the sums
    // land in d_result and the pinned h_result, both allocated by sn_create
    int st;
    if ((st = run_energy(c, -1, (double *)grad_dev, c->d_scratch, SN_NOALLOC_PARTIALS)) ||
        (st = fetch_result(c, (size_t)c->R * kNPart)))
        return st;
    for (int r = 0; r < c->R; r++) {
        const double *res = c->h_result + (size_t)r * kNPart;
        if (dparams) for (int i = 0; i < 6; i++) dparams[6 * r + i] = res[4 + i];
        if (energy) energy[r] = energy_total(res);
    }
    return SN_OK;
}

int sn_energy(sn_ctx *c, int32_t r, double *energy, double parts[4])
{
    int st = check_rep(c, r);
    if (st) return st;
    if ((st = run_energy(c, r, nullptr)) || (st = fetch_result(c, kNPart))) return st;
    if (parts) for (int i = 0; i < 4; i++) parts[i] = c->h_result[i];
    if (energy) *energy = energy_total(c->h_result);
    return SN_OK;
}

int sn_moment(sn_ctx *c, int32_t r, double M[3])
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!M) return fail(c, SN_ERR_INVALID, "null output");
    const RepHost &h = c->reps[r];
    int nblk = (int)std::min<long long>((h.N + 255) / 256, 1024);
    if ((st = ensure_partials(c, (size_t)nblk * 3))) return st;
    hipLaunchKernelGGL(moment_kernel, dim3(nblk), dim3(256), 0, c->stream, c->d_lat, c->d_reps, r, c->d_partials);
    HIPCHK(c, hipGetLastError());
    launch_reduce(c, nblk, 3);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(M, c->d_result, 3 * sizeof(double), hipMemcpyDeviceToHost));
    return SN_OK;
}

// Offsets of radial_order_parameter (src/starrynight-analysis.c:552-561): every
// (dx,dy,dz) in [-9,9]^3 with r^2 <= 80, grouped by r^2; n = r/|r| in FP64.
// The reference's r^2 = 81 bin is accumulated but never zeroed or printed
// (its arrays hold 82 entries, its loops run to 80), so it is left out.
static int ensure_radial(sn_ctx *c)
{
    if (c->d_radial) return SN_OK;
    std::vector<RadialOff> offs;
    std::vector<int> start(kRadialBins + 1, 0);
    for (int b = 0; b < kRadialBins; b++) {
        start[b] = (int)offs.size();
        for (int dx = -9; dx <= 9; dx++)
            for (int dy = -9; dy <= 9; dy++)
                for (int dz = -9; dz <= 9; dz++) {
                    if (dx * dx + dy * dy + dz * dz != b) continue;
                    double d = b ? std::sqrt((double)b) : 1.0;
                    offs.push_back(RadialOff{dx, dy, dz, 0, dx / d, dy / d, dz / d, 0.0});
                }
    }
    start[kRadialBins] = (int)offs.size();
    HIPCHK(c, hipMalloc((void **)&c->d_radial, offs.size() * sizeof(RadialOff)));
    HIPCHK(c, hipMalloc((void **)&c->d_radial_bins, start.size() * sizeof(int)));
    HIPCHK(c, hipMemcpy(c->d_radial, offs.data(), offs.size() * sizeof(RadialOff), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_radial_bins, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
    c->radial_bins = start;
    return SN_OK;
}

int sn_radial_order(sn_ctx *c, int32_t r, double fe[81], double afe[81], int64_t count[81])
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!fe || !afe || !count) return fail(c, SN_ERR_INVALID, "null output");
    if ((st = ensure_radial(c))) return st;
    const RepHost &h = c->reps[r];
    int nblk = (int)std::min<long long>((h.N + 255) / 256, 2048);
    if ((st = ensure_partials(c, (size_t)nblk * 2 * kRadialBins))) return st;
    hipLaunchKernelGGL(radial_order_kernel, dim3(nblk), dim3(256), 0, c->stream, c->d_lat, c->d_reps, r,
                       c->d_radial, c->d_radial_bins, c->d_partials);
    HIPCHK(c, hipGetLastError());
    launch_reduce(c, nblk, 2 * kRadialBins);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double sums[2 * kRadialBins];
    HIPCHK(c, hipMemcpy(sums, c->d_result, sizeof(sums), hipMemcpyDeviceToHost));
    for (int b = 0; b < kRadialBins; b++) {
        const long long cnt = h.N * (long long)(c->radial_bins[b + 1] - c->radial_bins[b]);
        count[b] = cnt;
        fe[b] = cnt ? sums[2 * b] / (double)cnt : 0.0;
        afe[b] = cnt ? sums[2 * b + 1] / (double)cnt : 0.0;
    }
    return SN_OK;
}

// Offsets of dipole_potential (kind 0: 0 < |r| <= 6, src/starrynight-analysis.c:72-80)
// or dipole_electricfield (kind = cutoff: 0 < |r| <= cutoff, :403-413), in the
// reference's dx, dy, dz loop order; the dz range runs when Z == 1.
static int ensure_stencil(sn_ctx *c, int kind)
{
    if (c->d_stencil && c->stencil_kind == kind) return SN_OK;
    const int R = kind == 0 ? 6 : kind;
    std::vector<StencilOff> offs;
    for (int dx = -R; dx <= R; dx++)
        for (int dy = -R; dy <= R; dy++)
            for (int dz = -R; dz <= R; dz++) {
                const int d2 = dx * dx + dy * dy + dz * dz;
                if (d2 == 0 || d2 > R * R) continue;
                const double d = std::sqrt((double)d2), w = 1.0 / (d * d * d);
                if (kind == 0) offs.push_back(StencilOff{dx, dy, dz, 0, dx * w, dy * w, dz * w, w});
                else offs.push_back(StencilOff{dx, dy, dz, 0, dx / d, dy / d, dz / d, w});
            }
    if (c->d_stencil) (void)hipFree(c->d_stencil);
    c->d_stencil = nullptr;
    HIPCHK(c, hipMalloc((void **)&c->d_stencil, offs.size() * sizeof(StencilOff)));
    HIPCHK(c, hipMemcpy(c->d_stencil, offs.data(), offs.size() * sizeof(StencilOff), hipMemcpyHostToDevice));
    c->stencil_kind = kind;
    c->stencil_n = (int)offs.size();
    return SN_OK;
}

static int run_stencil(sn_ctx *c, int32_t r, int kind, double *host_out, int width)
{
    const RepHost &h = c->reps[r];
    int st;
    if ((st = ensure_stencil(c, kind))) return st;
    const size_t need = (size_t)h.N * width;
    if (need > c->obs_cap) {
        if (c->d_obs) (void)hipFree(c->d_obs);
        c->d_obs = nullptr;
        HIPCHK(c, hipMalloc((void **)&c->d_obs, need * sizeof(double)));
        c->obs_cap = need;
    }
    const unsigned nb = (unsigned)((h.N + 255) / 256);
    if (kind == 0)
        hipLaunchKernelGGL(stencil_observable_kernel<false>, dim3(nb), dim3(256), 0, c->stream, c->d_lat, c->d_reps,
                           r, c->d_stencil, c->stencil_n, c->d_obs);
    else
        hipLaunchKernelGGL(stencil_observable_kernel<true>, dim3(nb), dim3(256), 0, c->stream, c->d_lat, c->d_reps,
                           r, c->d_stencil, c->stencil_n, c->d_obs);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(host_out, c->d_obs, need * sizeof(double), hipMemcpyDeviceToHost));
    return SN_OK;
}

int sn_dipole_potential(sn_ctx *c, int32_t r, double *pot)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!pot) return fail(c, SN_ERR_INVALID, "null output");
    return run_stencil(c, r, 0, pot, 1);
}

int sn_dipole_efield(sn_ctx *c, int32_t r, int32_t cutoff, double *E)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!E) return fail(c, SN_ERR_INVALID, "null output");
    if (cutoff < 1 || cutoff > 9) return fail(c, SN_ERR_RANGE, "field cutoff must be in 1..9");
    return run_stencil(c, r, cutoff, E, 3);
}

int sn_recombination(sn_ctx *c, int32_t r, double out[8])
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!out) return fail(c, SN_ERR_INVALID, "null output");
    const RepHost &h = c->reps[r];
    std::vector<double> pot(h.N);
    if ((st = run_stencil(c, r, 0, pot.data(), 1))) return st;   // leaves the potential in d_obs
    const int nblk = (int)std::min<long long>((h.N + 255) / 256, 1024);
    if ((st = ensure_partials(c, (size_t)nblk * kRecombParts))) return st;
    // the reference's constants: 0.165 eV per internal unit, screened by 5; kT = 0.025 eV
    hipLaunchKernelGGL(recombination_kernel, dim3(nblk), dim3(256), 0, c->stream, c->d_obs, (long long)h.N,
                       0.165 / 5, 1 / (0.025), c->d_partials);
    HIPCHK(c, hipGetLastError());
    launch_reduce(c, nblk, kRecombParts);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double v[kRecombParts];
    HIPCHK(c, hipMemcpy(v, c->d_result, sizeof(v), hipMemcpyDeviceToHost));
    const double N = (double)h.N;
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; out[3] = v[3];
    out[4] = N * N / (v[0] * v[1]);
    out[5] = N * v[4] / (v[2] * v[3]);
    out[6] = v[2] / v[2];   // sum of the normalised electron density (1 up to rounding)
    out[7] = v[3] / v[3];
    return SN_OK;
}

int sn_schedule_info(sn_ctx *c, int32_t r, int64_t *n_sites, int32_t *n_phases)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!c->can_sweep) return fail(c, SN_ERR_UNSUPPORTED, "sweeps unavailable: " + c->sweep_err);
    std::vector<int32_t> order, ps;
    if (c->torus) { order = c->torus_order[r]; ps = c->torus_cstart[r]; }
    else sweep_order(c->sched, c->geom[r], r, order, ps);
    if (n_sites) *n_sites = (int64_t)order.size();
    if (n_phases) *n_phases = (int32_t)ps.size() - 1;
    return SN_OK;
}

int sn_schedule_order(sn_ctx *c, int32_t r, int32_t *order, int32_t *phase_start)
{
    int st = check_rep(c, r);
    if (st) return st;
    if (!c->can_sweep) return fail(c, SN_ERR_UNSUPPORTED, "sweeps unavailable: " + c->sweep_err);
    std::vector<int32_t> o, ps;
    if (c->torus) { o = c->torus_order[r]; ps = c->torus_cstart[r]; }
    else sweep_order(c->sched, c->geom[r], r, o, ps);
    if (order) std::memcpy(order, o.data(), o.size() * sizeof(int32_t));
    if (phase_start) std::memcpy(phase_start, ps.data(), ps.size() * sizeof(int32_t));
    return SN_OK;
}

int sn_stream(sn_ctx *c, void **s)
{
    if (!c || !s) return SN_ERR_INVALID;
    *s = (void *)c->stream;
    return SN_OK;
}

int sn_kernel_choice(sn_ctx *c, int32_t *kernel, int32_t *block_threads, int32_t *lanes, int32_t *pairs)
{
    if (!c) return SN_ERR_INVALID;
    if (!c->can_sweep) return fail(c, SN_ERR_UNSUPPORTED, "sweeps unavailable: " + c->sweep_err);
    const int k = c->torus ? 3 : c->use_field ? 1 : 2;
    if (kernel) *kernel = k;
    if (block_threads) *block_threads = k == 3 ? c->torus_blk : k == 1 ? kBlock : c->blk;
    if (lanes) *lanes = k == 3 ? c->torus_G : k == 1 ? kFieldG : c->G;
    if (pairs) *pairs = k == 3 ? c->torus_NPL : k == 1 ? 0 : c->NPL;
    return SN_OK;
}

int sn_sweep_config(sn_ctx *c, int32_t *tiles, int32_t *launches, int32_t *lanes, int32_t *fine)
{
    if (!c) return SN_ERR_INVALID;
    int nl = 0;
    for (int cc = 0; cc < 8; cc++)
        if (c->sched.coarse_start[cc + 1] > c->sched.coarse_start[cc]) nl++;
    int maxcol = 0;
    for (const TileShape &s : c->sched.shapes) maxcol = std::max(maxcol, s.ncol);
    if (c->torus) { nl = 1; maxcol = c->torus_ncol; }
    if (tiles) *tiles = c->torus ? 0 : (int32_t)c->sched.tiles.size();
    if (launches) *launches = nl;
    if (lanes) *lanes = c->torus ? c->torus_G : c->use_field ? kFieldG : c->G;
    if (fine) *fine = maxcol;
    return SN_OK;
}

}  // extern "C"

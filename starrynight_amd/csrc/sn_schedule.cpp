// sn_schedule.cpp -- host-side colour schedule for the tile sweep.
#include "sn_schedule.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <tuple>

namespace sn {

std::vector<Offset> neighbour_offsets(int rc, bool two_d)
{
    // gen_neighbour (src/starrynight-montecarlo-core.c:38-72); the float
    // sqrt cutoff test there is the integer test |r|^2 <= rc^2.
    std::vector<Offset> v;
    int zc = two_d ? 0 : rc;
    for (int dx = -rc; dx <= rc; dx++)
        for (int dy = -rc; dy <= rc; dy++)
            for (int dz = -zc; dz <= zc; dz++) {
                int r2 = dx * dx + dy * dy + dz * dz;
                if (r2 == 0 || r2 > rc * rc) continue;
                v.push_back({dx, dy, dz, r2});
            }
    return v;
}

std::vector<Offset> neighbour_pairs(int rc, bool two_d)
{
    std::vector<Offset> all = neighbour_offsets(rc, two_d), p;
    for (const Offset &o : all) {
        bool pos = o.dx > 0 || (o.dx == 0 && (o.dy > 0 || (o.dy == 0 && o.dz > 0)));
        if (pos) p.push_back(o);
    }
    std::stable_sort(p.begin(), p.end(),
                     [](const Offset &a, const Offset &b) { return (a.r2 == 1) > (b.r2 == 1); });
    return p;
}

static inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }
static inline int pmod(int a, int b) { int r = a % b; return r < 0 ? r + b : r; }

int Sublattice::colour(int x, int y, int z) const
{
    int k3 = floordiv(z, f);
    x -= k3 * c; y -= k3 * e; z -= k3 * f;
    int k2 = floordiv(y, d);
    x -= k2 * b; y -= k2 * d;
    x = pmod(x, a);
    return (x * d + y) * f + z;
}

int Sublattice::min_norm2() const
{
    // exact up to 64: smallest |v|^2 over nonzero lattice vectors, found by
    // testing membership (colour(v) == colour(0)) of every integer vector in a
    // ball whose radius grows until one is found (returns 64 when none is
    // shorter; cutoffs up to 7 need only this range).
    const int c0 = colour(0, 0, 0);
    int bound = std::min(64, std::min(a * a, std::min(b * b + d * d, c * c + e * e + f * f)));
    for (int r2 = 1; r2 <= bound; r2++) {
        int R = (int)std::sqrt((double)r2) + 1;
        for (int x = -R; x <= R; x++)
            for (int y = -R; y <= R; y++)
                for (int z = -R; z <= R; z++)
                    if (x * x + y * y + z * z == r2 && colour(x, y, z) == c0) return r2;
    }
    return bound;
}

// Shortest nonzero in-plane (z == 0) lattice vector, squared.
static int min_norm2_plane(const Sublattice &s)
{
    const int c0 = s.colour(0, 0, 0);
    int bound = std::min(64, std::min(s.a * s.a, s.b * s.b + s.d * s.d));
    for (int r2 = 1; r2 <= bound; r2++) {
        int R = (int)std::sqrt((double)r2) + 1;
        for (int x = -R; x <= R; x++)
            for (int y = -R; y <= R; y++)
                if (x * x + y * y == r2 && s.colour(x, y, 0) == c0) return r2;
    }
    return bound;
}

std::string Sublattice::describe() const
{
    std::ostringstream s;
    s << "(" << a << ",0,0),(" << b << "," << d << ",0),(" << c << "," << e << "," << f << ")";
    return s.str();
}

Sublattice make_sublattice(int rc, bool two_d, int kind)
{
    int P = rc + 1;
    if (kind == 1) return two_d ? Sublattice{P, 0, 0, P, 0, 1} : Sublattice{P, 0, 0, P, 0, P};
    if (kind == 3 && !two_d && rc <= 3) return Sublattice{4, 0, 2, 4, 2, 2};
    // Minimum-index search over lower-triangular bases.  In 2D the z vector is
    // (0,0,1), which never changes a colour because dz == 0 there.
    int lim = two_d ? P * P : P * P * P;
    for (int idx = 1; idx <= lim; idx++) {
        for (int a = 1; a <= idx; a++) {
            if (idx % a) continue;
            for (int d = 1; d <= idx / a; d++) {
                if ((idx / a) % d) continue;
                int f = idx / (a * d);
                if (two_d && f != 1) continue;
                for (int b = 0; b < a; b++)
                    for (int c = 0; c < (two_d ? 1 : a); c++)
                        for (int e = 0; e < (two_d ? 1 : d); e++) {
                            Sublattice s{a, b, c, d, e, f};
                            // in 2D only in-plane vectors matter
                            int mn = two_d ? min_norm2_plane(s) : s.min_norm2();
                            if (mn > rc * rc) return s;
                        }
            }
        }
    }
    return make_sublattice(rc, two_d, 1);
}

bool split_axis(int L, int rc, int target, int tmax, std::vector<int> &start, std::vector<int> &size)
{
    start.clear(); size.clear();
    if (L == 1) { start.push_back(0); size.push_back(1); return true; }
    if (target < rc) target = rc;
    // the even tile count whose mean edge is closest to the target, first
    // among splits whose largest tile does not exceed the target (a larger
    // tile would raise every tile's LDS share in a context whose lengths are
    // not multiples of the target), then among all valid splits
    int best_n = -1;
    for (int pass = 0; pass < 2 && best_n < 0; pass++) {
        double best_cost = 1e30;
        for (int n = 2; n <= L; n += 2) {
            int lo = L / n, hi = (L + n - 1) / n;
            if (lo < rc) break;
            if (hi + 2 * rc > L || hi > tmax || (pass == 0 && hi > target)) continue;
            double cost = std::fabs((double)L / n - target);
            if (cost < best_cost - 1e-9) { best_cost = cost; best_n = n; }
        }
    }
    if (best_n < 0) return false;
    int base = L / best_n, extra = L % best_n, pos = 0;
    for (int i = 0; i < best_n; i++) {
        // spread the larger tiles evenly
        int sz = base + (((i + 1) * extra) / best_n - (i * extra) / best_n);
        start.push_back(pos); size.push_back(sz); pos += sz;
    }
    return pos == L;
}

bool build_schedule(const std::vector<ReplicaGeom> &reps, int rc, const int target[3],
                    int colouring_kind, int chunk, int max_region, int max_tile, Schedule &out)
{
    out = Schedule();
    out.rc = rc;
    if (reps.empty()) { out.error = "no replicas"; return false; }
    if (rc < 1 || rc > 8) { out.error = "DipoleCutOff must be in 1..8"; return false; }
    out.two_d = reps[0].Z == 1;
    for (const ReplicaGeom &g : reps)
        if ((g.Z == 1) != out.two_d) { out.error = "replicas mix 2D and 3D lattices"; return false; }
    out.lat = make_sublattice(rc, out.two_d, colouring_kind);
    int hz = out.two_d ? 0 : rc;
    // largest tile edge whose tile + halo stays within max_region sites and
    // whose tile stays within max_tile sites
    int tmax = 1;
    while (true) {
        int e = tmax + 1 + 2 * rc, t = tmax + 1;
        long long reg = out.two_d ? (long long)e * e : (long long)e * e * e;
        long long vol = out.two_d ? (long long)t * t : (long long)t * t * t;
        if (reg > max_region || vol > max_tile) break;
        tmax++;
    }

    std::map<std::tuple<int, int, int>, int> shape_id;
    std::vector<std::vector<Tile>> by_colour(8);
    for (size_t r = 0; r < reps.size(); r++) {
        const ReplicaGeom &g = reps[r];
        if (g.X < 1 || g.Y < 1 || g.Z < 1) { out.error = "lattice size must be positive"; return false; }
        std::vector<int> sx, lx, sy, ly, sz, lz;
        if (!split_axis(g.X, rc, target[0], tmax, sx, lx) || !split_axis(g.Y, rc, target[1], tmax, sy, ly) ||
            !split_axis(g.Z, rc, target[2], tmax, sz, lz)) {
            std::ostringstream m;
            m << "lattice " << g.X << "x" << g.Y << "x" << g.Z << " cannot be tiled for DipoleCutOff="
              << rc << " (each axis needs an even number of tiles of edge >= DipoleCutOff with tile + "
              << "2*DipoleCutOff <= length and tile <= " << tmax << ", or Z == 1)";
            out.error = m.str();
            return false;
        }
        for (size_t i = 0; i < sx.size(); i++)
            for (size_t j = 0; j < sy.size(); j++)
                for (size_t k = 0; k < sz.size(); k++) {
                    auto key = std::make_tuple(lx[i], ly[j], lz[k]);
                    auto it = shape_id.find(key);
                    int sid;
                    if (it == shape_id.end()) {
                        sid = (int)out.shapes.size();
                        shape_id[key] = sid;
                        TileShape s;
                        s.tx = lx[i]; s.ty = ly[j]; s.tz = lz[k];
                        out.shapes.push_back(s);
                    } else sid = it->second;
                    int coarse = (int)(i & 1) + 2 * (int)(j & 1) + 4 * (int)(k & 1);
                    by_colour[coarse].push_back({(int)r, sx[i], sy[j], sz[k], sid, coarse});
                }
    }
    // fine colour classes and proposal chunks per shape
    for (TileShape &s : out.shapes) {
        int C = out.lat.colours();
        std::vector<std::vector<int>> cls(C);
        for (int x = 0; x < s.tx; x++)
            for (int y = 0; y < s.ty; y++)
                for (int z = 0; z < s.tz; z++)
                    cls[out.lat.colour(x, y, z)].push_back((x * s.ty + y) * s.tz + z);
        s.class_start.assign(1, 0);
        s.class_sites.clear();
        s.max_class = 0;
        for (int c = 0; c < C; c++) {
            if (cls[c].empty()) continue;
            s.class_sites.insert(s.class_sites.end(), cls[c].begin(), cls[c].end());
            s.class_start.push_back((int)s.class_sites.size());
            s.max_class = std::max(s.max_class, (int)cls[c].size());
        }
        s.ncol = (int)s.class_start.size() - 1;
        if (s.max_class > chunk) { out.error = "fine colour class larger than a workgroup"; return false; }
        s.chunk_start.clear();
        int c0 = 0;
        while (c0 < s.ncol) {
            s.chunk_start.push_back(c0);
            int c1 = c0;
            while (c1 < s.ncol && s.class_start[c1 + 1] - s.class_start[c0] <= chunk) c1++;
            c0 = c1;
        }
        s.chunk_start.push_back(s.ncol);
        int reg = (s.tx + 2 * rc) * (s.ty + 2 * rc) * (s.tz + 2 * hz);
        out.max_region = std::max(out.max_region, reg);
    }
    out.coarse_start.assign(9, 0);
    for (int c = 0; c < 8; c++) {
        out.coarse_start[c] = (int)out.tiles.size();
        out.tiles.insert(out.tiles.end(), by_colour[c].begin(), by_colour[c].end());
    }
    out.coarse_start[8] = (int)out.tiles.size();
    return true;
}

void sweep_order(const Schedule &s, const ReplicaGeom &g, int replica, std::vector<int32_t> &order,
                 std::vector<int32_t> &phase_start)
{
    order.clear(); phase_start.clear();
    int maxcol = 0;
    for (const TileShape &sh : s.shapes) maxcol = std::max(maxcol, sh.ncol);
    for (int cc = 0; cc < 8; cc++) {
        for (int c = 0; c < maxcol; c++) {
            size_t before = order.size();
            for (int t = s.coarse_start[cc]; t < s.coarse_start[cc + 1]; t++) {
                const Tile &tl = s.tiles[t];
                if (tl.replica != replica) continue;
                const TileShape &sh = s.shapes[tl.shape];
                if (c >= sh.ncol) continue;
                for (int k = sh.class_start[c]; k < sh.class_start[c + 1]; k++) {
                    int li = sh.class_sites[k];
                    int lz = li % sh.tz, ly = (li / sh.tz) % sh.ty, lx = li / (sh.tz * sh.ty);
                    int x = tl.x0 + lx, y = tl.y0 + ly, z = tl.z0 + lz;
                    order.push_back((int32_t)(((int64_t)x * g.Y + y) * g.Z + z));
                }
            }
            if (order.size() > before) phase_start.push_back((int32_t)before);
        }
    }
    phase_start.push_back((int32_t)order.size());
}

bool periodic_sublattice(const std::vector<ReplicaGeom> &reps, int rc, int max_index, Sublattice &out)
{
    if (reps.empty()) return false;
    const bool two_d = reps[0].Z == 1;
    auto holds_boxes = [&](const Sublattice &s) {
        const int c0 = s.colour(0, 0, 0);
        for (const ReplicaGeom &g : reps)
            if (s.colour(g.X, 0, 0) != c0 || s.colour(0, g.Y, 0) != c0 || s.colour(0, 0, g.Z) != c0) return false;
        return true;
    };
    for (int idx = 1; idx <= max_index; idx++)
        for (int a = 1; a <= idx; a++) {
            if (idx % a) continue;
            for (int d = 1; d <= idx / a; d++) {
                if ((idx / a) % d) continue;
                const int f = idx / (a * d);
                if (two_d && f != 1) continue;
                for (int b = 0; b < a; b++)
                    for (int c = 0; c < (two_d ? 1 : a); c++)
                        for (int e = 0; e < (two_d ? 1 : d); e++) {
                            Sublattice s{a, b, c, d, e, f};
                            if (!holds_boxes(s)) continue;
                            const int mn = two_d ? min_norm2_plane(s) : s.min_norm2();
                            if (mn > rc * rc) { out = s; return true; }
                        }
            }
        }
    return false;
}

void torus_classes(const Sublattice &lat, const ReplicaGeom &g, std::vector<int32_t> &sites,
                   std::vector<int32_t> &start)
{
    const int nc = lat.colours();
    std::vector<std::vector<int32_t>> by(nc);
    for (int x = 0; x < g.X; x++)
        for (int y = 0; y < g.Y; y++)
            for (int z = 0; z < g.Z; z++) by[lat.colour(x, y, z)].push_back((x * g.Y + y) * g.Z + z);
    sites.clear();
    start.assign(1, 0);
    for (const auto &v : by) {
        sites.insert(sites.end(), v.begin(), v.end());
        start.push_back((int32_t)sites.size());
    }
}

// neighbour sites of every site (periodic wrap), as linear indices
static std::vector<std::vector<int32_t>> torus_neighbours(const ReplicaGeom &g, int rc)
{
    const std::vector<Offset> offs = neighbour_offsets(rc, g.Z == 1);
    const int N = g.X * g.Y * g.Z;
    std::vector<std::vector<int32_t>> nb(N);
    for (int x = 0; x < g.X; x++)
        for (int y = 0; y < g.Y; y++)
            for (int z = 0; z < g.Z; z++) {
                std::vector<int32_t> &v = nb[(x * g.Y + y) * g.Z + z];
                for (const Offset &o : offs) {
                    const int a = ((x + o.dx) % g.X + g.X) % g.X, b = ((y + o.dy) % g.Y + g.Y) % g.Y;
                    const int c = ((z + o.dz) % g.Z + g.Z) % g.Z;
                    v.push_back((a * g.Y + b) * g.Z + c);
                }
            }
    return nb;
}

int greedy_torus_classes(const ReplicaGeom &g, int rc, int max_class, std::vector<int32_t> &sites,
                         std::vector<int32_t> &start)
{
    const int N = g.X * g.Y * g.Z;
    const std::vector<std::vector<int32_t>> nb = torus_neighbours(g, rc);
    std::vector<int> col(N), size;
    std::vector<int> mark;   // mark[c] == s + 1: colour c is taken next to site s
    auto greedy = [&](const std::vector<int32_t> &order) {
        std::fill(col.begin(), col.end(), -1);
        size.clear();
        for (int32_t s : order) {
            for (int32_t t : nb[s])
                if (col[t] >= 0) mark[col[t]] = s + 1;
            int c = 0;
            while (c < (int)size.size() && (mark[c] == s + 1 || size[c] >= max_class)) c++;
            if (c == (int)size.size()) size.push_back(0);
            col[s] = c;
            size[c]++;
        }
    };
    mark.assign(N + 1, 0);
    // first pass: a sheared raster, (x + 4 y) mod X, then index
    std::vector<int32_t> order(N);
    for (int i = 0; i < N; i++) order[i] = i;
    auto key = [&](int32_t i) { const int x = i / (g.Y * g.Z), y = (i / g.Z) % g.Y; return (x + 4 * y) % g.X; };
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key(a) < key(b); });
    greedy(order);
    std::vector<int> best_col = col;
    int best = (int)size.size();
    uint32_t rng = 0x9E3779B9u;
    for (int it = 0; it < 60; it++) {
        // the classes of this pass, reordered: largest first, reversed, or shuffled
        const int nc = (int)size.size();
        std::vector<int> keys(nc);
        for (int k = 0; k < nc; k++) keys[k] = k;
        if (it % 3 == 0) std::stable_sort(keys.begin(), keys.end(), [&](int a, int b) { return size[a] > size[b]; });
        else if (it % 3 == 1) std::reverse(keys.begin(), keys.end());
        else
            for (int k = nc - 1; k > 0; k--) {
                rng = rng * 1664525u + 1013904223u;
                std::swap(keys[k], keys[(int)((rng >> 8) % (uint32_t)(k + 1))]);
            }
        std::vector<std::vector<int32_t>> cls(nc);
        for (int i = 0; i < N; i++) cls[col[i]].push_back(i);
        order.clear();
        for (int k : keys) order.insert(order.end(), cls[k].begin(), cls[k].end());
        greedy(order);
        if ((int)size.size() < best) { best = (int)size.size(); best_col = col; }
    }
    std::vector<std::vector<int32_t>> by(best);
    for (int i = 0; i < N; i++) by[best_col[i]].push_back(i);
    sites.clear();
    start.assign(1, 0);
    for (const auto &v : by) {
        sites.insert(sites.end(), v.begin(), v.end());
        start.push_back((int32_t)sites.size());
    }
    return best;
}

bool valid_torus_classes(const ReplicaGeom &g, int rc, const std::vector<int32_t> &sites,
                         const std::vector<int32_t> &start)
{
    const int N = g.X * g.Y * g.Z;
    if ((int)sites.size() != N || start.empty() || start.front() != 0 || start.back() != N) return false;
    std::vector<int> col(N, -1);
    for (size_t k = 0; k + 1 < start.size(); k++) {
        if (start[k + 1] < start[k]) return false;
        for (int32_t j = start[k]; j < start[k + 1]; j++) {
            const int32_t s = sites[j];
            if (s < 0 || s >= N || col[s] >= 0) return false;
            col[s] = (int)k;
        }
    }
    const std::vector<std::vector<int32_t>> nb = torus_neighbours(g, rc);
    for (int s = 0; s < N; s++)
        for (int32_t t : nb[s])
            if (t != s && col[t] == col[s]) return false;
    return true;
}

}  // namespace sn

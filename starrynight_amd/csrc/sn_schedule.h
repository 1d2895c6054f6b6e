// sn_schedule.h -- host-side colour schedule for the tile sweep (pure C++).
//
// A sweep visits every site once.  Tiles (blocks of sites, edge >= rc) get one
// of up to 8 coarse colours by the parity of their tile coordinates; tiles of
// one coarse colour are separated by a full tile (>= rc sites) along some axis,
// so they never interact and run concurrently, one workgroup each.  Inside a
// tile the sites are split into fine colours by a sublattice whose shortest
// vector is longer than rc; one fine colour is updated at a time with
// workgroup barriers in between.  The neighbour table follows gen_neighbour
// (src/starrynight-montecarlo-core.c:38-72): every r != 0 with |r| <= rc,
// dz == 0 when Z == 1.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace sn {

struct Offset {
    int dx, dy, dz;
    int r2;
};

// Neighbour offsets in gen_neighbour order (dx, dy, dz ascending).
std::vector<Offset> neighbour_offsets(int rc, bool two_d);
// One representative r of every +-r pair, nearest neighbours (|r| = 1) first.
std::vector<Offset> neighbour_pairs(int rc, bool two_d);

// Sublattice used for the fine colouring, as a lower-triangular basis
// v1 = (a,0,0), v2 = (b,d,0), v3 = (c,e,f); colours = a*d*f.
struct Sublattice {
    int a, b, c, d, e, f;
    int colours() const { return a * d * f; }
    int colour(int x, int y, int z) const;
    int min_norm2() const;
    std::string describe() const;
};
// kind: 1 cubic (rc+1)^3, 2 minimum index with shortest vector^2 > rc^2,
// 3 body-centred (4,0,0),(0,4,0),(2,2,2) (valid for rc <= 3).
Sublattice make_sublattice(int rc, bool two_d, int kind);

// Split an axis of length L into an even number of tiles, each of size >= rc
// and <= tmax, near `target`, with tile + 2*rc <= L (so a tile's halo never
// wraps onto itself).  L == 1 (the z axis of a 2D lattice) gives one tile of
// size 1.  Returns false when no such split exists.
bool split_axis(int L, int rc, int target, int tmax, std::vector<int> &start, std::vector<int> &size);

struct TileShape {
    int tx, ty, tz;             // tile size
    int ncol;                   // fine colours
    int max_class;              // largest fine class
    std::vector<int> class_start;  // ncol + 1
    std::vector<int> class_sites;  // tile-local linear index (lx*ty + ly)*tz + lz, by colour
    std::vector<int> chunk_start;  // colour index where each proposal chunk starts (+ ncol)
};

struct Tile {
    int replica;
    int x0, y0, z0;
    int shape;
    int coarse;   // 0..7
};

struct ReplicaGeom {
    int X, Y, Z;
};

struct Schedule {
    int rc = 0;
    bool two_d = false;
    Sublattice lat{};
    std::vector<TileShape> shapes;
    std::vector<Tile> tiles;                       // grouped by coarse colour
    std::vector<int> coarse_start;                 // 9 entries
    int max_region = 0;                            // largest (tx+2h)(ty+2h)(tz+2hz)
    std::string error;
};

// Build the schedule for all replicas.  target[3] = tile edge aimed for per
// axis; chunk = block size (sites whose proposals are drawn at once);
// max_region = largest tile + halo (sites) the LDS budget allows;
// max_tile = largest tile (sites) a workgroup handles.
bool build_schedule(const std::vector<ReplicaGeom> &reps, int rc, const int target[3],
                    int colouring_kind, int chunk, int max_region, int max_tile, Schedule &out);

// Visit order of replica r for one sweep and its phase boundaries
// (phase = one (coarse colour, fine colour) pair across all tiles).
void sweep_order(const Schedule &s, const ReplicaGeom &g, int replica, std::vector<int32_t> &order,
                 std::vector<int32_t> &phase_start);

// Colouring of whole periodic lattices (no tiles): a sublattice that holds
// every replica's box vectors (X,0,0), (0,Y,0), (0,0,Z) and whose shortest
// vector is longer than rc, so that sites of one colour are more than rc apart
// through the periodic boundary as well.  The smallest index up to max_index;
// false if there is none.
bool periodic_sublattice(const std::vector<ReplicaGeom> &reps, int rc, int max_index, Sublattice &out);

// The sites of one replica by colour of `lat` (linear index (x*Y+y)*Z+z,
// ascending within a colour) and the colour starts (colours + 1 entries).
void torus_classes(const Sublattice &lat, const ReplicaGeom &g, std::vector<int32_t> &sites,
                   std::vector<int32_t> &start);

// Greedy colouring of one periodic lattice, for sizes whose sublattice
// colourings need many colours (25 x 25 admits only index 25, 125 or 625):
// sites within rc of each other through the periodic boundary get different
// colours, and no colour holds more than max_class sites.  Iterated greedy
// (J. Culberson, "Iterated greedy graph coloring", 1992): each pass colours
// the sites class by class in a reordering of the previous pass's classes,
// which never needs more colours.  sites / start as torus_classes; returns the
// colour count.
int greedy_torus_classes(const ReplicaGeom &g, int rc, int max_class, std::vector<int32_t> &sites,
                         std::vector<int32_t> &start);

// True if no two sites of one colour are within rc (periodic wrap) and every
// site appears exactly once.
bool valid_torus_classes(const ReplicaGeom &g, int rc, const std::vector<int32_t> &sites,
                         const std::vector<int32_t> &start);

}  // namespace sn

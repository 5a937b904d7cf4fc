// tree.hpp — periodic unwrap, octree construction and node records.
#pragma once

#include "common.hpp"
#include "pm.hpp"

#include <vector>

namespace chui {

// Preorder node records. Node i's subtree is the index range [i, i + link.skip),
// so a stackless traversal moves to i + 1 to open a node and to i + skip to
// leave it; visiting nodes in increasing index reproduces the reference's DFS
// order (children pushed in reverse octant order, tree_pm_coupling.cpp:916).
struct alignas(16) NodeLink {
  uint32_t begin;        // first tree-order particle
  uint32_t count;        // particle count
  uint32_t skip;         // subtree size in nodes (1 for leaves)
  uint32_t child_count;  // 0 = leaf
};
struct alignas(32) NodeGeo64 { double cx, cy, cz, half; };
struct alignas(32) NodeMom64 { double comx, comy, comz, mass; };
// Traceless quadrupole, raw second-moment trace and |centre - COM|.
struct alignas(64) NodeQuad64 { double qxx, qxy, qxz, qyy, qyz, qzz, trace, offset; };

// Mixed-precision records: 32-bit fixed-point periodic coordinates
// (x -> x * 2^32 / L, wrapping modulo 2^32) so that minimum-image deltas are an
// exact int32 subtraction.
struct alignas(16) NodeGeo32 { uint32_t cx, cy, cz; float half; };
struct alignas(16) NodeMom32 { uint32_t comx, comy, comz; float mass; };
struct alignas(32) NodeQuad32 { float qxx, qxy, qxz, qyy, qyz, qzz, trace, offset; };
// Tight bounding box of a node's particles for the mixed walk's cutoff prune:
// centre in fixed point, half-extents in separation units (fixed-point units
// for a cubic box, lengths otherwise) rounded up to cover the fixed-point
// rounding of the centre and of every particle; plus the octree cell's half
// size in length units.
struct alignas(32) NodeBox32 { uint32_t cx, cy, cz; float hx, hy, hz, cell_half; uint32_t pad; };

struct UnwrapResult {
  double anchor[3];
  double umin[3], umax[3];
  int canonical[3];
};

// Bits of the sticky fault word (the walk counters' fault slot, copied to the
// host after every solve). Walks set kFaultWalk; a capture-safe build sets the
// others where the ordinary build throws, since it reads nothing back.
constexpr uint64_t kFaultWalk = 1;         // a walk lost progress (logic error)
constexpr uint64_t kFaultCoordinates = 2;  // non-finite source coordinates (invalid argument)
constexpr uint64_t kFaultMasses = 4;       // non-finite or negative masses (invalid argument)
constexpr uint64_t kFaultPrepared = 8;     // node capacity, depth or mass uniformity beyond the prepared build
constexpr uint64_t kFaultTooDeep = 16;     // octree deeper than 21 levels (unsupported)
constexpr uint64_t kFaultActive = 32;      // active index out of range (invalid argument)

struct TreeParams {
  std::size_t n = 0;
  double L[3] = {0, 0, 0};
  uint32_t leaf_size = 16;
  bool quadrupole = true;
  bool mixed = false;  // emit the fixed-point records; the FP64 records follow on demand
  // Capture-safe build: no read-back and no allocation once prepared; every
  // per-build value stays on the device and data-dependent failures set bits
  // of *fault (mixed precision only).
  bool capture_safe = false;
  unsigned long long* fault = nullptr;
  // An ordinary build that sizes every buffer as a capture-safe build of the
  // same sources will (node arrays by the capacity, unwrap buffers by their
  // bounds), so that the capture-safe builds after it allocate nothing.
  bool size_for_capture = false;
};

// Device state of one built tree. Buffers are grow-only and owned by Tree.
struct Tree {
  std::size_t n = 0;
  std::size_t node_count = 0;
  std::size_t leaf_count = 0;
  uint32_t max_depth = 0;
  int level_hint = 0;  // levels the next build issues before its first read-back (0: all)
  // Capture-safe builds: the node capacity (bfs_capacity), the source count,
  // the mass uniformity and the levels issued (safe_levels, the prepared
  // depth + 4) come from the last ordinary build; node_count is then the
  // capacity and the real count sits on the device (counts_on_device).
  int safe_levels = 0;
  std::size_t prepared_n = 0;
  bool counts_on_device = false;
  bool sized_for_capture = false;  // the last ordinary build sized its buffers for capture-safe builds
  DeviceBuffer frame;  // Frame of the last capture-safe build
  UnwrapResult unwrap{};
  bool mass_uniform = false;
  double uniform_mass = 0.0;
  double root_centre[3] = {0, 0, 0};
  double root_half = 0.0;
  // particle level
  DeviceBuffer order;      // uint32[n]: tree position -> source row
  DeviceBuffer pk;         // double4[n]: tree-order canonical coordinates and mass
  DeviceBuffer pk_in;      // double4[n]: the same in input order (build scratch)
  DeviceBuffer pos_of_row; // uint32[n]: source row -> tree position (built on demand)
  // preorder node level
  DeviceBuffer link, geo64, mom64, quad64, child_index;  // child_index: uint32[8*nodes]
  DeviceBuffer geo32, mom32, quad32, box32;
  bool records64 = false;  // geo64, mom64, quad64 and child_index are current
  bool cubic = false;      // equal axes: the mixed records' separation unit is fixed point
  double fscale[3] = {0, 0, 0};  // 2^32 / L per axis
  // construction scratch
  DeviceBuffer keys, keys_alt, vals_alt, sort_scratch;
  DeviceBuffer bucket_min, bucket_max, block_best, block_max, occ_bytes;
  DeviceBuffer bfs_begin, bfs_count, bfs_first_child, bfs_nchild, bfs_meta, bfs_geo, bfs_subtree, bfs_pre,
      bfs_mom, bfs_quad, bfs_box, bfs_scan, level_scratch, level_bounds, level_dev, flags, bounds;
  DeviceBuffer dep_flag, dep_index, dep_units, dep_blocks, dep_scan, dep_total;
  std::size_t bfs_capacity = 0;
  std::vector<std::size_t> level_start;  // BFS level ranges (host copy)
  PinnedReadback readback;               // counts and flags read back during a build
};

// Bounds/canonical check, periodic unwrap on all three axes, Morton/path keys,
// stable sort, level-synchronous octree, preorder numbering, in-leaf Morton
// order and multipoles. Inputs are device pointers (caller positions/masses).
void build_tree(Tree& t, const double* x, const double* y, const double* z, const double* m,
                const TreeParams& p, double softening, hipStream_t s);

// The node count of the last build on the device (capture-safe builds), or
// nullptr after an ordinary build.
const uint32_t* device_node_count(const Tree& t);

// FP64 preorder records (geo64, mom64, quad64, child_index) of the last built
// tree. A mixed-precision build leaves them out (its walks read the
// fixed-point records); they are written from the BFS arrays on first need.
void ensure_records64(Tree& t, hipStream_t s);

// Mesh-deposit units of the last built tree: the shallowest subtrees whose
// particles span at most d.max_cells cells (or leaves that never do). Reads
// back the unit and block counts, or after a capture-safe build leaves them on
// the device (DepositUnits::dev_counts) and sizes the launch by a bound.
DepositUnits build_deposit_units(Tree& t, const DepositTiling& d, hipStream_t s);

// Morton keys + stable order of given (unwrapped) device coordinates.
void morton_order_device(Tree& t, const double* x, const double* y, const double* z, std::size_t n,
                         uint64_t* keys_sorted, uint32_t* order, hipStream_t s);
// Single-axis periodic unwrap (parity entry point).
double unwrap_axis_device(Tree& t, const double* in, std::size_t n, double L, double* out, hipStream_t s);

// row -> tree position map for active-subset targets.
void build_pos_of_row(Tree& t, hipStream_t s);

}  // namespace chui

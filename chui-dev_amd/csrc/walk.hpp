// walk.hpp — periodic short-range residual walk (per-target semantics of
// tree_pm_coupling.cpp:4026-4272) with the PM interpolation fused in front.
#pragma once

#include <vector>

#include "common.hpp"
#include "tree.hpp"

namespace chui {

constexpr int kHTableCells = 128;  // cubic cells of the Gaussian coefficient H(t)

struct WalkParams {
  double L[3];
  double G;
  double a;         // split scale r_s
  double rcut, rcut2;
  double theta;
  double eps;
  int criterion;    // chui_opening_criterion
  int quad;         // multipole order 2
  uint32_t node_count;
  uint32_t n_targets;
  int target_per_wave;   // mixed strict walk: one target per wave (walkmix_target_kernel)
  int walk_waves;        // waves per block of the timed mixed walk (2, 4 or 8; the autotuner's knob)
  int count;             // accumulate walk counters (diagnostic launches)
  int identity_targets;  // targets are tree positions 0..n-1, output slot = order[k]
  float uniform_mass_f;  // mixed: mass when uniform
  double uniform_mass;   // used when the mass lane is null
  // PM interpolation (fused)
  int pm_enabled;
  int assignment;   // 0 CIC, 1 TSC
  int mesh[3];
  int mesh_log2[3];
  std::size_t mesh_row;   // padded real row stride (elements)
  // Mixed precision: fixed-point scale, cutoff radii in separation units
  // (fixed-point units when the box is cubic, lengths otherwise) and the
  // cubic-cell table of H(t) = L(sqrt t)/t^{3/2}, t = r^2/(4 a^2).
  int cubic;
  float fx_to_len[3];  // L / 2^32
  float rcut2_sep, prune2_sep, contain2_sep;
  float c4, c8;        // 1/(4 a^2), 1/(8 a^3)
  float t_to_cell;     // r^2 -> table cell coordinate
  // Largest half-extent (fixed-point units, per axis) of a wave's targets for
  // which the walk uses float offsets from the wave origin:
  // min(2^28, 2^31 - r_cut - 2^24) in units of L_a / 2^32.
  uint32_t fast_half_extent[3];
  // Mixed walk in table units (lengths times sqrt(t_to_cell), so r^2 is the
  // H-table coordinate): per-axis scale of fixed-point differences, cutoff^2,
  // softening^2, the H coefficient c8 / t_to_cell^1.5, the force scale
  // G * t_to_cell, and the node half-size from which no strict acceptance is
  // possible (0.041 r_cut).
  float tab_scale[3];
  float rcut2_tab, eps2_tab, c8_tab, force_scale;
  float tab_to_len2;   // table-unit r^2 -> length^2 (nearest-pair lane)
  float accept_half_max;
  // Box-classified mixed walk: largest half-extent of a wave's target box
  // (separation units, per axis) for which the wave walks the tree once for
  // its box instead of lane by lane.
  float box_half_max[3];
  // FP64 box-classified walk (solves without counters or nearest-pair lane):
  // degree-5 cells of H on [0, t_max] (h64_cells of them, t = r^2 / 4a^2),
  // r^2 -> cell coordinate, eps^2 and 1/(8 a^3).
  // Mixed timed walk with the residual table indexed by the float bits of
  // s = r^2 + eps^2 (log_cells > 0): cells (the last one, at s = 1, is zero)
  // and the (exponent, top-5-mantissa) field of the first cell.
  uint32_t log_cells, log_base;
  int box64;
  int h64_cells;
  double h64_scale;
  double eps2_64, c8_64;
  // Adaptive acceptance (TreePmAcceptancePolicy::kAdaptiveRelative; FP64 walk):
  // estimator, com-distance cutoff policy, Gaussian-table leaf pairs, radial-
  // coefficient node expansion, maximum opening angle, alpha, acceleration
  // floor, and 1/(2a)^2, 1/(8a^3) of the table lookup.
  int estimator;
  int com_cut;
  int lookup;
  int radial;
  double amax, alpha, accel_floor;
  double inv2a2, inv8a3;
};

// Group walk parameters (tree_pm_group_walk.hpp:78-341): residual block and
// target group sizes, largest group extent, alpha / G, the squared maximum
// opening angle, the cutoff containment rule and the caller's source
// coordinate bounds for the bounded-metric check.
struct GroupParams {
  uint32_t block;        // residual_block_size (visit positions per block)
  uint32_t group_size;   // 0: the active targets of one leaf
  double max_extent;
  double tolg;
  double amax2;
  int contain;           // kContainNode
  double lo[3], hi[3];
};

// Group walk counters on the device (chui_treepm_walk_counters order after
// its header): visited, pruned, pairs, multipoles, leaves, groups, grouped
// targets, largest group, fallback targets, fallback history, fallback
// geometry, candidate leaves, culled leaves, candidate particles, largest list.
constexpr int kGroupCounters = 15;

// Totals and PM lanes are per active slot; the short-range lanes and the
// nearest-pair lane are indexed by source row (TreePmForceAccumulatorView,
// tree_pm_coupling.hpp:127-165, :1957-1976).
struct WalkOutputs {
  void* ax;
  void* ay;
  void* az;
  void* pm_x;   // nullable
  void* pm_y;
  void* pm_z;
  void* sr_x;   // nullable, source-row indexed
  void* sr_y;
  void* sr_z;
  void* near2 = nullptr;  // nullable, source-row indexed: r^2 of the nearest direct pair in r_cut
};

// Counters accumulated on the device: visited, pruned, pairs, multipoles,
// leaves (per target), wave node visits, wave leaf-source slots and wave
// pair-kernel iterations (per wave).
constexpr int kWalkCounters = 8;
// Counter words after the per-solve counters: the sticky traversal-fault word,
// then the FP64 per-lane walk's checksum of its direct-pair set (a wrapping
// sum of a 64-bit mix of (target row, source row) over every pair inside
// r_cut, so it does not depend on the summation order).
constexpr int kFaultSlot = kWalkCounters;
constexpr int kPairChecksumSlot = kWalkCounters + 1;
constexpr int kWalkCounterWords = kWalkCounters + 2;

// Tree-ordered sources: FP64 (x, y, z, m); mixed (qx, qy, qz, float bits of m).
// p.box64 selects the box-classified FP64 walk (h64tab: its table, device).
void walk_fp64(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass,
               const uint32_t* target_pos, const uint32_t* target_slot, const double* fx, const double* fy,
               const double* fz, const WalkOutputs& out, unsigned long long* counters, const double2* h64tab,
               hipStream_t s);

// Device buffers of the force VJP's accepted-node terms (multipole order 2):
// U (3) and V (7) per preorder node, the accepted (target << 32 | node) list
// with its count and capacity, sort scratch for it, and per-source
// corrections (3 per source, tree order).
struct VjpNodeTerms {
  double* node_acc = nullptr;
  unsigned long long* accepted = nullptr;
  unsigned long long* accepted_alt = nullptr;
  uint32_t* vals = nullptr;
  uint32_t* vals_alt = nullptr;
  void* sort_scratch = nullptr;
  unsigned int* count = nullptr;
  unsigned int cap = 0;
  double* corr = nullptr;
};

// Transpose of the total-force Jacobian applied to lam (FP64, all targets, tree
// order identity). src and lam are tree-ordered; fx..fz the forward force
// meshes and nu the adjoint potential on the same mesh. p.box64 selects the
// box-classified walk with the polynomial table (h64tab). With p.quad the
// walk applies the forward's strict acceptance: an accepted node's terms for
// the target are added in the walk, and nt (required then) collects the rest.
void walk_fp64_vjp(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass, const double4* lam,
                   const double* fx, const double* fy, const double* fz, const double* nu, const WalkOutputs& out,
                   unsigned long long* counters, const double2* h64tab, const VjpNodeTerms* nt, hipStream_t s);
// The accepted-node terms after walk_fp64_vjp (count accepted pairs, read back
// by the caller): per accepted (t, N), the pairs between t and N's particles
// that the symmetric pair terms counted on one side only; U and V spread over
// N's particles; the corrections added to out (caller order).
// With count_dev (a recorded VJP) `count` is the list's capacity: the list is
// sorted over it (unused slots hold ~0 and sort last), the corrections read the
// real count on the device, and a count beyond the capacity sets
// kFaultPrepared in *fault.
void vjp_node_corrections(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass,
                          double uniform_mass_value, const double4* lam, const VjpNodeTerms& nt, unsigned int count,
                          const WalkOutputs& out, hipStream_t s, const unsigned int* count_dev = nullptr,
                          unsigned long long* fault = nullptr);
// Per-target FP64 walk under the adaptive acceptance. gtab = Gaussian Hermite
// knots (value, slope) x 4097 (required when p.lookup), prev = previous
// acceleration magnitude per output slot (nullable).
void walk_fp64_adaptive(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass,
                        const double2* gtab, const double* prev, const uint32_t* target_pos,
                        const uint32_t* target_slot, const double* fx, const double* fy, const double* fz,
                        const WalkOutputs& out, unsigned long long* counters, hipStream_t s);

// FP64 group walk over the residual blocks of the visit-ordered targets.
// leaf_of_pos = leaf node per tree position. Targets of groups that fall back
// are appended to (fallback_pos, fallback_slot) with their count in
// *fallback_count (device) for walk_fp64_adaptive.
void walk_fp64_group(const Tree& t, const WalkParams& p, const GroupParams& g, const double4* src,
                     int uniform_mass, const double2* gtab, const double4* gcoef, const double* prev,
                     const uint32_t* leaf_of_pos,
                     const uint32_t* target_pos, const uint32_t* target_slot, const double* fx, const double* fy,
                     const double* fz, const WalkOutputs& out, uint32_t* fallback_pos, uint32_t* fallback_slot,
                     uint32_t* fallback_count, unsigned long long* counters, hipStream_t s);

// Host FP64 Gaussian Hermite knots of tree_pm_split_kernel.hpp:130-183:
// 4097 (H(t_i), step * dH/dt(t_i)) pairs, t_i = i / 64.
std::vector<double> gaussian_hermite_knots();
// The same 4096 cubic cells as monomials in the cell coordinate u:
// (a0, a1, a2, a3) per cell (the group walk's pair kernel).
std::vector<double> gaussian_hermite_monomials(const std::vector<double>& knots);

// Leaf node of every tree position of the last built tree.
// Mixed-precision group walk: the FP64 group walk's groups, lists and
// decisions (src64, FP64 records) with FP32 listed-source pairs (src32 in
// fixed point, the Gaussian coefficient table htab in H-table units) and FP32
// PM meshes and outputs. Fallback targets as in walk_fp64_group.
void walk_mixed_group(const Tree& t, const WalkParams& p, const GroupParams& g, const double4* src64,
                      const uint4* src32, int uniform_mass, const float4* htab, const double* prev,
                      const uint32_t* leaf_of_pos, const uint32_t* target_pos, const uint32_t* target_slot,
                      const float* fx, const float* fy, const float* fz, const WalkOutputs& out,
                      uint32_t* fallback_pos, uint32_t* fallback_slot, uint32_t* fallback_count,
                      unsigned long long* counters, hipStream_t s);

// Mixed adaptive solves: FP32 outputs of n targets whose short-range sums an
// FP64 per-target walk wrote into sx/sy/sz (by output slot) and near64 (by
// source row, nullable unless out.near2), plus the FP32 PM interpolation.
// identity: targets are tree positions 0..n-1 with slot order[k].
void adaptive_mixed_finish(const Tree& t, const WalkParams& p, const uint4* src32, const uint32_t* pos,
                           const uint32_t* slot, uint32_t n, int identity, const double* sx, const double* sy,
                           const double* sz, const double* near64, const float* fx, const float* fy,
                           const float* fz, const WalkOutputs& out, hipStream_t s);

void build_leaf_of_pos(const Tree& t, uint32_t* leaf_of_pos, hipStream_t s);

// ltab: the bit-indexed residual table of the timed walk (used when p.log_cells > 0).
void walk_mixed(const Tree& t, const WalkParams& p, const uint4* src, int uniform_mass, const float4* htab,
                const float4* ltab, const uint32_t* target_pos, const uint32_t* target_slot, const float* fx,
                const float* fy, const float* fz, const WalkOutputs& out, unsigned long long* counters,
                hipStream_t s);

// Host FP64 table of H(t) on [0, t_max] (kHTableCells cubic cells, 4 floats
// each); sets p.t_to_cell.
std::vector<float> gaussian_coefficient_table(WalkParams& p, double t_max);
// Host table of the FP64 box walk: H(t) on [0, t_max] in `cells` cells, per
// cell the monomial coefficients a0..a5 of the degree-5 polynomial in the cell
// coordinate u in [0, 1) through the 6 Chebyshev points of the cell.
std::vector<double> gaussian_quintic_table(double t_max, int cells);
// Host table of the mixed timed walk: the residual coefficient as a function
// of s = (r^2 + eps^2) * scale on 32 cubic cells per octave from the octave of
// eps^2 * scale up to s = 1, plus one zero cell; 4 floats per cell (the
// polynomial in the integer position of s in its cell). Empty when eps is 0
// or more than kLogMaxOctaves octaves would be needed.
constexpr int kLogMaxOctaves = 24;
constexpr int kLogCellBits = 5;  // 32 cells per octave
constexpr int kLogCellsPerOctave = 1 << kLogCellBits;
constexpr int kLogMantissaShift = 23 - kLogCellBits;
std::vector<float> log_residual_table(double scale, double eps, double a, int32_t& first_octave, uint32_t& cells);
constexpr int kH64MaxCells = 256;

}  // namespace chui

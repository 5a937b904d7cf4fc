// walk.hpp — periodic short-range residual walk (per-target semantics of
// tree_pm_coupling.cpp:4026-4272) with the PM interpolation fused in front.
#pragma once

#include "common.hpp"
#include "tree.hpp"

namespace chui {

constexpr int kMaxPolyDegree = 24;

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
  int identity_targets;  // targets are tree positions 0..n-1, output slot = order[k]
  float uniform_mass_f;  // mixed: mass when uniform
  double uniform_mass;   // used when the mass lane is null
  // PM interpolation (fused)
  int pm_enabled;
  int assignment;   // 0 CIC, 1 TSC
  int mesh[3];
  int mesh_log2[3];
  std::size_t mesh_row;   // padded real row stride (elements)
  // Mixed precision: fixed-point scale and the H(t) polynomial (Horner in
  // s = (t - t_mid) * t_inv_half, ascending coefficients).
  float fx_to_len[3];  // L / 2^32
  float c4, c8;        // 1/(4 a^2), 1/(8 a^3)
  float t_mid, t_inv_half;
  int poly_degree;
  float poly[kMaxPolyDegree + 1];
};

struct WalkOutputs {
  void* ax;
  void* ay;
  void* az;
  void* pm_x;   // nullable
  void* pm_y;
  void* pm_z;
  void* sr_x;   // nullable
  void* sr_y;
  void* sr_z;
};

// Counters accumulated on the device: visited, pruned, pairs, multipoles, leaves.
constexpr int kWalkCounters = 5;

void walk_fp64(const Tree& t, const WalkParams& p, const double* sx, const double* sy, const double* sz,
               const double* sm, const uint32_t* target_pos, const uint32_t* target_slot, const double* fx,
               const double* fy, const double* fz, const WalkOutputs& out, unsigned long long* counters,
               hipStream_t s);

void walk_mixed(const Tree& t, const WalkParams& p, const uint32_t* qx, const uint32_t* qy, const uint32_t* qz,
                const float* sm, const uint32_t* target_pos, const uint32_t* target_slot, const float* fx,
                const float* fy, const float* fz, const WalkOutputs& out, unsigned long long* counters,
                hipStream_t s);

// Host FP64 fit of H(t) = L(sqrt t)/t^{3/2} on [0, t_max]; fills p.poly/degree.
void fit_gaussian_coefficient(WalkParams& p, double t_max);

}  // namespace chui

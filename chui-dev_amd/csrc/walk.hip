// walk.hip — wavefront-cooperative periodic short-range residual walk.
//
// One wave64 owns 64 consecutive tree-order targets. The tree is stored in
// preorder with subtree sizes, so each lane's depth-first walk is a monotone
// sequence of node indices: opening node i continues at i + 1, pruning or
// accepting it continues at i + skip(i). The wave always processes the smallest
// pending node index of its lanes (a wave-uniform index, so node records and
// leaf sources come through the scalar cache); lanes whose next node is larger
// sit that node out. Every lane therefore sees exactly the node sequence,
// acceptance decisions and summation order of the reference's per-target DFS
// (tree_pm_coupling.cpp:4026-4272), and leaf pairs are evaluated with the
// source shared by the whole wave.
//
// walk64_kernel  : FP64, the reference's arithmetic (erf/exp split kernel,
//                  nearbyint minimum image, monopole + quadrupole residual).
// walkmix_kernel : 32-bit fixed-point positions (minimum image = int32
//                  subtraction), FP32 pair kernel with a polynomial for the
//                  Gaussian coefficient H(t), per-leaf partial sums folded into
//                  a Kahan-compensated FP32 accumulator.
#include "walk.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace chui {
namespace {

constexpr double kInvSqrtPi = 0.564189583547756286948079451560772586;
constexpr int kWalkThreads = 256;

// ---- FP64 split kernel (tree_pm_split_kernel.hpp:66-228) -------------------
__device__ __forceinline__ double series_h(double t) {
  constexpr double c = 2.25675833419102514779231780624309034;
  return c * (1.0 / 3.0 + t * (-1.0 / 5.0 + t * (1.0 / 14.0 + t * (-1.0 / 54.0 + t * (1.0 / 264.0 +
         t * (-1.0 / 1560.0 + t * (1.0 / 10800.0 - t / 85680.0)))))));
}
__device__ __forceinline__ double short_factor(double r, double a) {
  if (r <= 0.0) return 1.0;
  const double q = (0.5 * r) / a;
  if (isinf(q)) return 0.0;
  return erfc(q) + 2.0 * kInvSqrtPi * q * exp(-q * q);
}
__device__ __forceinline__ double long_factor(double r, double a) {
  if (r <= 0.0) return 0.0;
  const double q = (0.5 * r) / a;
  if (isinf(q)) return 1.0;
  if (q < 0.125) {
    const double t = q * q;
    const double h = 4.0 * kInvSqrtPi * (1.0 / 3.0 + t * (-1.0 / 5.0 + t * (1.0 / 14.0 + t * (-1.0 / 54.0 +
        t * (1.0 / 264.0 + t * (-1.0 / 1560.0 + t * (1.0 / 10800.0 - t / 85680.0)))))));
    return q * q * q * h;
  }
  return erf(q) - 2.0 * kInvSqrtPi * q * exp(-q * q);
}
__device__ __forceinline__ double long_inv_r3(double r, double a) {
  const double q = (0.5 * r) / a;
  if (q < 0.125) return ((0.125 * series_h(q * q) / a) / a) / a;
  return ((long_factor(r, a) / r) / r) / r;
}
__device__ __forceinline__ double softened_inv_r3(double r2, double eps) {
  const double s = r2 + eps * eps;
  if (s <= 0.0) return 0.0;
  return 1.0 / (s * sqrt(s));
}
__device__ __forceinline__ double residual_inv_r3(double r2, double eps, double a) {
  if (r2 <= 0.0) {
    if (eps == 0.0) return softened_inv_r3(r2, eps);
    constexpr double k = 0.094031597257959381158013241926795431;
    if (eps <= a) {
      const double ratio = eps / a;
      return (((1.0 - k * ratio * ratio * ratio) / eps) / eps) / eps;
    }
    const double ratio = a / eps;
    return (((ratio * ratio * ratio - k) / a) / a) / a;
  }
  return softened_inv_r3(r2, eps) - long_inv_r3(sqrt(r2), a);
}

// Minimum image (tree_pm_geometry.hpp:12-25): nearbyint semantics, with the
// division skipped whenever |d| <= L lies outside the half-box tie band.
__device__ __forceinline__ double mi64(double d, double L) {
  const double ad = fabs(d), half = 0.5 * L;
  if (ad <= L && fabs(ad - half) > 4.0 * 2.220446049250313e-16 * L)
    return d > half ? d - L : (d < -half ? d + L : d);
  return d - L * rint(d / L);
}

// tree_pm_geometry.hpp:51-60
__device__ __forceinline__ bool within_cutoff(double d2, double c, double c2) {
  if (!isnormal(c2) || !isfinite(d2) || fabs(d2 - c2) <= 8.0 * 2.220446049250313e-16 * c2) return sqrt(d2) <= c;
  return d2 <= c2;
}

// monopolePlusQuadrupoleAccelPeriodicReference (tree_pm_coupling.cpp:743-827)
__device__ void node_accel64(const NodeMom64& M, const NodeQuad64& Q, double dx, double dy, double dz,
                             const WalkParams& P, double& ax, double& ay, double& az) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double safe_r2 = fmax(r2, 1.0e-30);
  const double r = sqrt(safe_r2);
  const double pe = P.eps;  // uniform softening: max(node eps, target eps) = eps
  const double eps2 = pe * pe;
  const double denom = fmax(r2 + eps2, 1.0e-30);
  const double soft = 1.0 / (denom * sqrt(denom));
  const double split = short_factor(r, P.a);
  const double newton = 1.0 / (safe_r2 * r);
  const double radial = soft + (split - 1.0) * newton;
  const double pref = P.G;
  double cx = pref * M.mass * radial * dx;
  double cy = pref * M.mass * radial * dy;
  double cz = pref * M.mass * radial * dz;
  if (P.quad) {
    const double tr = Q.trace;
    const double mxx = (Q.qxx + tr) / 3.0, mxy = Q.qxy / 3.0, mxz = Q.qxz / 3.0;
    const double myy = (Q.qyy + tr) / 3.0, myz = Q.qyz / 3.0, mzz = (Q.qzz + tr) / 3.0;
    const double mdx = mxx * dx + mxy * dy + mxz * dz;
    const double mdy = mxy * dx + myy * dy + myz * dz;
    const double mdz = mxz * dx + myz * dy + mzz * dz;
    const double dmd = dx * mdx + dy * mdy + dz * mdz;
    const double a = P.a, a2 = a * a;
    const double ex = exp(-r2 / (4.0 * a2));
    const double s1 = -0.5 * kInvSqrtPi * r2 * ex / (a * a2);
    const double s2 = kInvSqrtPi * ex * (-r / (a * a2) + 0.25 * r * r2 / (a * a2 * a2));
    const double f1 = -3.0 * r / (denom * denom * sqrt(denom));
    const double f2 = -3.0 / (denom * denom * sqrt(denom)) + 15.0 * r2 / (denom * denom * denom * sqrt(denom));
    const double n1 = -3.0 / (safe_r2 * safe_r2);
    const double n2 = 12.0 / (safe_r2 * safe_r2 * r);
    const double rf = f1 + s1 * newton + (split - 1.0) * n1;
    const double rs = f2 + s2 * newton + 2.0 * s1 * n1 + (split - 1.0) * n2;
    const double fo = rf / r;
    const double cr = rs / r2 - rf / (r2 * r);
    cx += pref * (mdx * fo + 0.5 * tr * dx * fo + 0.5 * dx * dmd * cr);
    cy += pref * (mdy * fo + 0.5 * tr * dy * fo + 0.5 * dy * dmd * cr);
    cz += pref * (mdz * fo + 0.5 * tr * dz * fo + 0.5 * dz * dmd * cr);
  }
  ax += cx;
  ay += cy;
  az += cz;
}

// Strict-reference acceptance of an internal node (tree_pm_acceptance.hpp:54-139,
// tree_interaction_common.hpp:36-89) for uniform softening and no history.
__device__ __forceinline__ bool accept64(double h, double offset, bool inside, bool contained, double r2,
                                         const WalkParams& P) {
  const double r = sqrt(fmax(r2, 1.0e-30));
  const bool env_ok = P.quad && (2.0 * h / r) < 0.08;
  bool mac = false;
  if (!inside) {
    if (P.criterion == 0) mac = (2.0 * h / sqrt(r2 + 1.0e-30)) < P.theta;
    else mac = ((2.0 * h + offset) / sqrt(r2 + 1.0e-30)) < P.theta;
  }
  return mac && env_ok && contained && !inside && isfinite(r2);
}

// ---- PM interpolation (pm_solver.cpp:925-951, :5900-5957) --------------------
__device__ __forceinline__ double wrap_position(double x, double L) {
  const double w = fmod(x, L);
  return w < 0.0 ? w + L : w;
}
__device__ __forceinline__ int wrap_index(long i, int n) {
  long w = i % n;
  if (w < 0) w += n;
  return static_cast<int>(w);
}
struct Stencil64 {
  int idx[3];
  double w[3];
  int count;
};
__device__ __forceinline__ Stencil64 stencil64(double g, int scheme, int n) {
  Stencil64 s;
  if (scheme == 0) {
    const double base = floor(g);
    const long i0 = static_cast<long>(base);
    const double t = g - base;
    s.idx[0] = wrap_index(i0, n);
    s.idx[1] = wrap_index(i0 + 1, n);
    s.idx[2] = 0;
    s.w[0] = 1.0 - t;
    s.w[1] = t;
    s.w[2] = 0.0;
    s.count = 2;
  } else {
    const double c = floor(g + 0.5);
    const long ic = static_cast<long>(c);
    const double d = g - c;
    s.idx[0] = wrap_index(ic - 1, n);
    s.idx[1] = wrap_index(ic, n);
    s.idx[2] = wrap_index(ic + 1, n);
    s.w[0] = 0.5 * (0.5 - d) * (0.5 - d);
    s.w[1] = 0.75 - d * d;
    s.w[2] = 0.5 * (0.5 + d) * (0.5 + d);
    s.count = 3;
  }
  return s;
}

__device__ void interp64(const WalkParams& P, const double* __restrict__ fx, const double* __restrict__ fy,
                         const double* __restrict__ fz, double x, double y, double z, double& gx, double& gy,
                         double& gz) {
  const Stencil64 sx = stencil64(wrap_position(x, P.L[0]) * (static_cast<double>(P.mesh[0]) / P.L[0]), P.assignment, P.mesh[0]);
  const Stencil64 sy = stencil64(wrap_position(y, P.L[1]) * (static_cast<double>(P.mesh[1]) / P.L[1]), P.assignment, P.mesh[1]);
  const Stencil64 sz = stencil64(wrap_position(z, P.L[2]) * (static_cast<double>(P.mesh[2]) / P.L[2]), P.assignment, P.mesh[2]);
  gx = gy = gz = 0.0;
  for (int a = 0; a < sx.count; ++a)
    for (int b = 0; b < sy.count; ++b) {
      const std::size_t row = (static_cast<std::size_t>(sx.idx[a]) * P.mesh[1] + sy.idx[b]) * P.mesh_row;
      for (int c = 0; c < sz.count; ++c) {
        const double w = sx.w[a] * sy.w[b] * sz.w[c];
        const std::size_t i = row + sz.idx[c];
        gx += w * fx[i];
        gy += w * fy[i];
        gz += w * fz[i];
      }
    }
}

// Stencil from a 32-bit fixed-point coordinate on a power-of-two mesh: the
// cell index is the top log2(n) bits, the fraction the remaining bits.
struct Stencil32 {
  int idx[3];
  float w[3];
  int count;
};
__device__ __forceinline__ Stencil32 stencil32(uint32_t q, int scheme, int n, int lg) {
  Stencil32 s;
  const int sh = 32 - lg;
  const int cell = static_cast<int>(q >> sh);
  const float t = static_cast<float>(q & ((sh == 32) ? 0xFFFFFFFFu : ((1u << sh) - 1u))) * __builtin_ldexpf(1.0f, -sh);
  if (scheme == 0) {
    s.idx[0] = cell;
    s.idx[1] = (cell + 1) & (n - 1);
    s.idx[2] = 0;
    s.w[0] = 1.0f - t;
    s.w[1] = t;
    s.w[2] = 0.0f;
    s.count = 2;
  } else {
    const int c = t >= 0.5f ? cell + 1 : cell;
    const float d = t >= 0.5f ? t - 1.0f : t;
    s.idx[0] = (c - 1) & (n - 1);
    s.idx[1] = c & (n - 1);
    s.idx[2] = (c + 1) & (n - 1);
    s.w[0] = 0.5f * (0.5f - d) * (0.5f - d);
    s.w[1] = 0.75f - d * d;
    s.w[2] = 0.5f * (0.5f + d) * (0.5f + d);
    s.count = 3;
  }
  return s;
}

__device__ void interp32(const WalkParams& P, const float* __restrict__ fx, const float* __restrict__ fy,
                         const float* __restrict__ fz, uint32_t qx, uint32_t qy, uint32_t qz, float& gx, float& gy,
                         float& gz) {
  const Stencil32 sx = stencil32(qx, P.assignment, P.mesh[0], P.mesh_log2[0]);
  const Stencil32 sy = stencil32(qy, P.assignment, P.mesh[1], P.mesh_log2[1]);
  const Stencil32 sz = stencil32(qz, P.assignment, P.mesh[2], P.mesh_log2[2]);
  gx = gy = gz = 0.0f;
  for (int a = 0; a < sx.count; ++a)
    for (int b = 0; b < sy.count; ++b) {
      const std::size_t row = (static_cast<std::size_t>(sx.idx[a]) * P.mesh[1] + sy.idx[b]) * P.mesh_row;
      const float wab = sx.w[a] * sy.w[b];
      for (int c = 0; c < sz.count; ++c) {
        const float w = wab * sz.w[c];
        const std::size_t i = row + sz.idx[c];
        gx = fmaf(w, fx[i], gx);
        gy = fmaf(w, fy[i], gy);
        gz = fmaf(w, fz[i], gz);
      }
    }
}

// XCD-aware block order: the dispatcher sends block b to XCD b % 8; give each
// XCD a contiguous range of target chunks so neighbouring targets share an L2.
__device__ __forceinline__ uint32_t xcd_logical_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb / 8, r = nb % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}

__device__ __forceinline__ void flush_counters(unsigned long long* counters, const uint32_t* c) {
  for (int i = 0; i < kWalkCounters; ++i) {
    const unsigned long long v = wave_sum(static_cast<unsigned long long>(c[i]));
    if (lane_id() == 0 && v) atomicAdd(&counters[i], v);
  }
}

__global__ __launch_bounds__(kWalkThreads) void walk64_kernel(
    WalkParams P, const NodeLink* __restrict__ link, const NodeGeo64* __restrict__ geo,
    const NodeMom64* __restrict__ mom, const NodeQuad64* __restrict__ quad, const uint32_t* __restrict__ order,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const double* __restrict__ sm, const uint32_t* __restrict__ tpos, const uint32_t* __restrict__ tslot,
    const double* __restrict__ fx, const double* __restrict__ fy, const double* __restrict__ fz, WalkOutputs out,
    unsigned long long* __restrict__ counters) {
  const uint32_t k = xcd_logical_block() * kWalkThreads + threadIdx.x;
  const bool valid = k < P.n_targets;
  const uint32_t tp = valid ? (P.identity_targets ? k : tpos[k]) : 0u;
  const uint32_t slot = valid ? (P.identity_targets ? order[k] : tslot[k]) : 0u;
  const double px = valid ? sx[tp] : 0.0, py = valid ? sy[tp] : 0.0, pz = valid ? sz[tp] : 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (P.pm_enabled && valid) interp64(P, fx, fy, fz, px, py, pz, gx, gy, gz);
  double ax = 0.0, ay = 0.0, az = 0.0;
  uint32_t cnt[kWalkCounters] = {0, 0, 0, 0, 0};
  uint32_t nxt = valid ? 0u : P.node_count;
  const double L0 = P.L[0], L1 = P.L[1], L2 = P.L[2];
  while (true) {
    const uint32_t n = __builtin_amdgcn_readfirstlane(wave_min(nxt));
    if (n >= P.node_count) break;
    const NodeLink lk = link[n];
    const NodeGeo64 g = geo[n];
    const bool act = nxt == n;
    bool do_leaf = false;
    if (act) {
      ++cnt[0];
      const double h = g.half;
      const double ndx = mi64(g.cx - px, L0), ndy = mi64(g.cy - py, L1), ndz = mi64(g.cz - pz, L2);
      const double ax_ = fabs(ndx), ay_ = fabs(ndy), az_ = fabs(ndz);
      const double ex = fmax(0.0, ax_ - h), ey = fmax(0.0, ay_ - h), ez = fmax(0.0, az_ - h);
      if (!within_cutoff(ex * ex + ey * ey + ez * ez, P.rcut, P.rcut2)) {
        ++cnt[1];
        nxt = n + lk.skip;
      } else if (lk.child_count == 0) {
        ++cnt[4];
        do_leaf = true;
        nxt = n + 1;
      } else {
        const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
        const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
        const bool contained = within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
        const NodeMom64 M = mom[n];
        const double dx = mi64(M.comx - px, L0), dy = mi64(M.comy - py, L1), dz = mi64(M.comz - pz, L2);
        const double r2 = dx * dx + dy * dy + dz * dz;
        const NodeQuad64 Q = quad[n];
        if (accept64(h, Q.offset, inside, contained, r2, P)) {
          ++cnt[3];
          node_accel64(M, Q, dx, dy, dz, P, ax, ay, az);
          nxt = n + lk.skip;
        } else {
          nxt = n + 1;
        }
      }
    }
    if (__ballot(do_leaf)) {
      const uint32_t b0 = lk.begin, e0 = lk.begin + lk.count;
      for (uint32_t j = b0; j < e0; ++j) {
        const double sxj = sx[j], syj = sy[j], szj = sz[j];
        const double mj = sm ? sm[j] : P.uniform_mass;
        if (do_leaf && j != tp) {
          const double qx = mi64(sxj - px, L0), qy = mi64(syj - py, L1), qz = mi64(szj - pz, L2);
          const double r2 = qx * qx + qy * qy + qz * qz;
          if (r2 <= P.rcut2) {
            ++cnt[2];
            if (!(qx == 0.0 && qy == 0.0 && qz == 0.0)) {
              const double f = residual_inv_r3(r2, P.eps, P.a) * P.G;
              ax += f * mj * qx;
              ay += f * mj * qy;
              az += f * mj * qz;
            }
          }
        }
      }
    }
  }
  flush_counters(counters, cnt);
  if (!valid) return;
  double* oax = static_cast<double*>(out.ax);
  double* oay = static_cast<double*>(out.ay);
  double* oaz = static_cast<double*>(out.az);
  oax[slot] = gx + ax;
  oay[slot] = gy + ay;
  oaz[slot] = gz + az;
  if (out.pm_x) {
    static_cast<double*>(out.pm_x)[slot] = gx;
    static_cast<double*>(out.pm_y)[slot] = gy;
    static_cast<double*>(out.pm_z)[slot] = gz;
  }
  if (out.sr_x) {
    static_cast<double*>(out.sr_x)[slot] = ax;
    static_cast<double*>(out.sr_y)[slot] = ay;
    static_cast<double*>(out.sr_z)[slot] = az;
  }
}

// ---- mixed precision ---------------------------------------------------------
__device__ __forceinline__ float fx_delta(uint32_t a, uint32_t b, float scale) {
  return static_cast<float>(static_cast<int32_t>(a - b)) * scale;
}

__device__ __forceinline__ float gaussian_h(float t, const WalkParams& P) {
  const float s = (t - P.t_mid) * P.t_inv_half;
  float h = P.poly[P.poly_degree];
  for (int i = P.poly_degree - 1; i >= 0; --i) h = fmaf(h, s, P.poly[i]);
  return h;
}

__device__ void node_accel32(const NodeMom32& M, const NodeQuad32& Q, float dx, float dy, float dz,
                             const WalkParams& P, float& ax, float& ay, float& az) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float safe_r2 = fmaxf(r2, 1.0e-30f);
  const float r = sqrtf(safe_r2);
  const float eps = static_cast<float>(P.eps);
  const float denom = fmaxf(r2 + eps * eps, 1.0e-30f);
  const float rs_d = rsqrtf(denom);
  const float soft = rs_d * rs_d * rs_d;
  const float a = static_cast<float>(P.a);
  const float q = 0.5f * r / a;
  const float split = erfcf(q) + 2.0f * static_cast<float>(kInvSqrtPi) * q * expf(-q * q);
  const float newton = 1.0f / (safe_r2 * r);
  const float radial = soft + (split - 1.0f) * newton;
  const float pref = static_cast<float>(P.G);
  float cx = pref * M.mass * radial * dx;
  float cy = pref * M.mass * radial * dy;
  float cz = pref * M.mass * radial * dz;
  if (P.quad) {
    const float tr = Q.trace;
    const float third = 1.0f / 3.0f;
    const float mxx = (Q.qxx + tr) * third, mxy = Q.qxy * third, mxz = Q.qxz * third;
    const float myy = (Q.qyy + tr) * third, myz = Q.qyz * third, mzz = (Q.qzz + tr) * third;
    const float mdx = mxx * dx + mxy * dy + mxz * dz;
    const float mdy = mxy * dx + myy * dy + myz * dz;
    const float mdz = mxz * dx + myz * dy + mzz * dz;
    const float dmd = dx * mdx + dy * mdy + dz * mdz;
    const float a2 = a * a;
    const float ex = expf(-r2 / (4.0f * a2));
    const float isp = static_cast<float>(kInvSqrtPi);
    const float s1 = -0.5f * isp * r2 * ex / (a * a2);
    const float s2 = isp * ex * (-r / (a * a2) + 0.25f * r * r2 / (a * a2 * a2));
    const float d52 = soft / (denom);  // denom^-5/2
    const float f1 = -3.0f * r * d52;
    const float f2 = -3.0f * d52 + 15.0f * r2 * d52 / denom;
    const float n1 = -3.0f / (safe_r2 * safe_r2);
    const float n2 = 12.0f / (safe_r2 * safe_r2 * r);
    const float rf = f1 + s1 * newton + (split - 1.0f) * n1;
    const float rsec = f2 + s2 * newton + 2.0f * s1 * n1 + (split - 1.0f) * n2;
    const float fo = rf / r;
    const float cr = rsec / r2 - rf / (r2 * r);
    cx += pref * (mdx * fo + 0.5f * tr * dx * fo + 0.5f * dx * dmd * cr);
    cy += pref * (mdy * fo + 0.5f * tr * dy * fo + 0.5f * dy * dmd * cr);
    cz += pref * (mdz * fo + 0.5f * tr * dz * fo + 0.5f * dz * dmd * cr);
  }
  ax += cx;
  ay += cy;
  az += cz;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

__global__ __launch_bounds__(kWalkThreads) void walkmix_kernel(
    WalkParams P, const NodeLink* __restrict__ link, const NodeGeo32* __restrict__ geo,
    const NodeMom32* __restrict__ mom, const NodeQuad32* __restrict__ quad, const uint32_t* __restrict__ order,
    const uint32_t* __restrict__ qxs, const uint32_t* __restrict__ qys, const uint32_t* __restrict__ qzs,
    const float* __restrict__ sm, const uint32_t* __restrict__ tpos, const uint32_t* __restrict__ tslot,
    const float* __restrict__ fx, const float* __restrict__ fy, const float* __restrict__ fz, WalkOutputs out,
    unsigned long long* __restrict__ counters) {
  const uint32_t k = xcd_logical_block() * kWalkThreads + threadIdx.x;
  const bool valid = k < P.n_targets;
  const uint32_t tp = valid ? (P.identity_targets ? k : tpos[k]) : 0u;
  const uint32_t slot = valid ? (P.identity_targets ? order[k] : tslot[k]) : 0u;
  const uint32_t px = valid ? qxs[tp] : 0u, py = valid ? qys[tp] : 0u, pz = valid ? qzs[tp] : 0u;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  if (P.pm_enabled && valid) interp32(P, fx, fy, fz, px, py, pz, gx, gy, gz);
  const float s0 = P.fx_to_len[0], s1 = P.fx_to_len[1], s2 = P.fx_to_len[2];
  const float rcut = static_cast<float>(P.rcut);
  const float rcut2 = static_cast<float>(P.rcut2);
  const float prune2 = static_cast<float>(P.rcut2 * (1.0 + 1.0e-5));
  const float contain2 = static_cast<float>(P.rcut2 * (1.0 - 1.0e-5));
  const float eps2 = static_cast<float>(P.eps * P.eps);
  const float theta = static_cast<float>(P.theta);
  const float gm = static_cast<float>(P.G);
  float sx_ = 0.0f, cx_ = 0.0f, sy_ = 0.0f, cy_ = 0.0f, sz_ = 0.0f, cz_ = 0.0f;
  uint32_t cnt[kWalkCounters] = {0, 0, 0, 0, 0};
  uint32_t nxt = valid ? 0u : P.node_count;
  (void)rcut;
  while (true) {
    const uint32_t n = __builtin_amdgcn_readfirstlane(wave_min(nxt));
    if (n >= P.node_count) break;
    const NodeLink lk = link[n];
    const NodeGeo32 g = geo[n];
    const bool act = nxt == n;
    bool do_leaf = false;
    if (act) {
      ++cnt[0];
      const float h = g.half;
      const float ax_ = fabsf(fx_delta(g.cx, px, s0)), ay_ = fabsf(fx_delta(g.cy, py, s1)),
                  az_ = fabsf(fx_delta(g.cz, pz, s2));
      const float ex = fmaxf(0.0f, ax_ - h), ey = fmaxf(0.0f, ay_ - h), ez = fmaxf(0.0f, az_ - h);
      if (ex * ex + ey * ey + ez * ez > prune2) {
        ++cnt[1];
        nxt = n + lk.skip;
      } else if (lk.child_count == 0) {
        ++cnt[4];
        do_leaf = true;
        nxt = n + 1;
      } else {
        const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
        const float mx = ax_ + h, my = ay_ + h, mz = az_ + h;
        const bool contained = mx * mx + my * my + mz * mz <= contain2;
        const NodeMom32 M = mom[n];
        const float dx = fx_delta(M.comx, px, s0), dy = fx_delta(M.comy, py, s1), dz = fx_delta(M.comz, pz, s2);
        const float r2 = dx * dx + dy * dy + dz * dz;
        const NodeQuad32 Q = quad[n];
        const float r = sqrtf(fmaxf(r2, 1.0e-30f));
        const bool env_ok = P.quad && (2.0f * h / r) < 0.08f;
        bool mac = false;
        if (!inside) {
          const float rr = sqrtf(r2 + 1.0e-30f);
          mac = P.criterion == 0 ? (2.0f * h / rr) < theta : ((2.0f * h + Q.offset) / rr) < theta;
        }
        if (mac && env_ok && contained && !inside) {
          ++cnt[3];
          float mxa = 0.0f, mya = 0.0f, mza = 0.0f;
          node_accel32(M, Q, dx, dy, dz, P, mxa, mya, mza);
          kahan_add(sx_, cx_, mxa);
          kahan_add(sy_, cy_, mya);
          kahan_add(sz_, cz_, mza);
          nxt = n + lk.skip;
        } else {
          nxt = n + 1;
        }
      }
    }
    if (__ballot(do_leaf)) {
      const uint32_t b0 = lk.begin, e0 = lk.begin + lk.count;
      float lx = 0.0f, ly = 0.0f, lz = 0.0f;
      for (uint32_t j = b0; j < e0; ++j) {
        const uint32_t qx = qxs[j], qy = qys[j], qz = qzs[j];
        const float mj = sm ? sm[j] : P.uniform_mass_f;
        const float dx = fx_delta(qx, px, s0), dy = fx_delta(qy, py, s1), dz = fx_delta(qz, pz, s2);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (do_leaf && j != tp && r2 <= rcut2) {
          ++cnt[2];
          if (r2 > 0.0f) {
            const float rinv = rsqrtf(r2 + eps2);
            const float soft = rinv * rinv * rinv;
            const float hcoef = gaussian_h(r2 * P.c4, P);
            const float f = fmaf(-hcoef, P.c8, soft) * mj;
            lx = fmaf(f, dx, lx);
            ly = fmaf(f, dy, ly);
            lz = fmaf(f, dz, lz);
          }
        }
      }
      if (do_leaf) {
        kahan_add(sx_, cx_, lx * gm);
        kahan_add(sy_, cy_, ly * gm);
        kahan_add(sz_, cz_, lz * gm);
      }
    }
  }
  flush_counters(counters, cnt);
  if (!valid) return;
  const float ax = sx_ - cx_, ay = sy_ - cy_, az = sz_ - cz_;
  static_cast<float*>(out.ax)[slot] = gx + ax;
  static_cast<float*>(out.ay)[slot] = gy + ay;
  static_cast<float*>(out.az)[slot] = gz + az;
  if (out.pm_x) {
    static_cast<float*>(out.pm_x)[slot] = gx;
    static_cast<float*>(out.pm_y)[slot] = gy;
    static_cast<float*>(out.pm_z)[slot] = gz;
  }
  if (out.sr_x) {
    static_cast<float*>(out.sr_x)[slot] = ax;
    static_cast<float*>(out.sr_y)[slot] = ay;
    static_cast<float*>(out.sr_z)[slot] = az;
  }
}

// H(t) = L(sqrt t) / t^{3/2} (tree_pm_split_kernel.hpp:109-128).
double gaussian_h_exact(double t) {
  if (t < 0.015625) {
    constexpr double c = 2.25675833419102514779231780624309034;
    return c * (1.0 / 3.0 + t * (-1.0 / 5.0 + t * (1.0 / 14.0 + t * (-1.0 / 54.0 + t * (1.0 / 264.0 +
           t * (-1.0 / 1560.0 + t * (1.0 / 10800.0 - t / 85680.0)))))));
  }
  const double q = std::sqrt(t);
  return (std::erf(q) - 2.0 * kInvSqrtPi * q * std::exp(-t)) / (t * q);
}

}  // namespace

void fit_gaussian_coefficient(WalkParams& p, double t_max) {
  const double mid = 0.5 * t_max, half = 0.5 * t_max;
  for (int deg = 6; deg <= kMaxPolyDegree; ++deg) {
    const int m = deg + 1;
    // Chebyshev interpolation at Chebyshev nodes, then conversion to monomials in s.
    std::vector<double> f(m), c(m, 0.0);
    for (int j = 0; j < m; ++j) {
      const double s = std::cos(M_PI * (j + 0.5) / m);
      f[j] = gaussian_h_exact(mid + half * s);
    }
    for (int k = 0; k < m; ++k) {
      double acc = 0.0;
      for (int j = 0; j < m; ++j) acc += f[j] * std::cos(M_PI * k * (j + 0.5) / m);
      c[k] = (k == 0 ? 1.0 : 2.0) * acc / m;
    }
    // T_k(s) in monomials: T0=1, T1=s, T_{k+1}=2 s T_k - T_{k-1}
    std::vector<std::vector<double>> T(m, std::vector<double>(m, 0.0));
    T[0][0] = 1.0;
    if (m > 1) T[1][1] = 1.0;
    for (int k = 1; k + 1 < m; ++k)
      for (int i = 0; i < m; ++i) {
        double v = -T[k - 1][i];
        if (i > 0) v += 2.0 * T[k][i - 1];
        T[k + 1][i] = v;
      }
    std::vector<double> mono(m, 0.0);
    for (int k = 0; k < m; ++k)
      for (int i = 0; i < m; ++i) mono[i] += c[k] * T[k][i];
    double worst = 0.0;
    for (int j = 0; j <= 4000; ++j) {
      const double t = t_max * j / 4000.0;
      const double s = (t - mid) / half;
      double h = mono[deg];
      for (int i = deg - 1; i >= 0; --i) h = h * s + mono[i];
      const double e = gaussian_h_exact(t);
      worst = std::max(worst, std::abs(h - e) / std::abs(e));
    }
    if (worst < 2.0e-9 || deg == kMaxPolyDegree) {
      p.poly_degree = deg;
      for (int i = 0; i <= deg; ++i) p.poly[i] = static_cast<float>(mono[i]);
      p.t_mid = static_cast<float>(mid);
      p.t_inv_half = static_cast<float>(1.0 / half);
      return;
    }
  }
}

void walk_fp64(const Tree& t, const WalkParams& p, const double* sx, const double* sy, const double* sz,
               const double* sm, const uint32_t* target_pos, const uint32_t* target_slot, const double* fx,
               const double* fy, const double* fz, const WalkOutputs& out, unsigned long long* counters,
               hipStream_t s) {
  if (p.n_targets == 0) return;
  walk64_kernel<<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, 0, s>>>(
      p, t.link.as<NodeLink>(), t.geo64.as<NodeGeo64>(), t.mom64.as<NodeMom64>(), t.quad64.as<NodeQuad64>(),
      t.order.as<uint32_t>(), sx, sy, sz, sm, target_pos, target_slot, fx, fy, fz, out, counters);
  CHUI_LAUNCH_CHECK("walk64_kernel");
}

void walk_mixed(const Tree& t, const WalkParams& p, const uint32_t* qx, const uint32_t* qy, const uint32_t* qz,
                const float* sm, const uint32_t* target_pos, const uint32_t* target_slot, const float* fx,
                const float* fy, const float* fz, const WalkOutputs& out, unsigned long long* counters,
                hipStream_t s) {
  if (p.n_targets == 0) return;
  walkmix_kernel<<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, 0, s>>>(
      p, t.link.as<NodeLink>(), t.geo32.as<NodeGeo32>(), t.mom32.as<NodeMom32>(), t.quad32.as<NodeQuad32>(),
      t.order.as<uint32_t>(), qx, qy, qz, sm, target_pos, target_slot, fx, fy, fz, out, counters);
  CHUI_LAUNCH_CHECK("walkmix_kernel");
}

}  // namespace chui

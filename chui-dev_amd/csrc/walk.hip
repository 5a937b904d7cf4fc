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
// walk64_kernel  : FP64, the reference's arithmetic (erf/exp split
//                  kernel, nearbyint minimum image, monopole + quadrupole residual).
// walkmix_kernel : 32-bit fixed-point positions (minimum image = int32
//                  subtraction), FP32 pair kernel with a cubic-cell LDS table
//                  for the Gaussian coefficient H(t), per-drain partial sums
//                  folded into a Kahan-compensated FP32 accumulator.
#include "walk.hpp"

#include "primitives.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
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

// d/dt of residual_inv_r3 at t = r^2. With q = r/(2a) and
// L(r) = erf(q) - 2q exp(-q^2)/sqrt(pi):
//   d/dt (t + eps^2)^{-3/2} = -1.5 (t + eps^2)^{-5/2},
//   d/dt [L(r)/r^3] = exp(-q^2) / (4 sqrt(pi) a^3 r^2) - 1.5 L(r)/r^5,
// the latter from the series of L(r)/r^3 = H(q^2)/(8a^3) below q = 1/8.
__device__ __forceinline__ double residual_inv_r3_dt(double r2, double eps, double a) {
  const double s = r2 + eps * eps;
  const double dsoft = s > 0.0 ? -1.5 / (s * s * sqrt(s)) : 0.0;
  const double r = sqrt(r2);
  const double q = (0.5 * r) / a;
  double dlong;
  if (q < 0.125) {
    constexpr double c = 2.25675833419102514779231780624309034;
    const double u = q * q;
    const double dh = c * (-1.0 / 5.0 + u * (2.0 / 14.0 + u * (-3.0 / 54.0 + u * (4.0 / 264.0 +
                      u * (-5.0 / 1560.0 + u * (6.0 / 10800.0 - 7.0 * u / 85680.0))))));
    dlong = (0.125 / (a * a * a)) * dh * (0.25 / (a * a));  // dH/du * du/dt, u = t / (4 a^2)
  } else {
    dlong = exp(-q * q) * kInvSqrtPi / (4.0 * a * a * a * r2) - 1.5 * long_factor(r, a) / (r2 * r2 * r);
  }
  return dsoft - dlong;
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

// Forward-mode derivative of a scalar with respect to the node offset d
// (three partials), for the Jacobian of the node expansion in the force VJP.
struct D3 {
  double v, x, y, z;
};
__device__ __forceinline__ D3 d3c(double c) { return D3{c, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.v + b.v, a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.v - b.v, a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator-(D3 a) { return D3{-a.v, -a.x, -a.y, -a.z}; }
__device__ __forceinline__ D3 operator*(D3 a, D3 b) {
  return D3{a.v * b.v, a.x * b.v + a.v * b.x, a.y * b.v + a.v * b.y, a.z * b.v + a.v * b.z};
}
__device__ __forceinline__ D3 operator*(D3 a, double c) { return D3{a.v * c, a.x * c, a.y * c, a.z * c}; }
__device__ __forceinline__ D3 operator*(double c, D3 a) { return a * c; }
__device__ __forceinline__ D3 operator+(D3 a, double c) { return D3{a.v + c, a.x, a.y, a.z}; }
__device__ __forceinline__ D3 operator+(double c, D3 a) { return a + c; }
__device__ __forceinline__ D3 operator-(D3 a, double c) { return D3{a.v - c, a.x, a.y, a.z}; }
__device__ __forceinline__ D3 operator-(double c, D3 a) { return D3{c - a.v, -a.x, -a.y, -a.z}; }
__device__ __forceinline__ D3 operator/(D3 a, D3 b) {
  const double q = a.v / b.v, ib = 1.0 / b.v;
  return D3{q, (a.x - q * b.x) * ib, (a.y - q * b.y) * ib, (a.z - q * b.z) * ib};
}
__device__ __forceinline__ D3 operator/(D3 a, double c) { return D3{a.v / c, a.x / c, a.y / c, a.z / c}; }
__device__ __forceinline__ D3 operator/(double c, D3 b) { return d3c(c) / b; }
__device__ __forceinline__ D3 d_scale(D3 a, double v, double dv) { return D3{v, a.x * dv, a.y * dv, a.z * dv}; }
__device__ __forceinline__ double f_max(double a, double c) { return fmax(a, c); }
__device__ __forceinline__ D3 f_max(D3 a, double c) { return a.v > c ? a : d3c(fmax(a.v, c)); }
__device__ __forceinline__ double f_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ D3 f_sqrt(D3 a) {
  const double v = sqrt(a.v);
  return d_scale(a, v, 0.5 / v);
}
__device__ __forceinline__ double f_exp(double a) { return exp(a); }
__device__ __forceinline__ D3 f_exp(D3 a) {
  const double v = exp(a.v);
  return d_scale(a, v, v);
}
// short_factor of a dual r: d/dr [erfc(q) + 2 q exp(-q^2)/sqrt(pi)] = -2 q^2 exp(-q^2) / (sqrt(pi) a), q = r/2a
__device__ __forceinline__ double f_short(double r, double a) { return short_factor(r, a); }
__device__ __forceinline__ D3 f_short(D3 r, double a) {
  const double q = (0.5 * r.v) / a;
  const double dv = (r.v <= 0.0 || isinf(q)) ? 0.0 : -2.0 * kInvSqrtPi * q * q * exp(-q * q) / a;
  return d_scale(r, short_factor(r.v, a), dv);
}

// monopolePlusQuadrupoleAccelPeriodicReference (tree_pm_coupling.cpp:743-827),
// on double (the walks) or on D3 (its Jacobian with respect to the offset).
template <class T>
__device__ __forceinline__ void node_accel_t(double mass, const NodeQuad64& Q, T dx, T dy, T dz,
                                             const WalkParams& P, T& cx, T& cy, T& cz) {
  const T r2 = dx * dx + dy * dy + dz * dz;
  const T safe_r2 = f_max(r2, 1.0e-30);
  const T r = f_sqrt(safe_r2);
  const double pe = P.eps;  // uniform softening: max(node eps, target eps) = eps
  const double eps2 = pe * pe;
  const T denom = f_max(r2 + eps2, 1.0e-30);
  const T soft = 1.0 / (denom * f_sqrt(denom));
  const T split = f_short(r, P.a);
  const T newton = 1.0 / (safe_r2 * r);
  const T radial = soft + (split - 1.0) * newton;
  const double pref = P.G;
  cx = pref * mass * radial * dx;
  cy = pref * mass * radial * dy;
  cz = pref * mass * radial * dz;
  if (P.quad) {
    const double tr = Q.trace;
    const double mxx = (Q.qxx + tr) / 3.0, mxy = Q.qxy / 3.0, mxz = Q.qxz / 3.0;
    const double myy = (Q.qyy + tr) / 3.0, myz = Q.qyz / 3.0, mzz = (Q.qzz + tr) / 3.0;
    const T mdx = mxx * dx + mxy * dy + mxz * dz;
    const T mdy = mxy * dx + myy * dy + myz * dz;
    const T mdz = mxz * dx + myz * dy + mzz * dz;
    const T dmd = dx * mdx + dy * mdy + dz * mdz;
    const double a = P.a, a2 = a * a;
    const T ex = f_exp(-r2 / (4.0 * a2));
    const T s1 = -0.5 * kInvSqrtPi * r2 * ex / (a * a2);
    const T s2 = kInvSqrtPi * ex * (-r / (a * a2) + 0.25 * r * r2 / (a * a2 * a2));
    const T f1 = -3.0 * r / (denom * denom * f_sqrt(denom));
    const T f2 = -3.0 / (denom * denom * f_sqrt(denom)) + 15.0 * r2 / (denom * denom * denom * f_sqrt(denom));
    const T n1 = -3.0 / (safe_r2 * safe_r2);
    const T n2 = 12.0 / (safe_r2 * safe_r2 * r);
    const T rf = f1 + s1 * newton + (split - 1.0) * n1;
    const T rs = f2 + s2 * newton + 2.0 * s1 * n1 + (split - 1.0) * n2;
    const T fo = rf / r;
    const T cr = rs / r2 - rf / (r2 * r);
    cx = cx + pref * (mdx * fo + 0.5 * tr * dx * fo + 0.5 * dx * dmd * cr);
    cy = cy + pref * (mdy * fo + 0.5 * tr * dy * fo + 0.5 * dy * dmd * cr);
    cz = cz + pref * (mdz * fo + 0.5 * tr * dz * fo + 0.5 * dz * dmd * cr);
  }
}

__device__ __forceinline__ void node_accel64(const NodeMom64& M, const NodeQuad64& Q, double dx, double dy, double dz,
                                             const WalkParams& P, double& ax, double& ay, double& az) {
  double cx, cy, cz;
  node_accel_t<double>(M.mass, Q, dx, dy, dz, P, cx, cy, cz);
  ax += cx;
  ay += cy;
  az += cz;
}

// Force VJP at an accepted node N of target t (offset d = COM_N - x_t, lam of
// t): the node's force a(d, Q) depends on x_t through d, and on the positions
// of N's particles through COM_N (= d + x_t) and the quadrupole Q.
//   g = J^T lam with J = da/dd (forward-mode D3);
//   s_t -= g (da/dx_t = -J);
//   U_N += g, V_N[p] += lam . da/dQ_p (a is linear in the seven stored Q
//   components qxx..qzz, trace: each column is the expansion with mass 0 and
//   Q = e_p).
// U_N and V_N are spread over N's particles afterwards (vjp_node_push).
__device__ __attribute__((noinline)) void jvp_node64(const NodeMom64& M, const NodeQuad64& Q, double dx, double dy,
                                                     double dz, const WalkParams& P, double lx, double ly,
                                                     double lz, double* __restrict__ acc, double& sx, double& sy,
                                                     double& sz) {
  D3 cx, cy, cz;
  node_accel_t<D3>(M.mass, Q, D3{dx, 1.0, 0.0, 0.0}, D3{dy, 0.0, 1.0, 0.0}, D3{dz, 0.0, 0.0, 1.0}, P, cx, cy, cz);
  const double gx = lx * cx.x + ly * cy.x + lz * cz.x;
  const double gy = lx * cx.y + ly * cy.y + lz * cz.y;
  const double gz = lx * cx.z + ly * cy.z + lz * cz.z;
  sx -= gx;
  sy -= gy;
  sz -= gz;
  atomicAdd(acc + 0, gx);
  atomicAdd(acc + 1, gy);
  atomicAdd(acc + 2, gz);
  if (!P.quad) return;
  for (int q = 0; q < 7; ++q) {
    NodeQuad64 e{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    switch (q) {
      case 0: e.qxx = 1.0; break;
      case 1: e.qxy = 1.0; break;
      case 2: e.qxz = 1.0; break;
      case 3: e.qyy = 1.0; break;
      case 4: e.qyz = 1.0; break;
      case 5: e.qzz = 1.0; break;
      default: e.trace = 1.0; break;
    }
    double ex, ey, ez;
    node_accel_t<double>(0.0, e, dx, dy, dz, P, ex, ey, ez);
    atomicAdd(acc + 3 + q, lx * ex + ly * ey + lz * ez);
  }
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

// ---- adaptive acceptance pieces (FP64) -------------------------------------------
// Gaussian Hermite table (tree_pm_split_kernel.hpp:130-183): knots (H, step dH/dt)
// at t = i / 64, i = 0..4096.
constexpr int kGaussCells = 4096;
__device__ __forceinline__ double gauss_hermite(const double2* __restrict__ tab, double t) {
  const double cell = t * 64.0;
  const int i = static_cast<int>(cell);
  const double u = cell - static_cast<double>(i), u2 = u * u, u3 = u2 * u;
  const double2 k0 = tab[i], k1 = tab[i + 1];
  return (2 * u3 - 3 * u2 + 1) * k0.x + (u3 - 2 * u2 + u) * k0.y + (-2 * u3 + 3 * u2) * k1.x + (u3 - u2) * k1.y;
}
// dimensionlessCoefficient: series below t = 1/64, analytic from t = 64
__device__ __forceinline__ double gauss_coefficient(const double2* __restrict__ tab, double t) {
  if (t < 0.015625) return series_h(t);
  if (!(t < 64.0)) return long_inv_r3(sqrt(t), 0.5);
  return gauss_hermite(tab, t);
}
// Leaf-pair residual of the per-target walk with the Gaussian lookup
// (tree_pm_coupling.cpp:4216-4227), r2 > 0.
__device__ __forceinline__ double table_residual64(const double2* __restrict__ tab, double r2, const WalkParams& P) {
  const double t = r2 * P.inv2a2;
  double lr;
  if (t < 64.0) {
    lr = gauss_coefficient(tab, t) * P.inv8a3;
  } else {
    const double tt = 0.25 * (r2 / P.a) / P.a;  // longRangeInvR3 beyond the table
    lr = (!(tt >= 0.0) || tt >= 64.0) ? long_inv_r3(sqrt(r2), P.a)
                                      : ((0.125 * gauss_coefficient(tab, tt) / P.a) / P.a) / P.a;
  }
  return softened_inv_r3(r2, P.eps) - lr;
}

// treePmResidualRadialCoefficients (tree_pm_split_kernel.hpp:245-282).
__device__ __forceinline__ void residual_radial64(double x, double eps, double a, double& value, double& first,
                                                  double& second) {
  const double d = x + eps * eps;
  const double soft = 1.0 / (d * sqrt(d));
  const double t = (0.25 * x / a) / a;
  constexpr double c = 2.25675833419102514779231780624309034;
  double h, dh, ddh;
  if (t < 0.015625) {
    constexpr double k[8] = {1.0 / 3.0, -1.0 / 5.0, 1.0 / 14.0, -1.0 / 54.0,
                             1.0 / 264.0, -1.0 / 1560.0, 1.0 / 10800.0, -1.0 / 85680.0};
    h = k[7]; dh = 0.0; ddh = 0.0;
#pragma unroll
    for (int i = 7; i > 0; --i) {
      ddh = ddh * t + 2.0 * dh;
      dh = dh * t + h;
      h = h * t + k[i - 1];
    }
    h *= c; dh *= c; ddh *= c;
  } else {
    const double q = sqrt(t);
    const double e = exp(-t);
    h = (erf(q) - 0.5 * c * q * e) / (t * q);
    dh = (0.5 * c * e - 1.5 * h) / t;
    ddh = (-0.5 * c * e - 2.5 * dh) / t;
  }
  const double scale = ((0.125 / a) / a) / a;
  const double dt_dx = (0.25 / a) / a;
  value = soft - scale * h;
  first = -3.0 * soft / d - 2.0 * scale * dt_dx * dh;
  second = 15.0 * soft / (d * d) - 4.0 * scale * dt_dx * dt_dx * ddh;
}

// monopolePlusQuadrupoleAccelPeriodic (tree_pm_coupling.cpp:829-894); non-finite
// coefficients use the reference expansion.
__device__ __forceinline__ void node_accel_radial64(const NodeMom64& M, const NodeQuad64& Q, double dx, double dy,
                                                    double dz, const WalkParams& P, double& ax, double& ay,
                                                    double& az) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double safe_r2 = fmax(r2, 1.0e-30);
  double v, fo, cr;
  residual_radial64(safe_r2, P.eps, P.a, v, fo, cr);
  if (!isfinite(v) || !isfinite(fo) || !isfinite(cr)) {
    node_accel64(M, Q, dx, dy, dz, P, ax, ay, az);
    return;
  }
  const double pref = P.G;
  double cx = pref * M.mass * v * dx;
  double cy = pref * M.mass * v * dy;
  double cz = pref * M.mass * v * dz;
  if (P.quad) {
    const double tr = Q.trace;
    const double mxx = (Q.qxx + tr) / 3.0, mxy = Q.qxy / 3.0, mxz = Q.qxz / 3.0;
    const double myy = (Q.qyy + tr) / 3.0, myz = Q.qyz / 3.0, mzz = (Q.qzz + tr) / 3.0;
    const double mdx = mxx * dx + mxy * dy + mxz * dz;
    const double mdy = mxy * dx + myy * dy + myz * dz;
    const double mdz = mxz * dx + myz * dy + mzz * dz;
    const double dmd = dx * mdx + dy * mdy + dz * mdz;
    cx += pref * (mdx * fo + 0.5 * tr * dx * fo + 0.5 * dx * dmd * cr);
    cy += pref * (mdy * fo + 0.5 * tr * dy * fo + 0.5 * dy * dmd * cr);
    cz += pref * (mdz * fo + 0.5 * tr * dz * fo + 0.5 * dz * dmd * cr);
  }
  ax += cx;
  ay += cy;
  az += cz;
}

// Out-of-line forms for kernels whose register budget the inlined expansions
// would exceed (accepted nodes are rare next to leaf pairs).
__device__ __attribute__((noinline)) void node_accel64_call(const NodeMom64& M, const NodeQuad64& Q, double dx,
                                                            double dy, double dz, const WalkParams& P, double& ax,
                                                            double& ay, double& az) {
  if (P.radial) node_accel_radial64(M, Q, dx, dy, dz, P, ax, ay, az);
  else node_accel64(M, Q, dx, dy, dz, P, ax, ay, az);
}

// residualAccuracyProxy (tree_pm_acceptance.hpp:14-51) for uniform softening.
__device__ __forceinline__ double residual_proxy64(double h, double offset, double mass, double r2,
                                                   const WalkParams& P) {
  if (!(r2 > 0.0) || !isfinite(r2)) return INFINITY;
  const double a = P.a;
  const double r = sqrt(r2);
  const double d = r2 + P.eps * P.eps;
  const double inv_d5 = 1.0 / (d * d * sqrt(d));
  const double inv_r3 = 1.0 / (r2 * r);
  const double inv_r4 = 1.0 / (r2 * r2);
  const double lf = long_factor(r, a);
  const double q = r / (2.0 * a);
  const double e = exp(-q * q);
  const double a3 = a * a * a;
  const double s1 = -0.5 * kInvSqrtPi * r2 * e / a3;
  const double s2 = kInvSqrtPi * e * (-r / a3 + 0.25 * r * r2 / (a3 * a * a));
  const double first = -3.0 * r * inv_d5 + s1 * inv_r3 - 3.0 * (-lf) * inv_r4;
  const double second = -3.0 * inv_d5 + 15.0 * r2 * inv_d5 / d + s2 * inv_r3 - 6.0 * s1 * inv_r4 +
                        12.0 * (-lf) * inv_r4 / r;
  const double width = 2.0 * h;
  const double rho = 1.7320508075688772935274463415058723 * h + offset;
  const double proxy = P.G * mass * fmax(width * width * inv_r4,
                                         0.5 * rho * rho * (3.0 * fabs(first) + fabs(r * second - first)));
  return isfinite(proxy) && proxy >= 0.0 ? proxy : INFINITY;
}

// acceptTreePmNode (tree_pm_acceptance.hpp:54-139) under the adaptive policy for
// an internal node that neither holds the target nor leaves the cutoff
// container (those are opened first, tree_pm_coupling.cpp:4094-4104).
__device__ __forceinline__ bool accept_adaptive64(double h, double offset, double mass, double r2, bool history,
                                                  double prev, const WalkParams& P) {
  const double r = sqrt(fmax(r2, 1.0e-30));
  const bool env_ok = (2.0 * h / r) < P.amax;
  bool mac;
  if (history && isfinite(prev) && prev > P.accel_floor) {
    const double allowance = P.alpha * prev;
    const double width = 2.0 * h;
    const double inv_r4 = 1.0 / (r2 * r2);
    const double lower = P.G * mass * (width * width * inv_r4);
    if (isfinite(allowance) && isfinite(lower) && lower > allowance) {
      mac = false;
    } else if (P.estimator == 1) {  // relative monopole
      mac = isfinite(allowance) && isfinite(lower);
    } else {
      const double est = residual_proxy64(h, offset, mass, r2, P);
      mac = isfinite(est) && isfinite(allowance) && est <= allowance;
    }
  } else {
    mac = ((2.0 * h + offset) / sqrt(r2 + 1.0e-30)) < P.theta;
  }
  return mac && env_ok && isfinite(r2);
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

// Stencil of one axis with a compile-time point count (CIC 2, TSC 3) so the
// gather below is fully unrolled and stays in registers.
template <int SCHEME>
__device__ __forceinline__ void stencil64(double g, int n, int (&idx)[SCHEME + 2], double (&w)[SCHEME + 2]) {
  if constexpr (SCHEME == 0) {
    const double base = floor(g);
    const long i0 = static_cast<long>(base);
    const double t = g - base;
    idx[0] = wrap_index(i0, n);
    idx[1] = wrap_index(i0 + 1, n);
    w[0] = 1.0 - t;
    w[1] = t;
  } else {
    const double c = floor(g + 0.5);
    const long ic = static_cast<long>(c);
    const double d = g - c;
    idx[0] = wrap_index(ic - 1, n);
    idx[1] = wrap_index(ic, n);
    idx[2] = wrap_index(ic + 1, n);
    w[0] = 0.5 * (0.5 - d) * (0.5 - d);
    w[1] = 0.75 - d * d;
    w[2] = 0.5 * (0.5 + d) * (0.5 + d);
  }
}

template <int SCHEME>
__device__ void interp64(const WalkParams& P, const double* __restrict__ fx, const double* __restrict__ fy,
                         const double* __restrict__ fz, double x, double y, double z, double& gx, double& gy,
                         double& gz) {
  constexpr int K = SCHEME + 2;
  int ix[K], iy[K], iz[K];
  double wx[K], wy[K], wz[K];
  stencil64<SCHEME>(wrap_position(x, P.L[0]) * (static_cast<double>(P.mesh[0]) / P.L[0]), P.mesh[0], ix, wx);
  stencil64<SCHEME>(wrap_position(y, P.L[1]) * (static_cast<double>(P.mesh[1]) / P.L[1]), P.mesh[1], iy, wy);
  stencil64<SCHEME>(wrap_position(z, P.L[2]) * (static_cast<double>(P.mesh[2]) / P.L[2]), P.mesh[2], iz, wz);
  gx = gy = gz = 0.0;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      const std::size_t row = (static_cast<std::size_t>(ix[a]) * P.mesh[1] + iy[b]) * P.mesh_row;
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const double w = wx[a] * wy[b] * wz[c];
        const std::size_t i = row + iz[c];
        gx += w * fx[i];
        gy += w * fy[i];
        gz += w * fz[i];
      }
    }
}

// Stencil weights and their derivatives with respect to the mesh coordinate g
// (multiply by n/L for a length derivative).
template <int SCHEME>
__device__ __forceinline__ void stencil64_grad(double g, int n, int (&idx)[SCHEME + 2], double (&w)[SCHEME + 2],
                                               double (&dw)[SCHEME + 2]) {
  stencil64<SCHEME>(g, n, idx, w);
  if constexpr (SCHEME == 0) {
    dw[0] = -1.0;
    dw[1] = 1.0;
  } else {
    const double d = g - floor(g + 0.5);
    dw[0] = -(0.5 - d);
    dw[1] = -2.0 * d;
    dw[2] = 0.5 + d;
  }
}

// Transpose of the PM force's Jacobian for one source k (FP64):
//   s_k = grad_{x_k} sum_c W(x_k, c) [lam_k . F(c) + m_k nu(c)]
// with F the force meshes (interpolation's own dependence on x_k) and nu the
// adjoint potential sum_a S_a^T mu_a of the deposited lam (deposit's
// dependence on x_k).
template <int SCHEME>
__device__ void vjp_gather64(const WalkParams& P, const double* __restrict__ fx, const double* __restrict__ fy,
                             const double* __restrict__ fz, const double* __restrict__ nu, double lx, double ly,
                             double lz, double mk, double x, double y, double z, double& sx, double& sy,
                             double& sz) {
  constexpr int K = SCHEME + 2;
  int ix[K], iy[K], iz[K];
  double wx[K], wy[K], wz[K], dx[K], dy[K], dz[K];
  const double hx = static_cast<double>(P.mesh[0]) / P.L[0];
  const double hy = static_cast<double>(P.mesh[1]) / P.L[1];
  const double hz = static_cast<double>(P.mesh[2]) / P.L[2];
  stencil64_grad<SCHEME>(wrap_position(x, P.L[0]) * hx, P.mesh[0], ix, wx, dx);
  stencil64_grad<SCHEME>(wrap_position(y, P.L[1]) * hy, P.mesh[1], iy, wy, dy);
  stencil64_grad<SCHEME>(wrap_position(z, P.L[2]) * hz, P.mesh[2], iz, wz, dz);
  sx = sy = sz = 0.0;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      const std::size_t row = (static_cast<std::size_t>(ix[a]) * P.mesh[1] + iy[b]) * P.mesh_row;
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const std::size_t i = row + iz[c];
        const double u = lx * fx[i] + ly * fy[i] + lz * fz[i] + mk * nu[i];
        sx += dx[a] * wy[b] * wz[c] * u;
        sy += wx[a] * dy[b] * wz[c] * u;
        sz += wx[a] * wy[b] * dz[c] * u;
      }
    }
  sx *= hx;
  sy *= hy;
  sz *= hz;
}

// Stencil from a 32-bit fixed-point coordinate q (x = q L / 2^32) on a mesh of
// n cells: q n / 2^32 = cell + t exactly in 64-bit integers (for n = 2^lg the
// cell is the top lg bits and t the remaining bits).
__device__ __forceinline__ int wrap_once(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <int SCHEME>
__device__ __forceinline__ void stencil32(uint32_t q, int n, int lg, int (&idx)[SCHEME + 2], float (&w)[SCHEME + 2]) {
  (void)lg;
  const uint64_t p = static_cast<uint64_t>(q) * static_cast<uint32_t>(n);
  const int cell = static_cast<int>(p >> 32);
  const float t = static_cast<float>(static_cast<uint32_t>(p)) * 0x1p-32f;
  if constexpr (SCHEME == 0) {
    idx[0] = cell;
    idx[1] = wrap_once(cell + 1, n);
    w[0] = 1.0f - t;
    w[1] = t;
  } else {
    const int c = t >= 0.5f ? cell + 1 : cell;
    const float d = t >= 0.5f ? t - 1.0f : t;
    idx[0] = wrap_once(c - 1, n);
    idx[1] = wrap_once(c, n);
    idx[2] = wrap_once(c + 1, n);
    w[0] = 0.5f * (0.5f - d) * (0.5f - d);
    w[1] = 0.75f - d * d;
    w[2] = 0.5f * (0.5f + d) * (0.5f + d);
  }
}

template <int SCHEME>
__device__ void interp32(const WalkParams& P, const float* __restrict__ fx, const float* __restrict__ fy,
                         const float* __restrict__ fz, uint32_t qx, uint32_t qy, uint32_t qz, float& gx, float& gy,
                         float& gz) {
  constexpr int K = SCHEME + 2;
  int ix[K], iy[K], iz[K];
  float wx[K], wy[K], wz[K];
  stencil32<SCHEME>(qx, P.mesh[0], P.mesh_log2[0], ix, wx);
  stencil32<SCHEME>(qy, P.mesh[1], P.mesh_log2[1], iy, wy);
  stencil32<SCHEME>(qz, P.mesh[2], P.mesh_log2[2], iz, wz);
  gx = gy = gz = 0.0f;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      const std::size_t row = (static_cast<std::size_t>(ix[a]) * P.mesh[1] + iy[b]) * P.mesh_row;
      const float wab = wx[a] * wy[b];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const float w = wab * wz[c];
        const std::size_t i = row + iz[c];
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

// ---- mixed-precision pair pieces -----------------------------------------------
__device__ __forceinline__ float fx_delta(uint32_t a, uint32_t b, float scale) {
  return static_cast<float>(static_cast<int32_t>(a - b)) * scale;
}

__device__ __forceinline__ void node_accel32(const NodeMom32& M, const NodeQuad32& Q, float dx, float dy, float dz,
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
    const float d52 = soft / denom;  // denom^-5/2
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
  ax = cx;
  ay = cy;
  az = cz;
}

// Node-test units of the mixed walk: per-axis scale of fixed-point
// differences (1 for a cubic box, whose tests run in fixed-point units),
// separation unit -> length and back, the containment radius^2 and theta^2.
struct MixFrame {
  float s0, s1, s2, to_len, inv_len, contain2, theta2;
};

// Strict acceptance of internal node n (tree_pm_acceptance.hpp:54-66: the
// node's cell inside r_cut, the opening criterion and 2h/r < 0.08 with
// quadrupoles) for the target at fixed point (px, py, pz), with the COM
// separation in lengths left in mf* for the multipole evaluation.
__device__ __forceinline__ bool strict_accept32(const WalkParams& P, const MixFrame& F, const NodeGeo32& g,
                                                const NodeMom32& M, const NodeQuad32* __restrict__ node_quad,
                                                uint32_t n, uint32_t px, uint32_t py, uint32_t pz, float& mfx,
                                                float& mfy, float& mfz) {
  const float h = g.half * F.inv_len;
  const float ax_ = fabsf(fx_delta(g.cx, px, F.s0)), ay_ = fabsf(fx_delta(g.cy, py, F.s1)),
              az_ = fabsf(fx_delta(g.cz, pz, F.s2));
  const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
  const float mx = ax_ + h, my = ay_ + h, mz = az_ + h;
  const bool contained = fmaf(mx, mx, fmaf(my, my, mz * mz)) <= F.contain2;
  mfx = fx_delta(M.comx, px, F.s0) * F.to_len;
  mfy = fx_delta(M.comy, py, F.s1) * F.to_len;
  mfz = fx_delta(M.comz, pz, F.s2) * F.to_len;
  const float r2 = mfx * mfx + mfy * mfy + mfz * mfz;
  const float w2 = 2.0f * g.half;
  // 2h/r < 0.08 and (2h [+ |c - COM|]) / r < theta, squared (no sqrt/div)
  const bool env_ok = P.quad && w2 * w2 < 0.0064f * fmaxf(r2, 1.0e-30f);
  const float lead = P.criterion == 0 ? w2 : w2 + node_quad[n].offset;
  const bool mac = !inside && lead * lead < F.theta2 * (r2 + 1.0e-30f);
  return mac && env_ok && contained;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

// ---- the walk -------------------------------------------------------------------
// Per node, every lane whose next node it is applies the reference's cutoff
// prune; whether the node is a leaf is wave-uniform, so leaf and internal-node
// work are scalar branches. While the wave streams a leaf's sources (wave-
// uniform scalar loads of packed records) each lane only forms the separation,
// applies the cutoff and self test and appends the source index to its own LDS
// queue; the residual kernel then runs over the queues, so lanes spend it only
// on pairs inside r_cut. A queue is drained before it can overflow and before
// any lane adds an accepted multipole, so each lane's contributions are summed
// in the order of the reference's sequential walk. The next node is i + 1
// whenever any lane is waiting on it (one ballot), otherwise the DPP wave
// minimum of the lanes' next nodes.
constexpr int kWalkWaves = kWalkThreads / kWave;
constexpr int kQueue = 32;
constexpr int kChunk = 16;

struct WalkData {
  const NodeLink* link;
  const NodeGeo64* geo64;
  const NodeMom64* mom64;
  const NodeQuad64* quad64;
  const NodeGeo32* geo32;
  const NodeMom32* mom32;
  const NodeQuad32* quad32;
  const NodeBox32* box32;
  const uint32_t* order;
  const double4* src64;
  const uint4* src32;
  const float4* htab;
  int uniform_mass;
  const uint32_t *tpos, *tslot;
  const double *f64x, *f64y, *f64z;
  const float *f32x, *f32y, *f32z;
  // force VJP: tree-ordered lam (x, y, z, unused) and the adjoint potential
  const double4* lam64;
  const double* nu64;
  // adaptive policy: Gaussian Hermite knots and |a_prev| per output slot (nullable)
  const double2* gtab;
  const double* prev;
  // group walk: the same Gaussian cells as monomial coefficients (a0..a3)
  const double4* gcoef;
  // FP64 box walk: degree-5 cells of H (a0..a5 as three double2 per cell)
  const double2* h64tab;
  // mixed timed walk: residual cells indexed by the float bits of r^2 + eps^2
  const float4* ltab;
  // preorder index of each node's child per octant (kInvalid for none)
  const uint32_t* child;
  // force VJP with accepted nodes: per preorder node U (3) and V (7)
  // accumulators, and the (target tree position << 32 | node) list
  double* vjp_node_acc;
  unsigned long long* vjp_accepted;
  unsigned int* vjp_accepted_count;
  unsigned int vjp_accepted_cap;
  // after a capture-safe build the node count is on the device (P.node_count
  // is then the capacity); the mixed walks bound their traversal with it
  const uint32_t* node_count_dev;
};

// splitmix64 finaliser of the pair (target row, source row): the pair-set
// checksum adds it over every direct pair inside r_cut.
__device__ __forceinline__ unsigned long long pair_mix(uint32_t target_row, uint32_t source_row) {
  unsigned long long z = (static_cast<unsigned long long>(target_row) << 32) | source_row;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t next_node(uint32_t n, uint32_t nxt) {
  return __ballot(nxt == n + 1) ? n + 1 : wave_min_u32(nxt);
}

// ---- FP64 walk (reference arithmetic) ---------------------------------------------
// JVP: instead of the acceleration, each target k accumulates the transpose of
// the acceleration Jacobian applied to lam (see vjp_gather64 for the PM part).
// For the residual pairs, with r = x_j - x_k, t = r^2 and the pair force
// G m_j F(t) r, the contribution is G (F w + 2 F'(t) (r . w) r) with
// w = m_k lam_j - m_j lam_k; the sum runs over the same interaction lists.
// ADAPT: the adaptive acceptance policy with the target's acceleration history,
// the cutoff node policy, optional Gaussian-table leaf pairs and the optional
// radial-coefficient node expansion (tree_pm_coupling.cpp:4026-4272 with
// TreePmAcceptancePolicy::kAdaptiveRelative).
template <int SCHEME, bool JVP, bool ADAPT = false>
__global__ __launch_bounds__(kWalkThreads) void walk64_kernel(WalkParams P, WalkData D, WalkOutputs out,
                                                               unsigned long long* __restrict__ counters) {
  __shared__ uint32_t queue[kWalkWaves][kQueue][kWave];
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const uint32_t k = xcd_logical_block() * kWalkThreads + threadIdx.x;
  const bool valid = k < P.n_targets;
  const uint32_t tp = valid ? (P.identity_targets ? k : D.tpos[k]) : 0u;
  const uint32_t slot = valid ? (P.identity_targets ? D.order[k] : D.tslot[k]) : 0u;
  uint32_t cnt[kWalkCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t qlen = 0, queued = 0;
  uint32_t nxt = valid ? 0u : P.node_count;
  double nearest = INFINITY;  // r^2 of the nearest distinct direct-summed source inside r_cut
  const uint32_t target_row = valid ? D.order[tp] : 0u;
  unsigned long long pair_sum = 0;  // pair-set checksum (wrapping)
  double px = 0.0, py = 0.0, pz = 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  double4 lk_ = make_double4(0.0, 0.0, 0.0, 0.0);  // JVP: lam of the target
  double mk = 0.0;                                  // JVP: mass of the target
  if (valid) {
    const double4 t = D.src64[tp];
    px = t.x;
    py = t.y;
    pz = t.z;
    if constexpr (JVP) {
      lk_ = D.lam64[tp];
      mk = D.uniform_mass ? P.uniform_mass : t.w;
      if (P.pm_enabled)
        vjp_gather64<SCHEME>(P, D.f64x, D.f64y, D.f64z, D.nu64, lk_.x, lk_.y, lk_.z, mk, px, py, pz, gx, gy, gz);
    } else {
      if (P.pm_enabled) interp64<SCHEME>(P, D.f64x, D.f64y, D.f64z, px, py, pz, gx, gy, gz);
    }
  }
  double ax = 0.0, ay = 0.0, az = 0.0;
  const double L0 = P.L[0], L1 = P.L[1], L2 = P.L[2];
  bool history = false;
  double prev = 0.0;
  if constexpr (ADAPT) {
    if (valid && D.prev != nullptr) {
      prev = D.prev[slot];
      history = isfinite(prev);
      if (!history) prev = 0.0;
    }
  }

  auto drain = [&]() {
    const uint32_t m = wave_max_u32(qlen);
    for (uint32_t q = 0; q < m; ++q) {
      if (q < qlen) {
        const uint32_t j = queue[w][q][lane];
        const double4 sj = D.src64[j];
        const double qx = mi64(sj.x - px, L0), qy = mi64(sj.y - py, L1), qz = mi64(sj.z - pz, L2);
        const double r2 = qx * qx + qy * qy + qz * qz;
        const double mj = D.uniform_mass ? P.uniform_mass : sj.w;
        double f;
        if constexpr (ADAPT) {
          f = (P.lookup ? table_residual64(D.gtab, r2, P) : residual_inv_r3(r2, P.eps, P.a)) * P.G;
        } else {
          f = residual_inv_r3(r2, P.eps, P.a) * P.G;
        }
        if constexpr (JVP) {
          const double4 lj = D.lam64[j];
          const double wx = mk * lj.x - mj * lk_.x, wy = mk * lj.y - mj * lk_.y, wz = mk * lj.z - mj * lk_.z;
          const double c = 2.0 * residual_inv_r3_dt(r2, P.eps, P.a) * P.G * (qx * wx + qy * wy + qz * wz);
          ax += f * wx + c * qx;
          ay += f * wy + c * qy;
          az += f * wz + c * qz;
        } else {
          ax += f * mj * qx;
          ay += f * mj * qy;
          az += f * mj * qz;
        }
      }
    }
    queued += qlen;
    qlen = 0;
  };

  uint32_t n = wave_min_u32(nxt);
  while (n < P.node_count) {
    if (lane == 0) ++cnt[5];
    const NodeLink lk = D.link[n];
    const NodeGeo64 g = D.geo64[n];
    const bool act = nxt == n;
    if (!__ballot(act)) {  // the chosen node must be some lane's next node
      if (lane == 0) atomicOr(&counters[kFaultSlot], 1ull);
      break;
    }
    cnt[0] += act;
    const double h = g.half;
    const double ndx = mi64(g.cx - px, L0), ndy = mi64(g.cy - py, L1), ndz = mi64(g.cz - pz, L2);
    const double ax_ = fabs(ndx), ay_ = fabs(ndy), az_ = fabs(ndz);
    const double ex = fmax(0.0, ax_ - h), ey = fmax(0.0, ay_ - h), ez = fmax(0.0, az_ - h);
    const bool keep = act && within_cutoff(ex * ex + ey * ey + ez * ez, P.rcut, P.rcut2);
    if (act && !keep) {
      ++cnt[1];
      nxt = n + lk.skip;
    }
    if (lk.child_count == 0) {
      if (keep) {
        ++cnt[4];
        nxt = n + 1;
      }
      if (__ballot(keep)) {
        const uint32_t e0 = lk.begin + lk.count;
        for (uint32_t c0 = lk.begin; c0 < e0; c0 += kChunk) {
          const uint32_t c1 = c0 + kChunk < e0 ? c0 + kChunk : e0;
          if (lane == 0) cnt[6] += c1 - c0;
          for (uint32_t j = c0; j < c1; ++j) {
            const double4 sj = D.src64[j];
            const double qx = mi64(sj.x - px, L0), qy = mi64(sj.y - py, L1), qz = mi64(sj.z - pz, L2);
            const double r2 = qx * qx + qy * qy + qz * qz;
            if (keep && j != tp && r2 <= P.rcut2) {
              nearest = fmin(nearest, r2);
              pair_sum += pair_mix(target_row, D.order[j]);
              // a coincident pair has a zero force but, in the JVP, the
              // Jacobian F(0) I
              if (!JVP && qx == 0.0 && qy == 0.0 && qz == 0.0) {
                ++cnt[2];  // counted coincident pair, zero vector
              } else {
                queue[w][qlen][lane] = j;
                ++qlen;
              }
            }
          }
          if (wave_max_u32(qlen) > static_cast<uint32_t>(kQueue - kChunk)) drain();
        }
      }
    } else if (__ballot(keep)) {
      bool take_mp = false;
      double mdx = 0.0, mdy = 0.0, mdz = 0.0;
      if (keep) {
        const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
        const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
        if constexpr (ADAPT) {
          const bool contained = P.com_cut || within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
          bool cut = false;
          if (!inside && contained) {
            const NodeMom64 M = D.mom64[n];
            mdx = mi64(M.comx - px, L0);
            mdy = mi64(M.comy - py, L1);
            mdz = mi64(M.comz - pz, L2);
            const double r2 = mdx * mdx + mdy * mdy + mdz * mdz;
            take_mp = accept_adaptive64(h, D.quad64[n].offset, M.mass, r2, history, prev, P);
            // kComDistance: an accepted node whose COM lies beyond r_cut is pruned
            if (take_mp && P.com_cut && r2 > P.rcut2) {
              take_mp = false;
              cut = true;
              ++cnt[1];
            }
          }
          nxt = (take_mp || cut) ? n + lk.skip : n + 1;
        } else {
          const bool contained = within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
          const NodeMom64 M = D.mom64[n];
          mdx = mi64(M.comx - px, L0);
          mdy = mi64(M.comy - py, L1);
          mdz = mi64(M.comz - pz, L2);
          const double r2 = mdx * mdx + mdy * mdy + mdz * mdz;
          take_mp = accept64(h, D.quad64[n].offset, inside, contained, r2, P);
          nxt = take_mp ? n + lk.skip : n + 1;
        }
      }
      if (__ballot(take_mp)) {
        drain();
        if (take_mp) {
          ++cnt[3];
          if (ADAPT && P.radial) node_accel_radial64(D.mom64[n], D.quad64[n], mdx, mdy, mdz, P, ax, ay, az);
          else node_accel64(D.mom64[n], D.quad64[n], mdx, mdy, mdz, P, ax, ay, az);
        }
      }
    }
    n = next_node(n, nxt);
  }
  drain();
  cnt[2] += queued;
  flush_counters(counters, cnt);
  {
    const unsigned long long v = wave_sum(pair_sum);
    if (lane == 0 && v) atomicAdd(&counters[kPairChecksumSlot], v);
  }
  if (!valid) return;
  static_cast<double*>(out.ax)[slot] = gx + ax;
  static_cast<double*>(out.ay)[slot] = gy + ay;
  static_cast<double*>(out.az)[slot] = gz + az;
  if (out.pm_x) {
    static_cast<double*>(out.pm_x)[slot] = gx;
    static_cast<double*>(out.pm_y)[slot] = gy;
    static_cast<double*>(out.pm_z)[slot] = gz;
  }
  const uint32_t row = D.order[tp];
  if (out.sr_x) {
    static_cast<double*>(out.sr_x)[row] = ax;
    static_cast<double*>(out.sr_y)[row] = ay;
    static_cast<double*>(out.sr_z)[row] = az;
  }
  if (out.near2) static_cast<double*>(out.near2)[row] = nearest;
}

// ---- group walk (FP64) -----------------------------------------------------------
// One wave per residual block of visit-ordered targets (tree_pm_coupling.cpp:
// 4451-4546). The wave splits its block into target groups (the targets of one
// leaf, or up to group_size consecutive targets whose box stays within
// max_extent per axis), then per group:
//   1. checks every target's history and bounded metric; a failing group is
//      appended to the per-target fallback list (tree_pm_coupling.cpp:4341-4370);
//   2. walks the tree once for the group box (buildTreePmGroupInteractionList,
//      tree_pm_group_walk.hpp:78-163). The walk is the stackless preorder walk
//      of the per-target kernel made wave-cooperative: the 64 lanes decide 64
//      consecutive preorder nodes at once, and the nodes actually visited are
//      the chain p -> next(p) -> ... through them (next = q + 1 to open, q + skip
//      otherwise). Listed leaves and accepted nodes are appended in preorder to
//      an LDS buffer;
//   3. whenever the buffer fills (and at the end) every target evaluates the
//      buffered list (accumulateTreePmGroupCulledLeaves, :290-341, and the
//      accepted-node expansion): a group of T targets gives each target S lanes
//      (the largest power of two with S T <= 64). The listed leaves' sources
//      are staged in LDS 64 at a time, each lane taking every S-th staged
//      source and every S-th accepted node; the lanes' sums are added at the
//      end. The per-target leaf cull is evaluated for its counters only: a
//      culled leaf lies beyond r_cut of the target, so its pairs are zero, and
//      evaluating every listed source keeps all S lanes of a target busy.
// Groups above 64 targets are evaluated 64 targets at a time, repeating the
// (identical) tree walk for each.
//
// Mixed precision (MIX, group_walkmix_kernel): groups, histories, the tree
// walk and every acceptance decision stay in FP64, so the lists equal the FP64
// walk's; the accepted nodes' expansions stay in FP64 as well (a few per
// target). Only the listed-source pairs change: sources are staged as 32-bit
// fixed-point coordinates (16 B instead of 32 B), a separation is the wrapped
// int32 difference scaled to H-table units, and the pair term is the FP32
// residual of the mixed strict walk (reciprocal square root, the 128-cell
// cubic table of H in LDS). Each staged chunk's FP32 sums are added to the
// FP64 totals, so no FP32 sum runs over more than 64 sources per lane. PM
// interpolation reads the FP32 force meshes at the fixed-point position.
constexpr int kGroupLeafCap = 192;
constexpr int kGroupNodeCap = 128;
constexpr int kGroupStage = 64;  // listed sources staged in LDS per evaluation pass

struct GroupLeaf {
  double ox, oy, oz, h;  // cell centre relative to the group centre, cell half size
  uint32_t b, e;         // tree-order source range
};

__device__ __forceinline__ double nearest_image(double d, double L) { return d - L * rint(d / L); }
// The same with 1/L precomputed: rint(d * iL) equals rint(d / L) unless d / L
// lies within an ulp of a half-integer, where both images are equally near.
__device__ __forceinline__ double nearest_image_inv(double d, double L, double iL) {
  return d - L * rint(d * iL);
}
__device__ __forceinline__ double wrap_half(double d, double L) {
  const double hl = 0.5 * L;
  return d > hl ? d - L : (d < -hl ? d + L : d);
}
template <class T> __device__ __forceinline__ T scan_min(T v, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const T o = __shfl_up(v, off, kWave);
    if (lane >= off) v = o < v ? o : v;
  }
  return v;
}
template <class T> __device__ __forceinline__ T scan_max(T v, int lane) {
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const T o = __shfl_up(v, off, kWave);
    if (lane >= off) v = o > v ? o : v;
  }
  return v;
}

template <int SCHEME, bool MIX, bool NODE32>
__device__ __forceinline__ void group_walk_body(const WalkParams& P, const GroupParams& G, const WalkData& D,
                                                const uint32_t* __restrict__ leaf_of_pos, const WalkOutputs& out,
                                                uint32_t* __restrict__ fb_pos, uint32_t* __restrict__ fb_slot,
                                                uint32_t* __restrict__ fb_count,
                                                unsigned long long* __restrict__ counters) {
  __shared__ GroupLeaf leaves[kGroupLeafCap];
  __shared__ uint32_t nodes[kGroupNodeCap];
  __shared__ uint32_t leaf_pre[kGroupLeafCap + 1];
  __shared__ double4 stage[MIX ? 1 : kGroupStage];
  __shared__ uint4 stage32[MIX ? kGroupStage : 1];
  __shared__ float4 htab[MIX ? kHTableCells : 1];
  __shared__ uint32_t stage_j[kGroupStage];
  const int lane = threadIdx.x;
  if constexpr (MIX) {
    for (int i = lane; i < kHTableCells; i += kWave) htab[i] = D.htab[i];
    __syncthreads();
  }
  const uint32_t bb = blockIdx.x * G.block;
  const uint32_t be = min(bb + G.block, P.n_targets);
  const double L0 = P.L[0], L1 = P.L[1], L2 = P.L[2];
  const double cull2 = P.rcut2 * (1.0 + 1.0e-9);
  const double tlim = 0x1.fffffffffffffp+5;  // largest double below 64
  const double iL0 = 1.0 / P.L[0], iL1 = 1.0 / P.L[1], iL2 = 1.0 / P.L[2];
  const double eps2 = P.eps * P.eps;
  uint32_t cnt[kGroupCounters] = {};
  unsigned long long size_max = 0, entries_max = 0;
  auto tpos = [&](uint32_t v) { return P.identity_targets ? v : D.tpos[v]; };
  auto tslot = [&](uint32_t v) { return P.identity_targets ? D.order[v] : D.tslot[v]; };

  uint32_t v = bb;
  while (v < be) {
    const double4 p0 = D.src64[tpos(v)];
    // ---- the group [v, ge) (tree_pm_coupling.cpp:4491-4523)
    uint32_t ge;
    if (G.group_size == 0) {
      const uint32_t leaf0 = leaf_of_pos[tpos(v)];
      ge = be;
      for (uint32_t c = v + 1; c < be; c += kWave) {
        const uint32_t k = c + lane;
        const unsigned long long m = __ballot(k < be && leaf_of_pos[tpos(k)] != leaf0);
        if (m) {
          ge = c + static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
          break;
        }
      }
    } else {
      const uint32_t lim = min(be, v + G.group_size);
      double lo0 = 0.0, lo1 = 0.0, lo2 = 0.0, hi0 = 0.0, hi1 = 0.0, hi2 = 0.0;
      ge = lim;
      for (uint32_t c = v + 1; c < lim; c += kWave) {
        const uint32_t k = c + lane;
        const bool in = k < lim;
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;  // 0 is neutral: every running box holds 0
        if (in) {
          const double4 q = D.src64[tpos(k)];
          d0 = nearest_image(q.x - p0.x, L0);
          d1 = nearest_image(q.y - p0.y, L1);
          d2 = nearest_image(q.z - p0.z, L2);
        }
        const double a0 = scan_min(fmin(lo0, d0), lane), a1 = scan_min(fmin(lo1, d1), lane),
                     a2 = scan_min(fmin(lo2, d2), lane);
        const double b0 = scan_max(fmax(hi0, d0), lane), b1 = scan_max(fmax(hi1, d1), lane),
                     b2 = scan_max(fmax(hi2, d2), lane);
        const bool within = b0 - a0 <= G.max_extent && b1 - a1 <= G.max_extent && b2 - a2 <= G.max_extent;
        const unsigned long long bad = __ballot(in && !within);
        if (bad) {
          ge = c + static_cast<uint32_t>(__ffsll(static_cast<long long>(bad)) - 1);
          break;
        }
        lo0 = readlane_f64(a0, 63); lo1 = readlane_f64(a1, 63); lo2 = readlane_f64(a2, 63);
        hi0 = readlane_f64(b0, 63); hi1 = readlane_f64(b1, 63); hi2 = readlane_f64(b2, 63);
      }
    }
    const uint32_t count = ge - v;

    // ---- history / bounded metric, minimum |a_prev| and the group box (:4333-4370)
    int fall = 0;  // 1 history, 2 geometry (first failing target decides)
    double amin = INFINITY, lo0 = 0.0, lo1 = 0.0, lo2 = 0.0, hi0 = 0.0, hi1 = 0.0, hi2 = 0.0;
    for (uint32_t c = v; c < ge; c += kWave) {
      const uint32_t k = c + lane;
      const bool in = k < ge;
      double pv = INFINITY, d0 = 0.0, d1 = 0.0, d2 = 0.0;
      bool hist = true, bounded = true;
      if (in) {
        pv = D.prev != nullptr ? D.prev[tslot(k)] : NAN;
        hist = isfinite(pv) && pv > P.accel_floor;
        const double4 q = D.src64[tpos(k)];
        bounded = G.lo[0] - q.x >= -L0 && G.hi[0] - q.x <= L0 && G.lo[1] - q.y >= -L1 && G.hi[1] - q.y <= L1 &&
                  G.lo[2] - q.z >= -L2 && G.hi[2] - q.z <= L2;
        d0 = nearest_image(q.x - p0.x, L0);
        d1 = nearest_image(q.y - p0.y, L1);
        d2 = nearest_image(q.z - p0.z, L2);
      }
      const unsigned long long bad = __ballot(in && !(hist && bounded));
      if (bad) {
        const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<long long>(bad)) - 1);
        fall = readlane_u32(hist ? 2u : 1u, l);
        break;
      }
      amin = fmin(amin, wave_min(in ? pv : INFINITY));
      lo0 = fmin(lo0, wave_min(d0)); lo1 = fmin(lo1, wave_min(d1)); lo2 = fmin(lo2, wave_min(d2));
      hi0 = fmax(hi0, wave_max(d0)); hi1 = fmax(hi1, wave_max(d1)); hi2 = fmax(hi2, wave_max(d2));
    }
    if (fall) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(fb_count, count);
      base = readlane_u32(base, 0);
      for (uint32_t c = v; c < ge; c += kWave) {
        const uint32_t k = c + lane;
        if (k < ge) {
          fb_pos[base + (k - v)] = tpos(k);
          fb_slot[base + (k - v)] = tslot(k);
        }
      }
      if (lane == 0) {
        cnt[8] += count;
        if (fall == 1) cnt[9] += count;
        else cnt[10] += count;
      }
      v = ge;
      continue;
    }
    // wave-uniform group box and minimum |a_prev| (scalar registers)
    amin = readlane_f64(amin, 0);
    const double c0x = readlane_f64(p0.x + 0.5 * (lo0 + hi0), 0), c0y = readlane_f64(p0.y + 0.5 * (lo1 + hi1), 0),
                 c0z = readlane_f64(p0.z + 0.5 * (lo2 + hi2), 0);
    const double hg0 = readlane_f64(0.5 * (hi0 - lo0), 0), hg1 = readlane_f64(0.5 * (hi1 - lo1), 0),
                 hg2 = readlane_f64(0.5 * (hi2 - lo2), 0);
    if (lane == 0) {
      cnt[5] += 1;
      cnt[6] += count;
    }
    size_max = count > size_max ? count : size_max;

    // ---- targets, 64 at a time
    for (uint32_t t0 = v; t0 < ge; t0 += kWave) {
      const uint32_t T = min(ge - t0, static_cast<uint32_t>(kWave));
      int S = 1;  // lanes per target: the largest power of two with S * T <= 64
      while (S < kWave && static_cast<uint32_t>(2 * S) * T <= static_cast<uint32_t>(kWave)) S *= 2;
      const int TT = kWave / S;
      const int t = lane & (TT - 1), s = lane / TT;
      const bool valid = static_cast<uint32_t>(t) < T;
      const bool first = t0 == v;  // count the (repeated) tree walk once per group
      uint32_t tp = 0, slot = 0;
      uint32_t tx = 0, ty = 0, tz = 0;  // MIX: fixed-point target position
      double px = 0.0, py = 0.0, pz = 0.0, ox = 0.0, oy = 0.0, oz = 0.0;
      if (valid) {
        tp = tpos(t0 + t);
        slot = tslot(t0 + t);
        if constexpr (MIX) {
          const uint4 f = D.src32[tp];
          tx = f.x;
          ty = f.y;
          tz = f.z;
        }
        const double4 q = D.src64[tp];
        px = q.x;
        py = q.y;
        pz = q.z;
        ox = nearest_image(px - c0x, L0);
        oy = nearest_image(py - c0y, L1);
        oz = nearest_image(pz - c0z, L2);
      }
      double ax = 0.0, ay = 0.0, az = 0.0, nearest = INFINITY;
      uint32_t listed = 0, nl = 0, nm = 0, p = 0;

      auto evaluate = [&]() {
        __syncthreads();
        // candidate and culled leaves per target (the counters of
        // accumulateTreePmGroupCulledLeaves): a culled leaf lies beyond r_cut
        // of the target, so its sources add nothing to the pair sums below
        for (uint32_t e = s; e < nl; e += S) {
          if (!valid) continue;
          const GroupLeaf Lf = leaves[e];
          const double gx = fmax(0.0, fabs(wrap_half(Lf.ox - ox, L0)) - Lf.h);
          const double gy = fmax(0.0, fabs(wrap_half(Lf.oy - oy, L1)) - Lf.h);
          const double gz = fmax(0.0, fabs(wrap_half(Lf.oz - oz, L2)) - Lf.h);
          const bool cull = gx * gx + gy * gy + gz * gz > cull2;
          ++cnt[11];
          if (cull) ++cnt[12];
          else cnt[13] += Lf.e - Lf.b;
        }
        // exclusive prefix of the listed leaves' source counts
        uint32_t total = 0;
        for (uint32_t c = 0; c < nl; c += kWave) {
          const uint32_t e = c + lane;
          const uint32_t own = e < nl ? leaves[e].e - leaves[e].b : 0u;
          uint32_t incl = own;
          for (int off = 1; off < kWave; off <<= 1) {
            const uint32_t o = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += o;
          }
          if (e < nl) leaf_pre[e] = total + incl - own;
          total += readlane_u32(incl, kWave - 1);
        }
        if (lane == 0) leaf_pre[nl] = total;
        // the listed sources, kGroupStage at a time: staged once in LDS and
        // evaluated by each target's S lanes (lane s takes every S-th source)
        for (uint32_t g0 = 0; g0 < total; g0 += kGroupStage) {
          const uint32_t gn = min(total - g0, static_cast<uint32_t>(kGroupStage));
          __syncthreads();
          if (static_cast<uint32_t>(lane) < gn) {
            const uint32_t g = g0 + lane;
            uint32_t lo = 0, hi = nl;  // the leaf e with leaf_pre[e] <= g < leaf_pre[e + 1]
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (leaf_pre[mid] <= g) lo = mid;
              else hi = mid;
            }
            const uint32_t j = leaves[lo].b + (g - leaf_pre[lo]);
            if constexpr (MIX) stage32[lane] = D.src32[j];
            else stage[lane] = D.src64[j];
            stage_j[lane] = j;
          }
          __syncthreads();
          if (!valid) continue;
          if constexpr (MIX) {
            // FP32 pairs in H-table units (the mixed strict walk's pair term)
            const float u0 = P.tab_scale[0], u1 = P.tab_scale[1], u2 = P.tab_scale[2];
            const float rc2 = P.rcut2_tab, e2 = P.eps2_tab, c8 = P.c8_tab;
            float lx = 0.0f, ly = 0.0f, lz = 0.0f, near_t = INFINITY;
            uint32_t inside_n = 0;
            for (uint32_t q = s; q < gn; q += S) {
              const uint4 sv = stage32[q];
              const uint32_t j = stage_j[q];
              const float dx = fx_delta(sv.x, tx, u0), dy = fx_delta(sv.y, ty, u1), dz = fx_delta(sv.z, tz, u2);
              const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
              const float rinv = __builtin_amdgcn_rsqf(r2 + e2);
              const float soft = rinv * rinv * rinv;
              const int ci = static_cast<int>(fminf(r2, static_cast<float>(kHTableCells - 1)));
              const float fu = __builtin_amdgcn_fractf(r2);
              const float4 c = htab[ci];
              const float hcoef = fmaf(fmaf(fmaf(c.w, fu, c.z), fu, c.y), fu, c.x);
              float fr = fmaf(-hcoef, c8, soft);
              if (!D.uniform_mass) fr *= __uint_as_float(sv.w);
              const bool inside = j != tp && r2 <= rc2;
              const float f = inside ? fr : 0.0f;
              lx = fmaf(f, dx, lx);
              ly = fmaf(f, dy, ly);
              lz = fmaf(f, dz, lz);
              inside_n += inside ? 1u : 0u;
              near_t = inside ? fminf(near_t, r2) : near_t;
            }
            const double gm = P.force_scale;
            ax = fma(gm, static_cast<double>(lx), ax);
            ay = fma(gm, static_cast<double>(ly), ay);
            az = fma(gm, static_cast<double>(lz), az);
            cnt[2] += inside_n;
            nearest = fmin(nearest, static_cast<double>(near_t) * static_cast<double>(P.tab_to_len2));
            continue;
          }
          for (uint32_t q = s; q < gn; q += S) {
            const double4 sj = stage[q];
            const uint32_t j = stage_j[q];
            // minimum image by rounding: the same separations as wrap_half
            // wherever |d| <= L/2 or d - L is exact (every pair inside r_cut)
            const double ex = sj.x - px, ey = sj.y - py, ez = sj.z - pz;
            const double dx = fma(-L0, rint(ex * iL0), ex), dy = fma(-L1, rint(ey * iL1), ey),
                         dz = fma(-L2, rint(ez * iL2), ez);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool inside = j != tp && r2 <= P.rcut2;
            // the Hermite cell as a cubic in u (Horner with fused steps)
            const double cell = fmin(r2 * P.inv2a2, tlim) * 64.0;
            const int ci = static_cast<int>(cell);
            const double u = cell - static_cast<double>(ci);
            const double4 hc = D.gcoef[ci];
            const double lr = fma(fma(fma(hc.w, u, hc.z), u, hc.y), u, hc.x) * P.inv8a3;
            // (r^2 + eps^2)^(-3/2): hardware reciprocal square root and two
            // Newton steps y <- y + y (1 - dd y^2) / 2 (r2 = 0 is masked below)
            const double dd = r2 + eps2;
            double y = __builtin_amdgcn_rsq(dd);
            y = fma(0.5 * y, fma(-dd * y, y, 1.0), y);
            y = fma(0.5 * y, fma(-dd * y, y, 1.0), y);
            const double soft = y * y * y;
            const double mj = D.uniform_mass ? P.uniform_mass : sj.w;
            const double f = inside && r2 > 0.0 ? (soft - lr) * P.G * mj : 0.0;
            ax += f * dx;
            ay += f * dy;
            az += f * dz;
            cnt[2] += inside ? 1u : 0u;
            nearest = inside ? fmin(nearest, r2) : nearest;
          }
        }
        for (uint32_t e = s; e < nm; e += S) {
          if (!valid) continue;
          const uint32_t nd = nodes[e];
          const NodeMom64 M = D.mom64[nd];
          const double dx = mi64(M.comx - px, L0), dy = mi64(M.comy - py, L1), dz = mi64(M.comz - pz, L2);
          if (P.com_cut && dx * dx + dy * dy + dz * dz > P.rcut2) {
            ++cnt[1];
            continue;
          }
          if constexpr (NODE32) {
            // the mixed strict walk's FP32 expansion from the fixed-point
            // centre of mass (the cutoff decision above stays FP64)
            const NodeMom32 M32 = D.mom32[nd];
            const float s0 = P.cubic ? 1.0f : P.fx_to_len[0], s1 = P.cubic ? 1.0f : P.fx_to_len[1],
                        s2 = P.cubic ? 1.0f : P.fx_to_len[2], tl = P.cubic ? P.fx_to_len[0] : 1.0f;
            float cx = 0.0f, cy = 0.0f, cz = 0.0f;
            node_accel32(M32, D.quad32[nd], fx_delta(M32.comx, tx, s0) * tl, fx_delta(M32.comy, ty, s1) * tl,
                         fx_delta(M32.comz, tz, s2) * tl, P, cx, cy, cz);
            ax += cx;
            ay += cy;
            az += cz;
          } else {
            // inline: an out-of-line call here saves and restores the live
            // registers through scratch on every accepted node
            if (P.radial) node_accel_radial64(M, D.quad64[nd], dx, dy, dz, P, ax, ay, az);
            else node_accel64(M, D.quad64[nd], dx, dy, dz, P, ax, ay, az);
          }
        }
        __syncthreads();
        nl = 0;
        nm = 0;
      };

      while (true) {
        const bool done = p >= P.node_count;
        if (done || nl > static_cast<uint32_t>(kGroupLeafCap - kWave) ||
            nm > static_cast<uint32_t>(kGroupNodeCap - kWave))
          evaluate();
        if (done) break;
        // decide preorder nodes p .. p + 63 for the group box
        const uint32_t q = p + lane;
        uint32_t next = q + 1, lb = 0, le = 0;
        int kind = 3;  // 0 pruned, 1 listed leaf, 2 accepted node, 3 opened
        double dc0 = 0.0, dc1 = 0.0, dc2 = 0.0, h = 0.0;
        if (q < P.node_count) {
          const NodeLink lk = D.link[q];
          const NodeGeo64 g = D.geo64[q];
          h = g.half;
          dc0 = nearest_image_inv(g.cx - c0x, L0, iL0);
          dc1 = nearest_image_inv(g.cy - c0y, L1, iL1);
          dc2 = nearest_image_inv(g.cz - c0z, L2, iL2);
          const double v0 = fabs(dc0), v1 = fabs(dc1), v2 = fabs(dc2);
          const double g0 = fmax(0.0, v0 - h - hg0), g1 = fmax(0.0, v1 - h - hg1), g2 = fmax(0.0, v2 - h - hg2);
          const double f0 = v0 + h + hg0, f1 = v1 + h + hg1, f2 = v2 + h + hg2;
          const double gap2 = g0 * g0 + g1 * g1 + g2 * g2;
          const double far2 = f0 * f0 + f1 * f1 + f2 * f2;
          if (gap2 > P.rcut2) {
            kind = 0;
          } else if (lk.child_count == 0) {
            kind = 1;
            lb = lk.begin;
            le = lk.begin + lk.count;
          } else {
            bool acc = gap2 > 0.0 && (!G.contain || far2 <= P.rcut2);
            if (acc) {
              const NodeMom64 M = D.mom64[q];
              const double m0 = fmax(0.0, fabs(nearest_image_inv(M.comx - c0x, L0, iL0)) - hg0);
              const double m1 = fmax(0.0, fabs(nearest_image_inv(M.comy - c0y, L1, iL1)) - hg1);
              const double m2 = fmax(0.0, fabs(nearest_image_inv(M.comz - c0z, L2, iL2)) - hg2);
              const double rmin2 = m0 * m0 + m1 * m1 + m2 * m2;
              const double width2 = 4.0 * h * h;
              acc = rmin2 > 0.0 && width2 < G.amax2 * rmin2 && M.mass * width2 <= G.tolg * amin * rmin2 * rmin2;
            }
            kind = acc ? 2 : 3;
          }
          next = kind == 3 ? q + 1 : q + lk.skip;
        }
        // the visited chain through the window
        unsigned long long chain = 0;
        uint32_t cur = p;
        while (cur < p + kWave && cur < P.node_count) {
          const uint32_t l = cur - p;
          chain |= 1ull << l;
          cur = readlane_u32(next, l);
        }
        p = cur;
        const bool on = (chain >> lane) & 1ull;
        const unsigned long long lm = __ballot(on && kind == 1);
        const unsigned long long mm = __ballot(on && kind == 2);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (on && kind == 1) {
          GroupLeaf& Lf = leaves[nl + __popcll(lm & below)];
          Lf.ox = dc0;
          Lf.oy = dc1;
          Lf.oz = dc2;
          Lf.h = h;
          Lf.b = lb;
          Lf.e = le;
        }
        if (on && kind == 2) nodes[nm + __popcll(mm & below)] = q;
        const unsigned long long pm = __ballot(on && kind == 0);
        nl += __popcll(lm);
        nm += __popcll(mm);
        listed += __popcll(lm) + __popcll(mm);
        if (first && lane == 0) {
          cnt[0] += __popcll(chain);
          cnt[1] += __popcll(pm);
          cnt[3] += __popcll(mm);
          cnt[4] += __popcll(lm);
        }
      }
      entries_max = listed > entries_max ? listed : entries_max;
      // sums of a target's lanes
      for (int off = TT; off < kWave; off <<= 1) {
        ax += __shfl_xor(ax, off, kWave);
        ay += __shfl_xor(ay, off, kWave);
        az += __shfl_xor(az, off, kWave);
        nearest = fmin(nearest, __shfl_xor(nearest, off, kWave));
      }
      if (MIX && valid && s == 0) {
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        if (P.pm_enabled) interp32<SCHEME>(P, D.f32x, D.f32y, D.f32z, tx, ty, tz, gx, gy, gz);
        const float sx = static_cast<float>(ax), sy = static_cast<float>(ay), sz = static_cast<float>(az);
        static_cast<float*>(out.ax)[slot] = gx + sx;
        static_cast<float*>(out.ay)[slot] = gy + sy;
        static_cast<float*>(out.az)[slot] = gz + sz;
        if (out.pm_x) {
          static_cast<float*>(out.pm_x)[slot] = gx;
          static_cast<float*>(out.pm_y)[slot] = gy;
          static_cast<float*>(out.pm_z)[slot] = gz;
        }
        const uint32_t row = D.order[tp];
        if (out.sr_x) {
          static_cast<float*>(out.sr_x)[row] = sx;
          static_cast<float*>(out.sr_y)[row] = sy;
          static_cast<float*>(out.sr_z)[row] = sz;
        }
        if (out.near2) static_cast<float*>(out.near2)[row] = static_cast<float>(nearest);
      } else if (!MIX && valid && s == 0) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        if (P.pm_enabled) interp64<SCHEME>(P, D.f64x, D.f64y, D.f64z, px, py, pz, gx, gy, gz);
        static_cast<double*>(out.ax)[slot] = gx + ax;
        static_cast<double*>(out.ay)[slot] = gy + ay;
        static_cast<double*>(out.az)[slot] = gz + az;
        if (out.pm_x) {
          static_cast<double*>(out.pm_x)[slot] = gx;
          static_cast<double*>(out.pm_y)[slot] = gy;
          static_cast<double*>(out.pm_z)[slot] = gz;
        }
        const uint32_t row = D.order[tp];
        if (out.sr_x) {
          static_cast<double*>(out.sr_x)[row] = ax;
          static_cast<double*>(out.sr_y)[row] = ay;
          static_cast<double*>(out.sr_z)[row] = az;
        }
        if (out.near2) static_cast<double*>(out.near2)[row] = nearest;
      }
    }
    v = ge;
  }
#pragma unroll
  for (int i = 0; i < kGroupCounters; ++i) {
    if (i == 7 || i == 14) continue;
    const unsigned long long sum = wave_sum(static_cast<unsigned long long>(cnt[i]));
    if (lane == 0 && sum) atomicAdd(&counters[i], sum);
  }
  if (lane == 0) {
    if (size_max) atomicMax(&counters[7], size_max);
    if (entries_max) atomicMax(&counters[14], entries_max);
  }
}

// Two waves per SIMD (256 VGPRs) where the FP64 node expansions are inline:
// the body needs about that many; at 168 VGPRs (3 waves) it spilled 384-512
// bytes per lane and ran 2% (FP64) to 29% (mixed) slower at 256^3. The mixed
// walk with FP32 node expansions (RADIAL false) holds 142 VGPRs and runs 3
// waves per SIMD (4 waves: 128 VGPRs with spills, 16% slower).
template <int SCHEME>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void group_walk64_kernel(
    WalkParams P, GroupParams G, WalkData D, const uint32_t* __restrict__ leaf_of_pos, WalkOutputs out,
    uint32_t* __restrict__ fb_pos, uint32_t* __restrict__ fb_slot, uint32_t* __restrict__ fb_count,
    unsigned long long* __restrict__ counters) {
  group_walk_body<SCHEME, false, false>(P, G, D, leaf_of_pos, out, fb_pos, fb_slot, fb_count, counters);
}

// RADIAL: the radial-coefficient node expansion (FP64; the FP32 expansion is
// the plain monopole + quadrupole one).
template <int SCHEME, bool RADIAL>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(RADIAL ? 2 : 3))) void group_walkmix_kernel(
    WalkParams P, GroupParams G, WalkData D, const uint32_t* __restrict__ leaf_of_pos, WalkOutputs out,
    uint32_t* __restrict__ fb_pos, uint32_t* __restrict__ fb_slot, uint32_t* __restrict__ fb_count,
    unsigned long long* __restrict__ counters) {
  group_walk_body<SCHEME, true, !RADIAL>(P, G, D, leaf_of_pos, out, fb_pos, fb_slot, fb_count, counters);
}

// Mixed adaptive solves: targets the group walk hands back (or every target
// when the group walk does not run) take the FP64 per-target adaptive walk
// without PM, written into FP64 lanes by output slot (nearest pair by source
// row). This kernel adds the FP32 PM interpolation at the fixed-point
// position and writes the FP32 outputs of those targets. pos/slot: the
// targets' tree positions and output slots (identity: position k, slot
// order[k]).
template <int SCHEME>
__global__ __launch_bounds__(256) void adaptive_mixed_finish_kernel(WalkParams P, WalkData D,
                                                                    const uint32_t* __restrict__ pos,
                                                                    const uint32_t* __restrict__ slot, uint32_t n,
                                                                    int identity, const double* __restrict__ sx,
                                                                    const double* __restrict__ sy,
                                                                    const double* __restrict__ sz,
                                                                    const double* __restrict__ near64,
                                                                    WalkOutputs out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t tp = identity ? i : pos[i];
  const uint32_t sl = identity ? D.order[i] : slot[i];
  const uint4 f = D.src32[tp];
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  if (P.pm_enabled) interp32<SCHEME>(P, D.f32x, D.f32y, D.f32z, f.x, f.y, f.z, gx, gy, gz);
  const float ax = static_cast<float>(sx[sl]), ay = static_cast<float>(sy[sl]), az = static_cast<float>(sz[sl]);
  static_cast<float*>(out.ax)[sl] = gx + ax;
  static_cast<float*>(out.ay)[sl] = gy + ay;
  static_cast<float*>(out.az)[sl] = gz + az;
  if (out.pm_x) {
    static_cast<float*>(out.pm_x)[sl] = gx;
    static_cast<float*>(out.pm_y)[sl] = gy;
    static_cast<float*>(out.pm_z)[sl] = gz;
  }
  const uint32_t row = D.order[tp];
  if (out.sr_x) {
    static_cast<float*>(out.sr_x)[row] = ax;
    static_cast<float*>(out.sr_y)[row] = ay;
    static_cast<float*>(out.sr_z)[row] = az;
  }
  if (out.near2) static_cast<float*>(out.near2)[row] = static_cast<float>(near64[row]);
}

__global__ void leaf_of_pos_kernel(const NodeLink* __restrict__ link, std::size_t nodes,
                                   uint32_t* __restrict__ leaf_of_pos) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  const NodeLink lk = link[i];
  if (lk.child_count != 0) return;
  for (uint32_t p = lk.begin; p < lk.begin + lk.count; ++p) leaf_of_pos[p] = static_cast<uint32_t>(i);
}

// ---- mixed walk -----------------------------------------------------------------
// CUBIC: equal box lengths, so separations stay in fixed-point units (one
// scale factor for all axes) and only the pair kernel converts to lengths.
// COUNT: keep the per-target counters (diagnostic launches only).
//
// A leaf that any lane keeps is staged in the wave's LDS window with one
// coalesced load (one source per lane). Every lane then evaluates the residual
// pair kernel against each staged source in turn, reading it with a broadcast
// LDS load, and keeps the term when the pair lies inside r_cut (a lane that
// pruned the leaf compares against a negative cutoff). A leaf's pair terms
// are summed in registers and added to a per-lane partial sum in LDS, which is
// folded into the lane's Kahan-compensated totals (also in LDS) every kWindow
// sources. Keeping these sums out of the node loop's registers spares the
// loop-carried copies the compiler would otherwise make on every node.
//
// Window format. Most waves hold 64 consecutive tree-order targets inside a
// small region: their window stores float offsets from the wave origin O (the
// centre of the targets' fixed-point bounding box), so a separation is three
// float subtractions. The exact minimum image is kept because every offset
// int32(q - O) is exact and, when all targets lie within min(L/16, L/2 - r_cut)
// of O (per axis), a pair inside r_cut cannot straddle the wrap of O's frame.
// Waves that straddle a jump of the space-filling order (targets far apart)
// store raw fixed-point coordinates and form int32(q_s - q_t) per pair.
//
// Box-classified traversal (BOX, the timed walk). A wave whose targets lie in
// a box of half-extent at most box_half_max (2 r_cut) per axis walks the tree
// once for that box instead of node by node for its 64 lanes: the lanes classify 64
// consecutive preorder nodes at a time against the box (pruned when the
// node's tight particle box lies beyond r_cut of the target box, i.e. beyond
// r_cut of every target), a scalar loop follows the visited chain through the
// window, and the kept leaves' sources go to a 128-entry ring in LDS that is
// evaluated one 64-source window at a time with the pair kernel above (each
// lane keeps only its own pairs inside r_cut). Only internal nodes of cells
// small enough for the strict acceptance (h < 0.041 r_cut) and not within
// 25 h of every target (where 2h/r < 0.08 fails for all of them) can be
// accepted by some lane; their subtrees are walked lane by lane as above.
// The pairs and multipoles are therefore those of the lane-by-lane walk;
// only the summation order differs.
constexpr int kWindow = 64;
constexpr int kPairGroup = 2;  // box walk: table reads in flight together

//
// Residual table indexed by the float bits (LOG, the timed walk): the whole
// pair coefficient F(s) = s^{-3/2} - c8 H((s/scale - eps^2) / 4a^2) as a
// function of s = r^2 + eps^2 in table units, whose scale puts r_cut^2 + eps^2
// at s = 1. The cell of s is its exponent and top 5 mantissa bits (32 cubic
// Hermite cells per octave, from the octave of eps^2 up to s = 1, then one
// zero cell), the position in the cell its low 18 mantissa bits as an integer
// (the coefficients carry the 2^-18 powers). eps^2 enters the FMA chain of
// r^2, the chain's last FMA clamps s to [0, 1] (so every pair beyond r_cut
// lands in the zero cell, the cell of s = 1), and the octave base is folded
// into the table's LDS address. A pair then costs a shift, an address add, an
// LDS read, the position (and, convert) and three FMAs instead of a
// reciprocal square root, the Gaussian table and a cutoff select. Cubic
// Hermite cells of relative width 1/32 to 1/64 keep F within 1.5e-7 of its
// value.
//
// The box walk stages the sources of its kept leaves 64 at a time and drops
// those farther than r_cut from the target box (they would all land in the
// zero cell) before they enter the LDS window: the kept records are compacted
// with a ballot and a lane prefix count, and the window is evaluated whenever
// 64 records are in it.

// 16-byte cell `cell` of an LDS table at LDS address `base`: the address is
// one shift-add (the compiler folds a constant shift of a shifted value into
// shift, and, add). Inline asm counts as convergent, so a loop holding it is
// unrolled only with a constant trip count: the box walk's full windows use
// it, the lane-by-lane walk (variable leaf counts) indexes the table plainly.
typedef float f32x4 __attribute__((ext_vector_type(4)));
using lds_f32x4 = __attribute__((address_space(3))) const f32x4;
__device__ __forceinline__ float4 lds_cell(uint32_t base, uint32_t cell) {
  uint32_t a;
  asm("v_lshl_add_u32 %0, %1, 4, %2" : "=v"(a) : "v"(cell), "s"(base));
  const f32x4 v = *(lds_f32x4*)static_cast<uintptr_t>(a);
  return make_float4(v.x, v.y, v.z, v.w);
}

// The box walk's last, partial window has a variable trip count, so its pair
// loop stays rolled (see lds_cell); the unroller's warning about it is muted.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
// DEVN: the traversal bound is the node count on the device (after a
// capture-safe build); otherwise the kernel argument, which the timed walk
// keeps (reading the bound from memory cost it 1 ms of 22 at 256^3).
template <int SCHEME, bool CUBIC, bool COUNT, bool BOX = false, bool LOG = false, int WAVES = kWalkWaves,
          bool DEVN = false>
__global__ __launch_bounds__(WAVES * kWave) __attribute__((amdgpu_waves_per_eu(COUNT ? 4 : 7)))
void walkmix_kernel(WalkParams P, WalkData D, WalkOutputs out, unsigned long long* __restrict__ counters,
                    const NodeLink* __restrict__ node_link, const NodeGeo32* __restrict__ node_geo,
                    const NodeMom32* __restrict__ node_mom, const NodeQuad32* __restrict__ node_quad,
                    const NodeBox32* __restrict__ node_box) {
  // node records are read-only for the kernel's lifetime and indexed by the
  // wave-uniform node number, so they come through the scalar cache
  // the box walk's window holds up to 127 compacted records between evaluations
  __shared__ float4 win[WAVES][BOX ? 2 * kWindow : kWindow];
  __shared__ float4 htab[LOG ? 1 : kHTableCells];
  __shared__ uint32_t list[WAVES][BOX ? kWindow : 1];  // box walk: listed source indices
  // per lane: force totals [0..2] and their Kahan compensations [3..5]
  __shared__ float ksum[WAVES][6][kWave];
  extern __shared__ float4 ltab[];  // LOG: P.log_cells cells
  if constexpr (LOG) {
    for (int i = threadIdx.x; i < static_cast<int>(P.log_cells); i += (WAVES * kWave)) ltab[i] = D.ltab[i];
  } else {
    for (int i = threadIdx.x; i < kHTableCells; i += (WAVES * kWave)) htab[i] = D.htab[i];
  }
  __syncthreads();
  const uint32_t node_count = DEVN ? *D.node_count_dev : P.node_count;
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const uint32_t k = xcd_logical_block() * (WAVES * kWave) + threadIdx.x;
  const bool valid = k < P.n_targets;
  const uint32_t tp = valid ? (P.identity_targets ? k : D.tpos[k]) : 0u;
  const uint32_t slot = valid ? (P.identity_targets ? D.order[k] : D.tslot[k]) : 0u;
  uint32_t cnt[kWalkCounters] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t nxt = valid ? 0u : node_count;
  float nearest = INFINITY;  // counting variant: nearest distinct pair, table units
  uint32_t px = 0, py = 0, pz = 0;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  if (valid) {
    const uint4 t = D.src32[tp];
    px = t.x;
    py = t.y;
    pz = t.z;
    if (P.pm_enabled) interp32<SCHEME>(P, D.f32x, D.f32y, D.f32z, px, py, pz, gx, gy, gz);
  }
  // node tests: fixed-point units (CUBIC) or lengths
  const float s0 = CUBIC ? 1.0f : P.fx_to_len[0];
  const float s1 = CUBIC ? 1.0f : P.fx_to_len[1];
  const float s2 = CUBIC ? 1.0f : P.fx_to_len[2];
  const float to_len = CUBIC ? P.fx_to_len[0] : 1.0f;          // node separation unit -> length
  const float inv_len = CUBIC ? 1.0f / P.fx_to_len[0] : 1.0f;  // length -> node separation unit
  const float prune2 = P.prune2_sep;
  const float contain2 = P.contain2_sep;
  const float theta2 = static_cast<float>(P.theta * P.theta);
  // pairs: table units (lengths * sqrt(t_to_cell)), so r^2 is the H-table coordinate
  const float u0 = P.tab_scale[0], u1 = P.tab_scale[1], u2 = P.tab_scale[2];
  const float rc2 = P.rcut2_tab;
  const float eps2 = P.eps2_tab;
  const float c8 = P.c8_tab;
  // wave origin: centre of the valid targets' bounding box in fixed point
  uint32_t org[3];
  bool near = true;
  float box_half[3];  // the box's half-extents, separation units (covers every target)
  float box_tab[3];   // the same in table units
  {
    const uint32_t pp[3] = {px, py, pz};
    const float sa[3] = {s0, s1, s2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const uint32_t o0 = readlane_u32(pp[a], 0);
      const int32_t off = valid ? static_cast<int32_t>(pp[a] - o0) : 0;
      // wave-uniform through readlane, so the box lives in scalar registers
      const int32_t lo = static_cast<int32_t>(readlane_u32(static_cast<uint32_t>(wave_min(off)), 0));
      const int32_t hi = static_cast<int32_t>(readlane_u32(static_cast<uint32_t>(wave_max(off)), 0));
      const int64_t mid = (static_cast<int64_t>(lo) + static_cast<int64_t>(hi)) / 2;
      org[a] = o0 + static_cast<uint32_t>(static_cast<int32_t>(mid));
      const int64_t half_extent = (static_cast<int64_t>(hi) - static_cast<int64_t>(lo)) / 2 + 1;
      near = near && half_extent <= static_cast<int64_t>(P.fast_half_extent[a]);
      box_half[a] = static_cast<float>(half_extent + 1) * sa[a];
      box_tab[a] = static_cast<float>(half_extent + 1) * P.tab_scale[a];
    }
  }
  const uint32_t ox = org[0], oy = org[1], oz = org[2];
  const float tx = fx_delta(px, ox, u0), ty = fx_delta(py, oy, u1), tz = fx_delta(pz, oz, u2);
#pragma unroll
  for (int c = 0; c < 6; ++c) ksum[w][c][lane] = 0.0f;
  auto accumulate = [&](float vx, float vy, float vz) {
    const float v[3] = {vx, vy, vz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float sum = ksum[w][c][lane], comp = ksum[w][3 + c][lane];
      kahan_add(sum, comp, v[c]);
      ksum[w][c][lane] = sum;
      ksum[w][3 + c][lane] = comp;
    }
  };

  // source -> window record, and window record -> separation from this lane
  auto stage_record = [&]<bool NEAR>(const uint4& sv, float mj) -> float4 {
    if constexpr (NEAR)
      return make_float4(fx_delta(sv.x, ox, u0), fx_delta(sv.y, oy, u1), fx_delta(sv.z, oz, u2), mj);
    else
      return make_float4(__uint_as_float(sv.x), __uint_as_float(sv.y), __uint_as_float(sv.z), mj);
  };
  auto separation = [&]<bool NEAR>(const float4& sj, float& dx, float& dy, float& dz) {
    if constexpr (NEAR) {
      dx = sj.x - tx;
      dy = sj.y - ty;
      dz = sj.z - tz;
    } else {
      dx = fx_delta(__float_as_uint(sj.x), px, u0);
      dy = fx_delta(__float_as_uint(sj.y), py, u1);
      dz = fx_delta(__float_as_uint(sj.z), pz, u2);
    }
  };

  uint32_t pend = 0;  // wave-uniform count of sources in the partial sums
  float psx = 0.0f, psy = 0.0f, psz = 0.0f;  // partial pair sums since the last fold
  auto fold = [&]() {
    const float gm = P.force_scale;
    accumulate(psx * gm, psy * gm, psz * gm);
    psx = psy = psz = 0.0f;
    pend = 0;
    if (COUNT && lane == 0) ++cnt[7];
  };

  // the residual pair kernel against the cn staged sources; UMASS: every
  // source has the uniform mass, which the force scale already carries; SEL:
  // lanes may carry a negative cutoff (they pruned the leaf or accepted an
  // ancestor), so their terms are dropped explicitly
  // LOG: cell (bits of s) >> 18 of the table is lcell[cell], at LDS address
  // lbase + 16 * cell
  const float4* const lcell = ltab - P.log_base;
  const uint32_t lbase = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_f32x4*)ltab - P.log_base));
  auto pairs = [&]<bool NEAR, bool UMASS, bool SEL = true>(uint32_t c0, uint32_t cn, float lim) {
    float lx = 0.0f, ly = 0.0f, lz = 0.0f;
#pragma unroll 4
    for (uint32_t si = 0; si < cn; ++si) {
      const float4 sj = win[w][si];
      float dx, dy, dz;
      separation.template operator()<NEAR>(sj, dx, dy, dz);
      float f;
      if constexpr (LOG) {
        // s = r^2 + eps^2 clamped to [0, 1]: s >= eps^2 >= 2^first_octave, and
        // s = 1 (r_cut and beyond) is the zero cell
        const float sc = __builtin_amdgcn_fmed3f(fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2))), 0.0f, 1.0f);
        const uint32_t bits = __float_as_uint(sc);
        float4 c;
        if constexpr (SEL) c = lcell[bits >> kLogMantissaShift];
        else c = lds_cell(lbase, bits >> kLogMantissaShift);
        const float u = static_cast<float>(bits & ((1u << kLogMantissaShift) - 1u));
        float fr = fmaf(fmaf(fmaf(c.w, u, c.z), u, c.y), u, c.x);
        if constexpr (!UMASS) fr *= sj.w;
        f = (SEL && lim < 0.0f) ? 0.0f : fr;
      } else {
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));  // = table coordinate
        const float rinv = __builtin_amdgcn_rsqf(r2 + eps2);
        const float soft = rinv * rinv * rinv;
        const int ci = static_cast<int>(fminf(r2, static_cast<float>(kHTableCells - 1)));
        const float fu = __builtin_amdgcn_fractf(r2);
        const float4 c = htab[ci];
        const float hcoef = fmaf(fmaf(fmaf(c.w, fu, c.z), fu, c.y), fu, c.x);
        float fr = fmaf(-hcoef, c8, soft);
        if constexpr (!UMASS) fr *= sj.w;
        // The target itself has r2 == 0 exactly (same fixed-point value), as
        // does any coincident source: its term is exactly zero (dx = dy = dz = 0
        // against a finite softened value), so no self test is needed. The
        // counting variant counts coincident sources as the reference does.
        const bool pair = r2 <= lim;
        if constexpr (COUNT) {
          const bool distinct = pair && (r2 != 0.0f || (c0 + si) != tp);
          cnt[2] += distinct;
          nearest = distinct ? fminf(nearest, r2) : nearest;
        }
        f = pair ? fr : 0.0f;
      }
      lx = fmaf(f, dx, lx);
      ly = fmaf(f, dy, ly);
      lz = fmaf(f, dz, lz);
    }
    psx += lx;
    psy += ly;
    psz += lz;
  };
  // A full window of the box walk (LOG): sources in groups of kPairGroup whose
  // table reads are in flight together before their polynomials are evaluated.
  auto window_pairs = [&]<bool UMASS>() {
    constexpr int G = kPairGroup;
    float lx = 0.0f, ly = 0.0f, lz = 0.0f;
#pragma unroll 2
    for (int s0 = 0; s0 < kWindow; s0 += G) {
      float dx[G], dy[G], dz[G], mj[G];
      uint32_t bits[G];
      float4 c[G];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const float4 sj = win[w][s0 + k];
        dx[k] = sj.x - tx;
        dy[k] = sj.y - ty;
        dz[k] = sj.z - tz;
        mj[k] = sj.w;
        bits[k] = __float_as_uint(
            __builtin_amdgcn_fmed3f(fmaf(dx[k], dx[k], fmaf(dy[k], dy[k], fmaf(dz[k], dz[k], eps2))), 0.0f, 1.0f));
      }
#pragma unroll
      for (int k = 0; k < G; ++k) c[k] = lds_cell(lbase, bits[k] >> kLogMantissaShift);
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const float u = static_cast<float>(bits[k] & ((1u << kLogMantissaShift) - 1u));
        float fr = fmaf(fmaf(fmaf(c[k].w, u, c[k].z), u, c[k].y), u, c[k].x);
        if constexpr (!UMASS) fr *= mj[k];
        lx = fmaf(fr, dx[k], lx);
        ly = fmaf(fr, dy[k], ly);
        lz = fmaf(fr, dz[k], lz);
      }
    }
    psx += lx;
    psy += ly;
    psz += lz;
  };

  // Strict acceptance of internal node n (cell centre and half from g) for
  // this lane, with the COM separation it needs left in mf*. Nodes with
  // h >= 0.041 r_cut are never accepted (2h/r < 0.08 with r inside r_cut), so
  // the caller skips the test for them.
  const MixFrame F{s0, s1, s2, to_len, inv_len, contain2, theta2};
  auto accepts = [&](uint32_t n, const NodeGeo32& g, float& mfx, float& mfy, float& mfz) -> bool {
    return strict_accept32(P, F, g, node_mom[n], node_quad, n, px, py, pz, mfx, mfy, mfz);
  };

  // The evaluating walk prunes against each node's tight particle box. Every
  // pair inside r_cut still lies in a kept leaf, and a node the reference
  // accepts as a multipole (its whole cell inside r_cut) is always kept, so
  // the pairs and multipoles are the reference's; only nodes whose particles
  // all lie beyond r_cut are skipped earlier than with the octree cells.
  auto walk = [&]<bool NEAR>(uint32_t end) {
    uint32_t n = wave_min_u32(nxt);
    while (n < end) {
      if (COUNT && lane == 0) ++cnt[5];
      const NodeLink lk = node_link[n];
      const NodeBox32 b = node_box[n];
      const bool act = nxt == n;
      if (!__ballot(act)) {
        if (lane == 0) atomicOr(&counters[kFaultSlot], 1ull);
        break;
      }
      const float ax_ = fabsf(fx_delta(b.cx, px, s0)), ay_ = fabsf(fx_delta(b.cy, py, s1)),
                  az_ = fabsf(fx_delta(b.cz, pz, s2));
      const float ex = fmaxf(0.0f, ax_ - b.hx), ey = fmaxf(0.0f, ay_ - b.hy), ez = fmaxf(0.0f, az_ - b.hz);
      const bool keep = act && !(fmaf(ex, ex, fmaf(ey, ey, ez * ez)) > prune2);
      if (lk.child_count == 0) {
        // a leaf's subtree is itself: every waiting lane moves on to n + 1,
        // so that is the wave's next node too
        nxt = act ? n + 1 : nxt;
        if (__ballot(keep)) {
          const float lim = keep ? rc2 : -1.0f;
          const uint32_t e0 = lk.begin + lk.count;
          for (uint32_t c0 = lk.begin; c0 < e0; c0 += kWindow) {
            const uint32_t cn = e0 - c0 < static_cast<uint32_t>(kWindow) ? e0 - c0 : static_cast<uint32_t>(kWindow);
            if (COUNT && lane == 0) cnt[6] += cn;
            if (pend + cn > static_cast<uint32_t>(kWindow)) fold();
            pend += cn;
            {
              const uint4 sv = D.src32[c0 + (static_cast<uint32_t>(lane) < cn ? lane : 0)];
              const float mj = D.uniform_mass ? 1.0f : __uint_as_float(sv.w);
              if (static_cast<uint32_t>(lane) < cn) win[w][lane] = stage_record.template operator()<NEAR>(sv, mj);
            }
            __builtin_amdgcn_wave_barrier();
            if (D.uniform_mass) pairs.template operator()<NEAR, true>(c0, cn, lim);
            else pairs.template operator()<NEAR, false>(c0, cn, lim);
            __builtin_amdgcn_wave_barrier();
          }
        }
        ++n;
        continue;
      }
      nxt = (act && !keep) ? n + lk.skip : nxt;
      if (__ballot(keep)) {
        bool take_mp = false;
        float mfx = 0.0f, mfy = 0.0f, mfz = 0.0f;
        if (b.cell_half < P.accept_half_max) {
          if (keep) take_mp = accepts(n, node_geo[n], mfx, mfy, mfz);
        }
        if (keep) nxt = take_mp ? n + lk.skip : n + 1;
        if (__ballot(take_mp) && take_mp) {
          float cx = 0.0f, cy = 0.0f, cz = 0.0f;
          node_accel32(node_mom[n], node_quad[n], mfx, mfy, mfz, P, cx, cy, cz);
          accumulate(cx, cy, cz);
        }
      }
      n = next_node(n, nxt);
    }
  };

  // Box classification of node q against the wave's target box: 0 pruned,
  // 1 kept leaf (sources [lb, lb + lc)), 3 opened, 4 a subtree some lane may
  // accept a multipole in (walked lane by lane); next is q + 1 to descend,
  // q + skip to move past the subtree.
  auto box_classify = [&](uint32_t q, uint32_t& next, uint32_t& lb, uint32_t& lc, uint32_t& span,
                          uint32_t& nchild) -> int {
    const NodeLink lk = node_link[q];
    const NodeBox32 b = node_box[q];
    span = lk.skip;
    nchild = lk.child_count;
    const float ex = fmaxf(0.0f, fabsf(fx_delta(b.cx, ox, s0)) - b.hx - box_half[0]);
    const float ey = fmaxf(0.0f, fabsf(fx_delta(b.cy, oy, s1)) - b.hy - box_half[1]);
    const float ez = fmaxf(0.0f, fabsf(fx_delta(b.cz, oz, s2)) - b.hz - box_half[2]);
    const bool keep = !(fmaf(ex, ex, fmaf(ey, ey, ez * ez)) > prune2);
    if (lk.child_count == 0) {
      lb = lk.begin;
      lc = lk.count;
      return keep ? 1 : 0;
    }
    if (!keep) {
      next = q + lk.skip;
      return 0;
    }
    if (b.cell_half < P.accept_half_max) {
      // every target within 25 h of the COM fails 2h/r < 0.08
      const NodeMom32 M = node_mom[q];
      const float dx = (fabsf(fx_delta(M.comx, ox, s0)) + box_half[0]) * to_len;
      const float dy = (fabsf(fx_delta(M.comy, oy, s1)) + box_half[1]) * to_len;
      const float dz = (fabsf(fx_delta(M.comz, oz, s2)) + box_half[2]) * to_len;
      const float far2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      if (!(far2 <= 624.9f * b.cell_half * b.cell_half)) {
        next = q + lk.skip;
        return 4;
      }
    }
    return 3;
  };

  auto box_walk = [&]() {
    const float lim = valid ? rc2 : -1.0f;
    // a source whose offset from the box centre lies beyond r_cut of the box
    // (outside half-extent + 1 fixed-point unit per axis) is beyond r_cut of
    // every target; the 1e-5 margin covers the rounding of the offsets, so
    // every dropped source is one whose s reaches the zero cell
    const float cull2 = rc2 * 1.00001f;
    uint32_t nl = 0;            // list[w][0, nl): listed sources, staged 64 at a time
    uint32_t wn = 0;            // compacted records in win[w][0, wn)
    auto evaluate = [&](uint32_t cn) {
      __builtin_amdgcn_wave_barrier();
      if (pend + cn > static_cast<uint32_t>(kWindow)) fold();
      pend += cn;
      // every lane sums every staged source (pairs beyond r_cut are zero);
      // lanes without a target are never written out
      if (LOG && cn == static_cast<uint32_t>(kWindow)) {
        if (D.uniform_mass) window_pairs.template operator()<true>();
        else window_pairs.template operator()<false>();
      } else if (D.uniform_mass) {
        pairs.template operator()<true, true, false>(0, cn, lim);
      } else {
        pairs.template operator()<true, false, false>(0, cn, lim);
      }
      __builtin_amdgcn_wave_barrier();
    };
    // stage the next cn listed sources, keep those within r_cut of the box
    auto eval_window = [&](uint32_t cn) {
      __builtin_amdgcn_wave_barrier();
      const bool on = static_cast<uint32_t>(lane) < cn;
      const uint4 sv = D.src32[list[w][on ? lane : 0]];
      const float mj = D.uniform_mass ? 1.0f : __uint_as_float(sv.w);
      const float4 rec = stage_record.template operator()<true>(sv, mj);
      const float ex = fmaxf(fabsf(rec.x) - box_tab[0], 0.0f);
      const float ey = fmaxf(fabsf(rec.y) - box_tab[1], 0.0f);
      const float ez = fmaxf(fabsf(rec.z) - box_tab[2], 0.0f);
      const bool keep = on && !(fmaf(ex, ex, fmaf(ey, ey, ez * ez)) > cull2);
      const unsigned long long km = __ballot(keep);
      const uint32_t at =
          wn + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(km >> 32),
                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(km), 0u));
      if (keep) win[w][at] = rec;
      wn += static_cast<uint32_t>(__popcll(km));
      nl = 0;
      if (wn >= static_cast<uint32_t>(kWindow)) {
        evaluate(kWindow);
        wn -= kWindow;
        // wn < 64: move the records past the window to its front
        if (static_cast<uint32_t>(lane) < wn) {
          const float4 t = win[w][kWindow + lane];
          win[w][lane] = t;
        }
      }
    };
    // the lanes' classified nodes: kept leaves' sources go to the list (a
    // leaf may continue in the next batch), subtrees some lane may accept in
    // are walked lane by lane
    auto take = [&](bool on, int kind, uint32_t q, uint32_t lb, uint32_t lc, uint32_t next) {
      unsigned long long rest = __ballot(on && kind == 1);
      while (rest) {
        const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<long long>(rest)) - 1);
        rest &= rest - 1ull;
        const uint32_t b0 = readlane_u32(lb, l), c = readlane_u32(lc, l);
        for (uint32_t c0 = 0; c0 < c;) {
          // nl < kWindow here: fill the list up to kWindow entries
          const uint32_t room = static_cast<uint32_t>(kWindow) - nl;
          const uint32_t cn = c - c0 < room ? c - c0 : room;
          if (static_cast<uint32_t>(lane) < cn) list[w][nl + lane] = b0 + c0 + lane;
          nl += cn;
          c0 += cn;
          if (nl == static_cast<uint32_t>(kWindow)) eval_window(kWindow);
        }
      }
      unsigned long long sub = __ballot(on && kind == 4);
      while (sub) {
        const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<long long>(sub)) - 1);
        sub &= sub - 1ull;
        const uint32_t q4 = readlane_u32(q, l), end = readlane_u32(next, l);
        // the lane-by-lane walk stages its leaves in the window's front
        if (wn) {
          evaluate(wn);
          wn = 0;
        }
        nxt = valid ? q4 : end;
        walk.template operator()<true>(end);
      }
    };
    // windows of 64 consecutive preorder nodes of [start, end), following the
    // visited chain through each window
    auto chain_walk = [&](uint32_t start, uint32_t end) {
      uint32_t p = start;
      while (p < end) {
        const uint32_t q = p + lane;
        uint32_t next = q + 1, lb = 0, lc = 0, span = 1, nch = 0;
        const int kind = q < end ? box_classify(q, next, lb, lc, span, nch) : 3;
        unsigned long long chain = 0;
        uint32_t cur = p;
        while (cur < p + kWave && cur < end) {
          const uint32_t l = cur - p;
          chain |= 1ull << l;
          cur = readlane_u32(next, l);
        }
        p = cur;
        take((chain >> lane) & 1ull, kind, q, lb, lc, next);
      }
    };
    chain_walk(0, node_count);
    if (nl) eval_window(nl);
    if (wn) evaluate(wn);
  };

  bool boxed = false;
  if constexpr (BOX) {
    boxed = near && box_half[0] <= P.box_half_max[0] && box_half[1] <= P.box_half_max[1] &&
            box_half[2] <= P.box_half_max[2];
  }
  if (boxed) {
    if constexpr (BOX) box_walk();
  } else if (near) {
    walk.template operator()<true>(node_count);
  } else {
    walk.template operator()<false>(node_count);
  }
  fold();

  // Diagnostic launches also replay the reference's walk over the octree
  // cells (no pair work) for its per-target node counters.
  if constexpr (COUNT) {
    uint32_t cnxt = valid ? 0u : node_count;
    uint32_t n = wave_min_u32(cnxt);
    while (n < node_count) {
      const NodeLink lk = node_link[n];
      const NodeGeo32 g = node_geo[n];
      const bool act = cnxt == n;
      if (!__ballot(act)) break;  // reported by the evaluating walk
      cnt[0] += act;
      const float h = g.half * inv_len;
      const float ax_ = fabsf(fx_delta(g.cx, px, s0)), ay_ = fabsf(fx_delta(g.cy, py, s1)),
                  az_ = fabsf(fx_delta(g.cz, pz, s2));
      const float ex = fmaxf(0.0f, ax_ - h), ey = fmaxf(0.0f, ay_ - h), ez = fmaxf(0.0f, az_ - h);
      const bool keep = act && !(fmaf(ex, ex, fmaf(ey, ey, ez * ez)) > prune2);
      cnt[1] += act && !keep;
      if (lk.child_count == 0) {
        cnt[4] += keep;
        cnxt = act ? n + 1 : cnxt;
        ++n;
        continue;
      }
      cnxt = (act && !keep) ? n + lk.skip : cnxt;
      bool take_mp = false;
      if (keep && g.half < P.accept_half_max) {
        float mfx, mfy, mfz;
        take_mp = accepts(n, g, mfx, mfy, mfz);
      }
      cnt[3] += take_mp;
      if (keep) cnxt = take_mp ? n + lk.skip : n + 1;
      n = next_node(n, cnxt);
    }
  }

  if (COUNT) flush_counters(counters, cnt);
  if (!valid) return;
  const float ax = ksum[w][0][lane] - ksum[w][3][lane];
  const float ay = ksum[w][1][lane] - ksum[w][4][lane];
  const float az = ksum[w][2][lane] - ksum[w][5][lane];
  static_cast<float*>(out.ax)[slot] = gx + ax;
  static_cast<float*>(out.ay)[slot] = gy + ay;
  static_cast<float*>(out.az)[slot] = gz + az;
  if (out.pm_x) {
    static_cast<float*>(out.pm_x)[slot] = gx;
    static_cast<float*>(out.pm_y)[slot] = gy;
    static_cast<float*>(out.pm_z)[slot] = gz;
  }
  const uint32_t row = D.order[tp];
  if (out.sr_x) {
    static_cast<float*>(out.sr_x)[row] = ax;
    static_cast<float*>(out.sr_y)[row] = ay;
    static_cast<float*>(out.sr_z)[row] = az;
  }
  if (COUNT && out.near2) static_cast<float*>(out.near2)[row] = nearest * P.tab_to_len2;
}

#pragma clang diagnostic pop

// ---- mixed walk, one target per wave --------------------------------------------
// For sparse active sets the 64 targets of a walkmix_kernel wave lie far
// apart, and the wave walks the union of 64 distant neighbourhoods. Here a
// wave serves one target. Its lanes classify 64 consecutive preorder nodes at
// a time with walkmix_kernel's tests (the cutoff prune against the tight
// particle boxes, strict_accept32 for internal nodes whose cell is small
// enough), follow the visited chain through that window, list the kept
// leaves' sources and the accepted nodes in LDS and evaluate the lists 64
// entries at a time with walkmix_kernel's pair and multipole arithmetic. The
// pairs and multipoles are walkmix_kernel's; per-lane partial sums are
// reduced in FP64 at the end.
constexpr int kTargetList = 1024;  // listed sources per wave before an evaluation
constexpr int kTargetNodes = 128;  // listed multipoles per wave before an evaluation

// The tree walk of the per-wave kernels: the lanes classify preorder nodes
// p .. p + 63 (classify(q, next, lb, lc) returns 0 pruned, 1 kept leaf with
// sources [lb, lb + lc), 2 accepted node or 3 opened, and sets next to q + 1
// to descend or q + skip to move past the subtree), a scalar loop follows the
// visited chain through the window with readlane, and the chain's kept
// leaves' sources and accepted nodes are appended to the wave's LDS lists
// (list: kTargetList entries, mps: kTargetNodes). evaluate() consumes the
// lists and resets nl and nm; it runs whenever a list would overflow and at
// the end.
template <class Classify, class Evaluate>
__device__ __forceinline__ void walk_target_wave(uint32_t node_count, int lane, uint32_t* list, uint32_t* mps,
                                                 uint32_t& nl, uint32_t& nm, Classify&& classify,
                                                 Evaluate&& evaluate) {
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t p = 0;
  while (p < node_count) {
    const uint32_t q = p + lane;
    uint32_t next = q + 1, lb = 0, lc = 0;
    const int kind = q < node_count ? classify(q, next, lb, lc) : 3;
    unsigned long long chain = 0;
    uint32_t cur = p;
    while (cur < p + kWave && cur < node_count) {
      const uint32_t l = cur - p;
      chain |= 1ull << l;
      cur = readlane_u32(next, l);
    }
    p = cur;
    const bool on = (chain >> lane) & 1ull;
    const unsigned long long mm = __ballot(on && kind == 2);
    if (nm + static_cast<uint32_t>(__popcll(mm)) > static_cast<uint32_t>(kTargetNodes)) evaluate();
    if (on && kind == 2) mps[nm + __popcll(mm & below)] = q;
    nm += __popcll(mm);
    unsigned long long rest = __ballot(on && kind == 1);
    while (rest) {
      const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<long long>(rest)) - 1);
      rest &= rest - 1ull;
      const uint32_t b0 = readlane_u32(lb, l), c = readlane_u32(lc, l);
      for (uint32_t c0 = 0; c0 < c; c0 += kWave) {
        const uint32_t cn = c - c0 < static_cast<uint32_t>(kWave) ? c - c0 : static_cast<uint32_t>(kWave);
        if (nl + cn > static_cast<uint32_t>(kTargetList)) evaluate();
        if (static_cast<uint32_t>(lane) < cn) list[nl + lane] = b0 + c0 + lane;
        nl += cn;
      }
    }
  }
  evaluate();
}

template <int SCHEME, bool CUBIC>
__global__ __launch_bounds__(kWalkThreads) void walkmix_target_kernel(WalkParams P, WalkData D, WalkOutputs out) {
  __shared__ float4 htab[kHTableCells];
  __shared__ uint32_t list[kWalkWaves][kTargetList];
  __shared__ uint32_t mps[kWalkWaves][kTargetNodes];
  for (int i = threadIdx.x; i < kHTableCells; i += kWalkThreads) htab[i] = D.htab[i];
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const uint32_t k = xcd_logical_block() * kWalkWaves + w;
  if (k >= P.n_targets) return;  // wave-uniform; no block barrier follows
  const uint32_t tp = P.identity_targets ? k : D.tpos[k];
  const uint32_t slot = P.identity_targets ? D.order[k] : D.tslot[k];
  const uint4 t = D.src32[tp];
  const uint32_t px = t.x, py = t.y, pz = t.z;
  const float s0 = CUBIC ? 1.0f : P.fx_to_len[0];
  const float s1 = CUBIC ? 1.0f : P.fx_to_len[1];
  const float s2 = CUBIC ? 1.0f : P.fx_to_len[2];
  const float to_len = CUBIC ? P.fx_to_len[0] : 1.0f;
  const float inv_len = CUBIC ? 1.0f / P.fx_to_len[0] : 1.0f;
  const MixFrame F{s0, s1, s2, to_len, inv_len, P.contain2_sep, static_cast<float>(P.theta * P.theta)};
  const float prune2 = P.prune2_sep;
  const float u0 = P.tab_scale[0], u1 = P.tab_scale[1], u2 = P.tab_scale[2];
  const float rc2 = P.rcut2_tab, eps2 = P.eps2_tab, c8 = P.c8_tab;

  float lx = 0.0f, ly = 0.0f, lz = 0.0f;  // pair sums (force-scale units)
  float mx = 0.0f, my = 0.0f, mz = 0.0f;  // multipole accelerations
  uint32_t nl = 0, nm = 0;
  auto evaluate = [&]() {
    __builtin_amdgcn_wave_barrier();
    for (uint32_t e = lane; e < nl; e += kWave) {
      const uint4 sv = D.src32[list[w][e]];
      const float dx = fx_delta(sv.x, px, u0), dy = fx_delta(sv.y, py, u1), dz = fx_delta(sv.z, pz, u2);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));  // = H-table coordinate
      const float rinv = __builtin_amdgcn_rsqf(r2 + eps2);
      const float soft = rinv * rinv * rinv;
      const int ci = static_cast<int>(fminf(r2, static_cast<float>(kHTableCells - 1)));
      const float fu = __builtin_amdgcn_fractf(r2);
      const float4 c = htab[ci];
      const float hcoef = fmaf(fmaf(fmaf(c.w, fu, c.z), fu, c.y), fu, c.x);
      float fr = fmaf(-hcoef, c8, soft);
      if (!D.uniform_mass) fr *= __uint_as_float(sv.w);
      const float f = r2 <= rc2 ? fr : 0.0f;  // the target itself: dx = dy = dz = 0
      lx = fmaf(f, dx, lx);
      ly = fmaf(f, dy, ly);
      lz = fmaf(f, dz, lz);
    }
    for (uint32_t e = lane; e < nm; e += kWave) {
      const uint32_t n = mps[w][e];
      const NodeMom32 M = D.mom32[n];
      const float dx = fx_delta(M.comx, px, s0) * to_len, dy = fx_delta(M.comy, py, s1) * to_len,
                  dz = fx_delta(M.comz, pz, s2) * to_len;
      float cx = 0.0f, cy = 0.0f, cz = 0.0f;
      node_accel32(M, D.quad32[n], dx, dy, dz, P, cx, cy, cz);
      mx += cx;
      my += cy;
      mz += cz;
    }
    __builtin_amdgcn_wave_barrier();
    nl = 0;
    nm = 0;
  };

  // walkmix_kernel's tests: the prune against the tight particle box and the
  // strict acceptance of small internal nodes
  auto classify = [&](uint32_t q, uint32_t& next, uint32_t& lb, uint32_t& lc) -> int {
    const NodeLink lk = D.link[q];
    const NodeBox32 b = D.box32[q];
    const float ax_ = fabsf(fx_delta(b.cx, px, s0)), ay_ = fabsf(fx_delta(b.cy, py, s1)),
                az_ = fabsf(fx_delta(b.cz, pz, s2));
    const float ex = fmaxf(0.0f, ax_ - b.hx), ey = fmaxf(0.0f, ay_ - b.hy), ez = fmaxf(0.0f, az_ - b.hz);
    const bool keep = !(fmaf(ex, ex, fmaf(ey, ey, ez * ez)) > prune2);
    if (lk.child_count == 0) {
      lb = lk.begin;
      lc = lk.count;
      return keep ? 1 : 0;
    }
    if (!keep) {
      next = q + lk.skip;
      return 0;
    }
    bool take = false;
    if (b.cell_half < P.accept_half_max) {
      float mfx, mfy, mfz;
      take = strict_accept32(P, F, D.geo32[q], D.mom32[q], D.quad32, q, px, py, pz, mfx, mfy, mfz);
    }
    next = take ? q + lk.skip : q + 1;
    return take ? 2 : 3;
  };
  walk_target_wave(D.node_count_dev ? *D.node_count_dev : P.node_count, lane, list[w], mps[w], nl, nm, classify, evaluate);
  const double gm = P.force_scale;
  const double sx = gm * wave_sum(static_cast<double>(lx)) + wave_sum(static_cast<double>(mx));
  const double sy = gm * wave_sum(static_cast<double>(ly)) + wave_sum(static_cast<double>(my));
  const double sz = gm * wave_sum(static_cast<double>(lz)) + wave_sum(static_cast<double>(mz));
  if (lane != 0) return;
  const float ax = static_cast<float>(sx), ay = static_cast<float>(sy), az = static_cast<float>(sz);
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  if (P.pm_enabled) interp32<SCHEME>(P, D.f32x, D.f32y, D.f32z, px, py, pz, gx, gy, gz);
  static_cast<float*>(out.ax)[slot] = gx + ax;
  static_cast<float*>(out.ay)[slot] = gy + ay;
  static_cast<float*>(out.az)[slot] = gz + az;
  if (out.pm_x) {
    static_cast<float*>(out.pm_x)[slot] = gx;
    static_cast<float*>(out.pm_y)[slot] = gy;
    static_cast<float*>(out.pm_z)[slot] = gz;
  }
  const uint32_t row = D.order[tp];
  if (out.sr_x) {
    static_cast<float*>(out.sr_x)[row] = ax;
    static_cast<float*>(out.sr_y)[row] = ay;
    static_cast<float*>(out.sr_z)[row] = az;
  }
}

// ---- FP64 walk, one target per wave ----------------------------------------------
// walkmix_target_kernel's mapping for the FP64 strict walk: the lanes classify
// 64 consecutive preorder nodes with walk64_kernel's tests (the cutoff prune
// against the octree cell, accept64 for internal nodes), the visited chain
// lists kept leaves' sources and accepted nodes in LDS, and the lists are
// evaluated 64 entries at a time with walk64_kernel's pair and node
// arithmetic. The pairs and multipoles are walk64_kernel's; the sum runs in
// another order (per-lane partial sums, then a wave reduction).
template <int SCHEME>
__global__ __launch_bounds__(kWalkThreads) void walk64_target_kernel(WalkParams P, WalkData D, WalkOutputs out) {
  __shared__ uint32_t list[kWalkWaves][kTargetList];
  __shared__ uint32_t mps[kWalkWaves][kTargetNodes];
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const uint32_t k = xcd_logical_block() * kWalkWaves + w;
  if (k >= P.n_targets) return;  // wave-uniform
  const uint32_t tp = P.identity_targets ? k : D.tpos[k];
  const uint32_t slot = P.identity_targets ? D.order[k] : D.tslot[k];
  const double4 t = D.src64[tp];
  const double px = t.x, py = t.y, pz = t.z;
  const double L0 = P.L[0], L1 = P.L[1], L2 = P.L[2];

  double ax = 0.0, ay = 0.0, az = 0.0;
  uint32_t nl = 0, nm = 0;
  auto evaluate = [&]() {
    __builtin_amdgcn_wave_barrier();
    for (uint32_t e = lane; e < nl; e += kWave) {
      const uint32_t j = list[w][e];
      const double4 sj = D.src64[j];
      const double qx = mi64(sj.x - px, L0), qy = mi64(sj.y - py, L1), qz = mi64(sj.z - pz, L2);
      const double r2 = qx * qx + qy * qy + qz * qz;
      // coincident pairs have a zero vector; the target itself is skipped
      if (j != tp && r2 <= P.rcut2 && !(qx == 0.0 && qy == 0.0 && qz == 0.0)) {
        const double f = residual_inv_r3(r2, P.eps, P.a) * P.G;
        const double mj = D.uniform_mass ? P.uniform_mass : sj.w;
        ax += f * mj * qx;
        ay += f * mj * qy;
        az += f * mj * qz;
      }
    }
    for (uint32_t e = lane; e < nm; e += kWave) {
      const uint32_t n = mps[w][e];
      const NodeMom64 M = D.mom64[n];
      const double dx = mi64(M.comx - px, L0), dy = mi64(M.comy - py, L1), dz = mi64(M.comz - pz, L2);
      node_accel64_call(M, D.quad64[n], dx, dy, dz, P, ax, ay, az);  // out of line: register pressure
    }
    __builtin_amdgcn_wave_barrier();
    nl = 0;
    nm = 0;
  };

  // walk64_kernel's tests: the prune against the octree cell and accept64
  auto classify = [&](uint32_t q, uint32_t& next, uint32_t& lb, uint32_t& lc) -> int {
    const NodeLink lk = D.link[q];
    const NodeGeo64 g = D.geo64[q];
    const double h = g.half;
    const double ax_ = fabs(mi64(g.cx - px, L0)), ay_ = fabs(mi64(g.cy - py, L1)), az_ = fabs(mi64(g.cz - pz, L2));
    const double ex = fmax(0.0, ax_ - h), ey = fmax(0.0, ay_ - h), ez = fmax(0.0, az_ - h);
    const bool keep = within_cutoff(ex * ex + ey * ey + ez * ez, P.rcut, P.rcut2);
    if (lk.child_count == 0) {
      lb = lk.begin;
      lc = lk.count;
      return keep ? 1 : 0;
    }
    if (!keep) {
      next = q + lk.skip;
      return 0;
    }
    const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
    const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
    const bool contained = within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
    const NodeMom64 M = D.mom64[q];
    const double mdx = mi64(M.comx - px, L0), mdy = mi64(M.comy - py, L1), mdz = mi64(M.comz - pz, L2);
    const bool take = accept64(h, D.quad64[q].offset, inside, contained, mdx * mdx + mdy * mdy + mdz * mdz, P);
    next = take ? q + lk.skip : q + 1;
    return take ? 2 : 3;
  };
  walk_target_wave((D.node_count_dev ? *D.node_count_dev : P.node_count), lane, list[w], mps[w], nl, nm, classify, evaluate);
  ax = wave_sum(ax);
  ay = wave_sum(ay);
  az = wave_sum(az);
  if (lane != 0) return;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (P.pm_enabled) interp64<SCHEME>(P, D.f64x, D.f64y, D.f64z, px, py, pz, gx, gy, gz);
  static_cast<double*>(out.ax)[slot] = gx + ax;
  static_cast<double*>(out.ay)[slot] = gy + ay;
  static_cast<double*>(out.az)[slot] = gz + az;
  if (out.pm_x) {
    static_cast<double*>(out.pm_x)[slot] = gx;
    static_cast<double*>(out.pm_y)[slot] = gy;
    static_cast<double*>(out.pm_z)[slot] = gz;
  }
  const uint32_t row = D.order[tp];
  if (out.sr_x) {
    static_cast<double*>(out.sr_x)[row] = ax;
    static_cast<double*>(out.sr_y)[row] = ay;
    static_cast<double*>(out.sr_z)[row] = az;
  }
}

// ---- FP64 strict walk, box-classified (the timed FP64 path) ----------------------
// walkmix_kernel's box-classified traversal in FP64 for solves without
// counters: a wave whose 64 targets lie in a box of half-extent at most 2 r_cut
// (and below L/2 - r_cut, so that every pair inside r_cut has its minimum
// image relative to the box centre) classifies 64 preorder nodes at a time
// against the box (octree cells beyond r_cut of the box are pruned; internal
// nodes that no target can accept are opened), lists the kept leaves'
// sources in an LDS ring and evaluates them one 64-source window at a time;
// subtrees some target may accept as a multipole (cells with h < 0.041 r_cut
// farther than 25 h from some target) and waves that do not fit a box are
// walked lane by lane with walk64_kernel's tests (cell prune, accept64,
// node_accel64). Pair term: (r^2 + eps^2)^{-3/2} from v_rsq_f64 and two Newton
// steps, L(r)/r^3 = H(r^2 / 4a^2) / 8a^3 from a degree-5 polynomial per cell
// of an LDS table (Chebyshev interpolation, error below 1e-14 of H), in FP64
// throughout. The pairs and multipoles are walk64_kernel's; the pair
// arithmetic and the summation order differ at the round-off level.
constexpr int kH64Cells = kH64MaxCells;

// s^{-1/2} for s in [1e-30, 1e30]: the FP32 reciprocal square root (about
// 2^-23 relative) refined by two Newton steps in FP64 (error squared each
// step), so the result is within an ulp or two of FP64.
__device__ __forceinline__ double rsq64(double s) {
  double y = static_cast<double>(__builtin_amdgcn_rsqf(static_cast<float>(s)));
  double e = fma(-s * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-s * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// JVP: the transpose of the acceleration Jacobian applied to lam, as in
// walk64_kernel<..., JVP> (pair term G (F w + 2 F'(t) (r . w) r), w = m_k
// lam_j - m_j lam_k, with F' = dF/d(r^2) from the same table), for
// multipole_order 0 (no accepted nodes).
template <int SCHEME, bool JVP = false>
__global__ __launch_bounds__(kWalkThreads) __attribute__((amdgpu_waves_per_eu(4))) void walk64_box_kernel(
    WalkParams P, WalkData D, WalkOutputs out, unsigned long long* __restrict__ counters) {
  // forward: the box walk's window holds up to 127 compacted records
  __shared__ double4 win[kWalkWaves][JVP ? kWindow : 2 * kWindow];
  __shared__ double4 winl[JVP ? kWalkWaves : 1][JVP ? kWindow : 1];  // JVP: lam of the staged sources
  __shared__ uint32_t list[kWalkWaves][kWindow];                       // box walk: listed source indices
  __shared__ double2 htab[3 * kH64Cells];
  for (int i = threadIdx.x; i < 3 * P.h64_cells; i += kWalkThreads) htab[i] = D.h64tab[i];
  __syncthreads();
  const uint32_t node_count = D.node_count_dev ? *D.node_count_dev : P.node_count;
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const uint32_t k = xcd_logical_block() * kWalkThreads + threadIdx.x;
  const bool valid = k < P.n_targets;
  const uint32_t tp = valid ? (P.identity_targets ? k : D.tpos[k]) : 0u;
  const uint32_t slot = valid ? (P.identity_targets ? D.order[k] : D.tslot[k]) : 0u;
  double px = 0.0, py = 0.0, pz = 0.0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  double4 lk_ = make_double4(0.0, 0.0, 0.0, 0.0);  // JVP: lam of the target
  double mk = 0.0;                                  // JVP: mass of the target
  if (valid) {
    const double4 t = D.src64[tp];
    px = t.x;
    py = t.y;
    pz = t.z;
    if constexpr (JVP) {
      lk_ = D.lam64[tp];
      mk = D.uniform_mass ? P.uniform_mass : t.w;
      if (P.pm_enabled)
        vjp_gather64<SCHEME>(P, D.f64x, D.f64y, D.f64z, D.nu64, lk_.x, lk_.y, lk_.z, mk, px, py, pz, gx, gy, gz);
    } else {
      if (P.pm_enabled) interp64<SCHEME>(P, D.f64x, D.f64y, D.f64z, px, py, pz, gx, gy, gz);
    }
  }
  const double L0 = P.L[0], L1 = P.L[1], L2 = P.L[2];
  const double iL0 = 1.0 / L0, iL1 = 1.0 / L1, iL2 = 1.0 / L2;
  // minimum image by a reciprocal multiply: d - L rint(d / L) up to the
  // rounding of d / L, which only matters at |d| within an ulp of L/2, far
  // outside r_cut
  auto mi_fast = [&](double d, int a) -> double {
    const double L = a == 0 ? L0 : (a == 1 ? L1 : L2);
    const double iL = a == 0 ? iL0 : (a == 1 ? iL1 : iL2);
    return fma(-L, rint(d * iL), d);
  };
  const double rc = P.rcut, rc2 = P.rcut2;
  // target box: offsets from lane 0's target (lane 0 is valid whenever any lane is)
  const double o0 = readlane_f64(px, 0), o1 = readlane_f64(py, 0), o2 = readlane_f64(pz, 0);
  const double ux = valid ? mi_fast(px - o0, 0) : 0.0, uy = valid ? mi_fast(py - o1, 1) : 0.0,
               uz = valid ? mi_fast(pz - o2, 2) : 0.0;
  // wave-uniform values through readlane, so they live in scalar registers
  const double lox = readlane_f64(wave_min(ux), 0), hix = readlane_f64(wave_max(ux), 0),
               loy = readlane_f64(wave_min(uy), 0), hiy = readlane_f64(wave_max(uy), 0),
               loz = readlane_f64(wave_min(uz), 0), hiz = readlane_f64(wave_max(uz), 0);
  const double cx = 0.5 * (lox + hix), cy = 0.5 * (loy + hiy), cz = 0.5 * (loz + hiz);
  const double ex = 0.5 * (hix - lox) * (1.0 + 1.0e-12), ey = 0.5 * (hiy - loy) * (1.0 + 1.0e-12),
               ez = 0.5 * (hiz - loz) * (1.0 + 1.0e-12);
  const bool any = __ballot(valid) != 0ull;
  const double emax = 2.0 * rc;  // the mixed walk's bound on the box half-extent
  const bool boxed = any && ex <= emax && ey <= emax && ez <= emax && ex + rc < 0.5 * L0 * (1.0 - 1.0e-9) &&
                     ey + rc < 0.5 * L1 * (1.0 - 1.0e-9) && ez + rc < 0.5 * L2 * (1.0 - 1.0e-9);
  // box centre in the tree frame. Boxed waves stage each source as its image
  // nearest the box centre and take the target's image likewise: for waves
  // away from the seam both are the tree-frame coordinates themselves (the
  // shift is zero), so r = x_j - x_k is the reference's difference exactly.
  const double bx = o0 + cx, by = o1 + cy, bz = o2 + cz;
  auto image = [&](double v, double b, int a) -> double {
    const double L = a == 0 ? L0 : (a == 1 ? L1 : L2);
    const double iL = a == 0 ? iL0 : (a == 1 ? iL1 : iL2);
    return fma(-L, rint((v - b) * iL), v);
  };
  const double qx = image(px, bx, 0), qy = image(py, by, 1), qz = image(pz, bz, 2);
  const double eps2 = P.eps2_64, c8 = P.c8_64, tscale = P.h64_scale;
  const double cmax = static_cast<double>(P.h64_cells) - 1.0;
  double ax = 0.0, ay = 0.0, az = 0.0;

  // residual pair kernel in FP64: G m_j [(r^2+eps^2)^{-3/2} - H(r^2/4a^2)/8a^3] r_vec (G m_j applied by the
  // caller); with the derivative fp = dF/d(r^2) when asked (JVP)
  auto residual = [&](double r2, double* fp) -> double {
    const double sft = fmax(r2 + eps2, 1.0e-30);  // coincident unsoftened pairs: finite, times a zero vector
    const double y = rsq64(sft);
    const double soft = y * y * y;
    const double t = fmin(r2 * tscale, cmax + 0.999999999999);
    const int ci = static_cast<int>(t);
    const double u = t - static_cast<double>(ci);
    const double2 a01 = htab[3 * ci], a23 = htab[3 * ci + 1], a45 = htab[3 * ci + 2];
    const double hv = fma(fma(fma(fma(fma(a45.y, u, a45.x), u, a23.y), u, a23.x), u, a01.y), u, a01.x);
    if (fp) {
      const double dh = fma(fma(fma(fma(5.0 * a45.y, u, 4.0 * a45.x), u, 3.0 * a23.y), u, 2.0 * a23.x), u, a01.y);
      *fp = fma(-1.5 * soft, y * y, -c8 * dh * tscale);
    }
    return fma(-hv, c8, soft);
  };
  // F at r = 0 (the reference's limit: 0 without softening), for coincident pairs in the JVP
  const double f_zero = JVP ? residual_inv_r3(0.0, P.eps, P.a) : 0.0;
  // the staged window against this lane's target: NEAR = images nearest the box
  // centre (boxed waves), else tree-frame coordinates with a minimum image per pair
  auto pairs = [&]<bool NEAR>(uint32_t cn, double lim) {
    double lx = 0.0, ly = 0.0, lz = 0.0;
#pragma unroll 2
    for (uint32_t si = 0; si < cn; ++si) {
      const double4 sj = win[w][si];
      double dx, dy, dz;
      if constexpr (NEAR) {
        dx = sj.x - qx;
        dy = sj.y - qy;
        dz = sj.z - qz;
      } else {
        dx = mi_fast(sj.x - px, 0);
        dy = mi_fast(sj.y - py, 1);
        dz = mi_fast(sj.z - pz, 2);
      }
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      if constexpr (JVP) {
        // the target itself: w = 0; a coincident source: F(0) w
        double fp = 0.0;
        double f = residual(r2, &fp);
        f = r2 > 0.0 ? f : f_zero;
        const double4 lj = winl[w][si];
        double wx, wy, wz;
        if (D.uniform_mass) {  // w / m (the mass goes into the final scale)
          wx = lj.x - lk_.x;
          wy = lj.y - lk_.y;
          wz = lj.z - lk_.z;
        } else {
          wx = mk * lj.x - sj.w * lk_.x;
          wy = mk * lj.y - sj.w * lk_.y;
          wz = mk * lj.z - sj.w * lk_.z;
        }
        const bool in = r2 <= lim;
        f = in ? f : 0.0;
        const double cr = in ? 2.0 * fp * fma(dx, wx, fma(dy, wy, dz * wz)) : 0.0;
        lx = fma(f, wx, fma(cr, dx, lx));
        ly = fma(f, wy, fma(cr, dy, ly));
        lz = fma(f, wz, fma(cr, dz, lz));
      } else {
        double f = residual(r2, nullptr);
        if (!D.uniform_mass) f *= sj.w;
        f = r2 <= lim ? f : 0.0;  // the target itself and coincident sources: zero vector
        lx = fma(f, dx, lx);
        ly = fma(f, dy, ly);
        lz = fma(f, dz, lz);
      }
    }
    const double gm = D.uniform_mass ? P.G * P.uniform_mass : P.G;
    ax = fma(gm, lx, ax);
    ay = fma(gm, ly, ay);
    az = fma(gm, lz, az);
  };
  auto stage = [&]<bool NEAR>(uint32_t j, bool on) {
    const double4 sv = D.src64[j];
    if constexpr (JVP) {
      const double4 lv = D.lam64[j];
      if (on) winl[w][lane] = lv;
    }
    if (on) {
      if constexpr (NEAR)
        win[w][lane] = make_double4(image(sv.x, bx, 0), image(sv.y, by, 1), image(sv.z, bz, 2), sv.w);
      else
        win[w][lane] = sv;
    }
  };

  // lane-by-lane walk of [nxt, end) with walk64_kernel's tests
  uint32_t nxt = valid ? 0u : node_count;
  auto walk = [&]<bool NEAR>(uint32_t end) {
    uint32_t n = wave_min_u32(nxt);
    while (n < end) {
      const NodeLink lk = D.link[n];
      const NodeGeo64 g = D.geo64[n];
      const bool act = nxt == n;
      if (!__ballot(act)) {
        if (lane == 0) atomicOr(&counters[kFaultSlot], 1ull);
        break;
      }
      const double h = g.half;
      const double ndx = mi_fast(g.cx - px, 0), ndy = mi_fast(g.cy - py, 1), ndz = mi_fast(g.cz - pz, 2);
      const double ax_ = fabs(ndx), ay_ = fabs(ndy), az_ = fabs(ndz);
      const double fx_ = fmax(0.0, ax_ - h), fy_ = fmax(0.0, ay_ - h), fz_ = fmax(0.0, az_ - h);
      const bool keep = act && within_cutoff(fx_ * fx_ + fy_ * fy_ + fz_ * fz_, P.rcut, P.rcut2);
      if (act && !keep) nxt = n + lk.skip;
      if (lk.child_count == 0) {
        if (keep) nxt = n + 1;
        if (__ballot(keep)) {
          const double lim = keep ? rc2 : -1.0;
          const uint32_t e0 = lk.begin + lk.count;
          for (uint32_t c0 = lk.begin; c0 < e0; c0 += kWindow) {
            const uint32_t cn = e0 - c0 < static_cast<uint32_t>(kWindow) ? e0 - c0 : static_cast<uint32_t>(kWindow);
            __builtin_amdgcn_wave_barrier();
            stage.template operator()<NEAR>(c0 + (static_cast<uint32_t>(lane) < cn ? lane : 0),
                                            static_cast<uint32_t>(lane) < cn);
            __builtin_amdgcn_wave_barrier();
            pairs.template operator()<NEAR>(cn, lim);
          }
          __builtin_amdgcn_wave_barrier();
        }
      } else if (__ballot(keep)) {
        if constexpr (JVP) {
          // the forward's strict acceptance; an accepted node's Jacobian terms
          // go to the target and to the node's accumulators
          bool take_mp = false;
          double mdx = 0.0, mdy = 0.0, mdz = 0.0;
          if (keep && P.quad) {
            const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
            const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
            const bool contained = within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
            const NodeMom64 M = D.mom64[n];
            mdx = mi_fast(M.comx - px, 0);
            mdy = mi_fast(M.comy - py, 1);
            mdz = mi_fast(M.comz - pz, 2);
            take_mp = accept64(h, D.quad64[n].offset, inside, contained, mdx * mdx + mdy * mdy + mdz * mdz, P);
          }
          if (keep) nxt = take_mp ? n + lk.skip : n + 1;
          if (__ballot(take_mp) && take_mp) {
            jvp_node64(D.mom64[n], D.quad64[n], mdx, mdy, mdz, P, lk_.x, lk_.y, lk_.z, D.vjp_node_acc + 10ull * n,
                       ax, ay, az);
            const unsigned int slot_i = atomicAdd(D.vjp_accepted_count, 1u);
            if (slot_i < D.vjp_accepted_cap)
              D.vjp_accepted[slot_i] = (static_cast<unsigned long long>(tp) << 32) | n;
          }
        } else {
          bool take_mp = false;
          double mdx = 0.0, mdy = 0.0, mdz = 0.0;
          if (keep) {
            const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
            const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
            const bool contained = within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
            const NodeMom64 M = D.mom64[n];
            mdx = mi_fast(M.comx - px, 0);
            mdy = mi_fast(M.comy - py, 1);
            mdz = mi_fast(M.comz - pz, 2);
            take_mp = accept64(h, D.quad64[n].offset, inside, contained, mdx * mdx + mdy * mdy + mdz * mdz, P);
            nxt = take_mp ? n + lk.skip : n + 1;
          }
          if (__ballot(take_mp) && take_mp) node_accel64(D.mom64[n], D.quad64[n], mdx, mdy, mdz, P, ax, ay, az);
        }
      }
      n = next_node(n, nxt);
    }
  };

  if (boxed) {
    // box classification of node q: 0 pruned, 1 kept leaf, 3 opened, 4 a
    // subtree walked lane by lane
    auto classify = [&](uint32_t q, uint32_t& next, uint32_t& lb, uint32_t& lc) -> int {
      const NodeLink lk = D.link[q];
      const NodeGeo64 g = D.geo64[q];
      const double h = g.half;
      const double dx = fmax(0.0, fabs(mi_fast(g.cx - bx, 0)) - h - ex);
      const double dy = fmax(0.0, fabs(mi_fast(g.cy - by, 1)) - h - ey);
      const double dz = fmax(0.0, fabs(mi_fast(g.cz - bz, 2)) - h - ez);
      const bool keep = dx * dx + dy * dy + dz * dz <= rc2 * (1.0 + 1.0e-9);
      if (lk.child_count == 0) {
        lb = lk.begin;
        lc = lk.count;
        return keep ? 1 : 0;
      }
      if (!keep) {
        next = q + lk.skip;
        return 0;
      }
      if (P.quad && h < 0.041 * rc) {
        // accept64 needs 2h/r < 0.08: every target within 25 h of the COM fails it
        const NodeMom64 M = D.mom64[q];
        const double fx_ = fabs(mi_fast(M.comx - bx, 0)) + ex, fy_ = fabs(mi_fast(M.comy - by, 1)) + ey,
                     fz_ = fabs(mi_fast(M.comz - bz, 2)) + ez;
        if (!(fx_ * fx_ + fy_ * fy_ + fz_ * fz_ <= 625.0 * h * h * (1.0 - 1.0e-9))) {
          next = q + lk.skip;
          return 4;
        }
      }
      return 3;
    };
    const double lim = valid ? rc2 : -1.0;
    // forward: staged sources beyond r_cut of the target box (every pair with
    // them lies beyond r_cut, so their terms are zero) are dropped before the
    // window, as in walkmix_kernel's box walk
    const double cull2 = rc2 * (1.0 + 1.0e-9);
    uint32_t nl = 0;  // list[w][0, nl): listed sources, staged 64 at a time
    uint32_t wn = 0;  // forward: compacted records in win[w][0, wn)
    auto evaluate = [&](uint32_t cn) {
      __builtin_amdgcn_wave_barrier();
      pairs.template operator()<true>(cn, lim);
      __builtin_amdgcn_wave_barrier();
    };
    auto eval_window = [&](uint32_t cn) {
      __builtin_amdgcn_wave_barrier();
      const bool on = static_cast<uint32_t>(lane) < cn;
      const uint32_t j = list[w][on ? lane : 0];
      nl = 0;
      if constexpr (JVP) {
        // the same drop without a second window: kept sources beyond the
        // window's free room go back to the front of the list (as indices)
        // and are staged again with the next batch
        const double4 sv = D.src64[j];
        const double4 rec = make_double4(image(sv.x, bx, 0), image(sv.y, by, 1), image(sv.z, bz, 2), sv.w);
        const double fx_ = fmax(0.0, fabs(rec.x - bx) - ex), fy_ = fmax(0.0, fabs(rec.y - by) - ey),
                     fz_ = fmax(0.0, fabs(rec.z - bz) - ez);
        const bool keep = on && fx_ * fx_ + fy_ * fy_ + fz_ * fz_ <= cull2;
        const unsigned long long km = __ballot(keep);
        const uint32_t pos = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(km >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(km), 0u));
        const uint32_t room = static_cast<uint32_t>(kWindow) - wn;
        const uint32_t kept = static_cast<uint32_t>(__popcll(km));
        if (keep && pos < room) {
          win[w][wn + pos] = rec;
          winl[w][wn + pos] = D.lam64[j];
        }
        if (keep && pos >= room) list[w][pos - room] = j;
        const uint32_t taken = kept < room ? kept : room;
        wn += taken;
        nl = kept - taken;
        if (wn == static_cast<uint32_t>(kWindow)) {
          evaluate(kWindow);
          wn = 0;
        }
      } else {
        const double4 sv = D.src64[j];
        const double4 rec = make_double4(image(sv.x, bx, 0), image(sv.y, by, 1), image(sv.z, bz, 2), sv.w);
        const double fx_ = fmax(0.0, fabs(rec.x - bx) - ex), fy_ = fmax(0.0, fabs(rec.y - by) - ey),
                     fz_ = fmax(0.0, fabs(rec.z - bz) - ez);
        const bool keep = on && fx_ * fx_ + fy_ * fy_ + fz_ * fz_ <= cull2;
        const unsigned long long km = __ballot(keep);
        const uint32_t at =
            wn + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(km >> 32),
                                           __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(km), 0u));
        if (keep) win[w][at] = rec;
        wn += static_cast<uint32_t>(__popcll(km));
        if (wn >= static_cast<uint32_t>(kWindow)) {
          evaluate(kWindow);
          wn -= kWindow;
          if (static_cast<uint32_t>(lane) < wn) {
            const double4 t = win[w][kWindow + lane];
            win[w][lane] = t;
          }
        }
      }
    };
    uint32_t p = 0;
    while (p < node_count) {
      const uint32_t q = p + lane;
      uint32_t next = q + 1, lb = 0, lc = 0;
      const int kind = q < node_count ? classify(q, next, lb, lc) : 3;
      unsigned long long chain = 0;
      uint32_t cur = p;
      while (cur < p + kWave && cur < node_count) {
        const uint32_t l = cur - p;
        chain |= 1ull << l;
        cur = readlane_u32(next, l);
      }
      const uint32_t base = p;
      p = cur;
      const bool on = (chain >> lane) & 1ull;
      unsigned long long rest = __ballot(on && kind == 1);
      while (rest) {
        const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<long long>(rest)) - 1);
        rest &= rest - 1ull;
        const uint32_t b0 = readlane_u32(lb, l), c = readlane_u32(lc, l);
        for (uint32_t c0 = 0; c0 < c;) {
          const uint32_t room = static_cast<uint32_t>(kWindow) - nl;
          const uint32_t cn = c - c0 < room ? c - c0 : room;
          if (static_cast<uint32_t>(lane) < cn) list[w][nl + lane] = b0 + c0 + lane;
          nl += cn;
          c0 += cn;
          if (nl == static_cast<uint32_t>(kWindow)) eval_window(kWindow);
        }
      }
      unsigned long long sub = __ballot(on && kind == 4);
      while (sub) {
        const uint32_t l = static_cast<uint32_t>(__ffsll(static_cast<long long>(sub)) - 1);
        sub &= sub - 1ull;
        const uint32_t end = readlane_u32(next, l);
        // the lane-by-lane walk stages its leaves in the window's front
        if (wn) {
          evaluate(wn);
          wn = 0;
        }
        nxt = valid ? base + l : end;
        walk.template operator()<true>(end);
      }
    }
    while (nl) eval_window(nl);  // JVP: each call takes at least one source or evaluates a full window
    if (wn) evaluate(wn);
  } else {
    walk.template operator()<false>(node_count);
  }
  if (!valid) return;
  static_cast<double*>(out.ax)[slot] = gx + ax;
  static_cast<double*>(out.ay)[slot] = gy + ay;
  static_cast<double*>(out.az)[slot] = gz + az;
  if (out.pm_x) {
    static_cast<double*>(out.pm_x)[slot] = gx;
    static_cast<double*>(out.pm_y)[slot] = gy;
    static_cast<double*>(out.pm_z)[slot] = gz;
  }
  const uint32_t row = D.order[tp];
  if (out.sr_x) {
    static_cast<double*>(out.sr_x)[row] = ax;
    static_cast<double*>(out.sr_y)[row] = ay;
    static_cast<double*>(out.sr_z)[row] = az;
  }
}

// ---- force VJP: accepted-node terms ----------------------------------------
// For a direct pair (t sees p) the symmetric pair term of p's walk adds
// M_tp lam_t to s_p, M_tp = d a_t / d x_p = G m_p [F I + 2 F' r r^T]; it is
// right only if t also sees p directly. For an accepted (t, N) and p in N
// within r_cut that sees t directly (p accepts no node holding t): remove
// M_tp lam_t from s_p, and add to s_t the term p's force contributes,
// M_pt lam_p = G m_t [F lam_p + 2 F' (r . lam_p) r]. rec is sorted.
// A recorded VJP: more accepted (target, node) pairs than list slots.
__global__ void vjp_count_check_kernel(const unsigned int* __restrict__ count, unsigned int cap,
                                       unsigned long long* __restrict__ fault) {
  if (threadIdx.x == 0 && *count > cap) atomicOr(fault, kFaultPrepared);
}

// nrec_dev (a recorded VJP): the list's real length on the device, nrec its capacity.
__global__ __launch_bounds__(256) void vjp_pair_fix_kernel(const unsigned long long* __restrict__ rec, uint32_t nrec_cap,
                                                           const unsigned int* __restrict__ nrec_dev,
                                                           const NodeLink* __restrict__ link,
                                                           const double4* __restrict__ src,
                                                           const double4* __restrict__ lam, int uniform, double um,
                                                           WalkParams P, double* __restrict__ corr) {
  const uint32_t nrec = nrec_dev ? min(*nrec_dev, nrec_cap) : nrec_cap;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrec) return;
  const uint32_t t = static_cast<uint32_t>(rec[i] >> 32), node = static_cast<uint32_t>(rec[i] & 0xFFFFFFFFull);
  const double4 xt = src[t], lt = lam[t];
  const double mt = uniform ? um : xt.w;
  const NodeLink lk = link[node];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (uint32_t pp = lk.begin; pp < lk.begin + lk.count; ++pp) {
    const double4 xp = src[pp];
    const double dx = mi64(xp.x - xt.x, P.L[0]), dy = mi64(xp.y - xt.y, P.L[1]), dz = mi64(xp.z - xt.z, P.L[2]);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 <= P.rcut2) || !(r2 > 0.0)) continue;
    const unsigned long long key = static_cast<unsigned long long>(pp) << 32;
    uint32_t lo = 0, hi = nrec;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if (rec[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    bool mutual = false;
    for (uint32_t j = lo; j < nrec && static_cast<uint32_t>(rec[j] >> 32) == pp; ++j) {
      const NodeLink l2 = link[static_cast<uint32_t>(rec[j] & 0xFFFFFFFFull)];
      if (t >= l2.begin && t < l2.begin + l2.count) {
        mutual = true;
        break;
      }
    }
    if (mutual) continue;
    const double f = residual_inv_r3(r2, P.eps, P.a), fd = residual_inv_r3_dt(r2, P.eps, P.a);
    const double4 lp = lam[pp];
    const double mp = uniform ? um : xp.w;
    const double rt = dx * lt.x + dy * lt.y + dz * lt.z, rp = dx * lp.x + dy * lp.y + dz * lp.z;
    const double gp = P.G * mp, gt = P.G * mt;
    atomicAdd(&corr[3ull * pp + 0], -gp * (f * lt.x + 2.0 * fd * rt * dx));
    atomicAdd(&corr[3ull * pp + 1], -gp * (f * lt.y + 2.0 * fd * rt * dy));
    atomicAdd(&corr[3ull * pp + 2], -gp * (f * lt.z + 2.0 * fd * rt * dz));
    s0 += gt * (f * lp.x + 2.0 * fd * rp * dx);
    s1 += gt * (f * lp.y + 2.0 * fd * rp * dy);
    s2 += gt * (f * lp.z + 2.0 * fd * rp * dz);
  }
  atomicAdd(&corr[3ull * t + 0], s0);
  atomicAdd(&corr[3ull * t + 1], s1);
  atomicAdd(&corr[3ull * t + 2], s2);
}

// U and V of node N spread over its particles p (d_p = x_p - COM_N):
// s_p += (m_p / M) U + m_p G_V d_p, with G_V the symmetric matrix of
// sum_q V_q d Q_q / d d_p for the stored components (Q_q of one particle:
// m (3 d d^T - |d|^2 I) traceless parts qxx..qzz and the trace m |d|^2).
__global__ __launch_bounds__(256) void vjp_node_push_kernel(uint32_t nodes, const double* __restrict__ acc,
                                                            const NodeLink* __restrict__ link,
                                                            const NodeMom64* __restrict__ mom,
                                                            const double4* __restrict__ src, int uniform, double um,
                                                            WalkParams P, double* __restrict__ corr) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  const double* a = acc + 10ull * n;
  bool any = false;
#pragma unroll
  for (int k = 0; k < 10; ++k) any = any || a[k] != 0.0;
  if (!any) return;
  const NodeMom64 M = mom[n];
  if (!(M.mass > 0.0)) return;
  const double ux = a[0] / M.mass, uy = a[1] / M.mass, uz = a[2] / M.mass;
  const double v0 = a[3], v1 = a[4], v2 = a[5], v3 = a[6], v4 = a[7], v5 = a[8], v6 = a[9];
  const double gxx = 4.0 * v0 - 2.0 * v3 - 2.0 * v5 + 2.0 * v6, gxy = 3.0 * v1, gxz = 3.0 * v2;
  const double gyy = -2.0 * v0 + 4.0 * v3 - 2.0 * v5 + 2.0 * v6, gyz = 3.0 * v4;
  const double gzz = -2.0 * v0 - 2.0 * v3 + 4.0 * v5 + 2.0 * v6;
  const NodeLink lk = link[n];
  for (uint32_t pp = lk.begin; pp < lk.begin + lk.count; ++pp) {
    const double4 xp = src[pp];
    const double mp = uniform ? um : xp.w;
    const double dx = mi64(xp.x - M.comx, P.L[0]), dy = mi64(xp.y - M.comy, P.L[1]), dz = mi64(xp.z - M.comz, P.L[2]);
    atomicAdd(&corr[3ull * pp + 0], mp * (ux + gxx * dx + gxy * dy + gxz * dz));
    atomicAdd(&corr[3ull * pp + 1], mp * (uy + gxy * dx + gyy * dy + gyz * dz));
    atomicAdd(&corr[3ull * pp + 2], mp * (uz + gxz * dx + gyz * dy + gzz * dz));
  }
}

__global__ __launch_bounds__(256) void vjp_apply_kernel(uint32_t n, const uint32_t* __restrict__ order,
                                                        const double* __restrict__ corr, double* __restrict__ ox,
                                                        double* __restrict__ oy, double* __restrict__ oz) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t row = order[k];
  ox[row] += corr[3ull * k + 0];
  oy[row] += corr[3ull * k + 1];
  oz[row] += corr[3ull * k + 2];
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

// dH/dt: derivative of the series below t = 1/64, of the closed form above
// (the same split as tree_pm_split_kernel.hpp:248-285).
double gaussian_h_slope(double t) {
  constexpr double c = 2.25675833419102514779231780624309034;
  if (t < 0.015625) {
    const double coef[8] = {1.0 / 3.0, -1.0 / 5.0, 1.0 / 14.0, -1.0 / 54.0,
                            1.0 / 264.0, -1.0 / 1560.0, 1.0 / 10800.0, -1.0 / 85680.0};
    double d = 0.0;
    for (int i = 7; i >= 1; --i) d = d * t + i * coef[i];
    return c * d;
  }
  return (0.5 * c * std::exp(-t) - 1.5 * gaussian_h_exact(t)) / t;
}

template <int SCHEME, bool JVP = false>
void launch64(const WalkParams& p, const WalkData& d, const WalkOutputs& out, unsigned long long* counters,
              hipStream_t s) {
  walk64_kernel<SCHEME, JVP><<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, 0, s>>>(p, d, out, counters);
  CHUI_LAUNCH_CHECK("walk64_kernel");
}
template <int SCHEME, bool CUBIC>
void launchmix(const WalkParams& p, const WalkData& d, const WalkOutputs& out, unsigned long long* counters,
               hipStream_t s) {
  const bool devn = d.node_count_dev != nullptr;
  if (p.count && !devn)
    walkmix_kernel<SCHEME, CUBIC, true><<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, 0, s>>>(
        p, d, out, counters, d.link, d.geo32, d.mom32, d.quad32, d.box32);
  else if (p.count)
    throw std::logic_error("counting walk after a capture-safe build");
  else if (p.target_per_wave)
    walkmix_target_kernel<SCHEME, CUBIC><<<grid_for(p.n_targets, kWalkWaves), kWalkThreads, 0, s>>>(p, d, out);
  else if (p.log_cells > 0 && devn)
    walkmix_kernel<SCHEME, CUBIC, false, true, true, kWalkWaves, true>
        <<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, p.log_cells * sizeof(float4), s>>>(
            p, d, out, counters, d.link, d.geo32, d.mom32, d.quad32, d.box32);
  else if (p.log_cells > 0 && p.walk_waves == 8)
    walkmix_kernel<SCHEME, CUBIC, false, true, true, 8>
        <<<grid_for(p.n_targets, 8 * kWave), 8 * kWave, p.log_cells * sizeof(float4), s>>>(
            p, d, out, counters, d.link, d.geo32, d.mom32, d.quad32, d.box32);
  else if (p.log_cells > 0 && p.walk_waves == 2)
    walkmix_kernel<SCHEME, CUBIC, false, true, true, 2>
        <<<grid_for(p.n_targets, 2 * kWave), 2 * kWave, p.log_cells * sizeof(float4), s>>>(
            p, d, out, counters, d.link, d.geo32, d.mom32, d.quad32, d.box32);
  else if (p.log_cells > 0)
    walkmix_kernel<SCHEME, CUBIC, false, true, true>
        <<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, p.log_cells * sizeof(float4), s>>>(
            p, d, out, counters, d.link, d.geo32, d.mom32, d.quad32, d.box32);
  else if (devn)
    walkmix_kernel<SCHEME, CUBIC, false, true, false, kWalkWaves, true>
        <<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, 0, s>>>(p, d, out, counters, d.link, d.geo32, d.mom32,
                                                                       d.quad32, d.box32);
  else
    walkmix_kernel<SCHEME, CUBIC, false, true><<<grid_for(p.n_targets, kWalkThreads), kWalkThreads, 0, s>>>(
        p, d, out, counters, d.link, d.geo32, d.mom32, d.quad32, d.box32);
  CHUI_LAUNCH_CHECK("walkmix_kernel");
}

}  // namespace

std::vector<float> log_residual_table(double scale, double eps, double a, int32_t& first_octave,
                                      uint32_t& cells) {
  // s = (r^2 + eps^2) * scale; F(s) = s^{-3/2} - c H(x(s)) with
  // x(s) = (s / scale - eps^2) / (4 a^2) and c = scale^{-3/2} / (8 a^3).
  cells = 0;
  const float eps2_tab = static_cast<float>(eps * eps * scale);
  if (!(eps2_tab > 0.0f) || !std::isfinite(eps2_tab)) return {};
  const int e0 = std::ilogb(eps2_tab);
  const int octaves = -e0;
  if (octaves < 1 || octaves > kLogMaxOctaves) return {};
  const double c = std::pow(scale, -1.5) / (8.0 * a * a * a);
  const double dxds = 1.0 / (scale * 4.0 * a * a);
  auto F = [&](double sv) { return std::pow(sv, -1.5) - c * gaussian_h_exact((sv / scale - eps * eps) / (4.0 * a * a)); };
  auto dF = [&](double sv) {
    return -1.5 * std::pow(sv, -2.5) - c * gaussian_h_slope((sv / scale - eps * eps) / (4.0 * a * a)) * dxds;
  };
  const int K = octaves * kLogCellsPerOctave;
  std::vector<float> tab(4 * static_cast<std::size_t>(K + 1), 0.0f);
  const double u1 = std::ldexp(1.0, -kLogMantissaShift), u2 = u1 * u1, u3 = u2 * u1;
  for (int o = 0; o < octaves; ++o) {
    const double base = std::ldexp(1.0, e0 + o);
    for (int j = 0; j < kLogCellsPerOctave; ++j) {
      const double s0 = base * (1.0 + static_cast<double>(j) / kLogCellsPerOctave);
      const double s1 = base * (1.0 + static_cast<double>(j + 1) / kLogCellsPerOctave);
      const double h = s1 - s0;
      const double f0 = F(s0), f1 = F(s1), d0 = h * dF(s0), d1 = h * dF(s1);
      float* t = &tab[4 * static_cast<std::size_t>(o * kLogCellsPerOctave + j)];
      t[0] = static_cast<float>(f0);
      t[1] = static_cast<float>(d0 * u1);
      t[2] = static_cast<float>((3.0 * (f1 - f0) - 2.0 * d0 - d1) * u2);
      t[3] = static_cast<float>((2.0 * (f0 - f1) + d0 + d1) * u3);
    }
  }
  first_octave = e0;
  cells = static_cast<uint32_t>(K + 1);  // the last cell is zero: s >= 1, beyond r_cut
  return tab;
}

std::vector<double> gaussian_quintic_table(double t_max, int cells) {
  // Interpolation at the Chebyshev points of each cell: error below
  // 2 (h/4)^6 max|H^(6)| / 6!, under 1e-16 for cells of width 0.04 and
  // about 1e-13 at width 0.2 (t_max = 50 on 256 cells). The 6 x 6
  // Vandermonde system in u is solved in long double.
  const int K = cells;
  const double h = t_max / K;
  std::vector<double> tab(6 * static_cast<std::size_t>(K));
  constexpr int D = 6;
  long double un[D];
  for (int j = 0; j < D; ++j)
    un[j] = 0.5L + 0.5L * std::cos(static_cast<long double>(3.14159265358979323846264338327950288L) *
                                   (2 * j + 1) / (2 * D));
  for (int i = 0; i < K; ++i) {
    long double A[D][D + 1];
    for (int j = 0; j < D; ++j) {
      long double pw = 1.0L;
      for (int c = 0; c < D; ++c) {
        A[j][c] = pw;
        pw *= un[j];
      }
      A[j][D] = gaussian_h_exact((i + static_cast<double>(un[j])) * h);
    }
    for (int c = 0; c < D; ++c) {  // Gaussian elimination with partial pivoting
      int piv = c;
      for (int r = c + 1; r < D; ++r)
        if (std::fabs(static_cast<double>(A[r][c])) > std::fabs(static_cast<double>(A[piv][c]))) piv = r;
      for (int q = 0; q <= D; ++q) std::swap(A[c][q], A[piv][q]);
      for (int r = 0; r < D; ++r) {
        if (r == c) continue;
        const long double f = A[r][c] / A[c][c];
        for (int q = c; q <= D; ++q) A[r][q] -= f * A[c][q];
      }
    }
    for (int c = 0; c < D; ++c) tab[6 * static_cast<std::size_t>(i) + c] = static_cast<double>(A[c][D] / A[c][c]);
  }
  return tab;
}

std::vector<float> gaussian_coefficient_table(WalkParams& p, double t_max) {
  // Cubic Hermite cells on [0, t_max]: value c0 + c1 u + c2 u^2 + c3 u^3 with u
  // the position inside the cell; interpolation error <= h^4/384 max|d4H/dt4|,
  // about 2e-9 for 128 cells on t <= 6, far below FP32 round-off.
  const int K = kHTableCells;
  const double h = t_max / K;
  std::vector<float> tab(4 * K);
  for (int i = 0; i < K; ++i) {
    const double t0 = i * h, t1 = (i + 1) * h;
    const double f0 = gaussian_h_exact(t0), f1 = gaussian_h_exact(t1);
    const double d0 = h * gaussian_h_slope(t0), d1 = h * gaussian_h_slope(t1);
    tab[4 * i + 0] = static_cast<float>(f0);
    tab[4 * i + 1] = static_cast<float>(d0);
    tab[4 * i + 2] = static_cast<float>(3.0 * (f1 - f0) - 2.0 * d0 - d1);
    tab[4 * i + 3] = static_cast<float>(2.0 * (f0 - f1) + d0 + d1);
  }
  p.t_to_cell = static_cast<float>(static_cast<double>(p.c4) / h);
  return tab;
}

void walk_fp64(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass,
               const uint32_t* target_pos, const uint32_t* target_slot, const double* fx, const double* fy,
               const double* fz, const WalkOutputs& out, unsigned long long* counters, const double2* h64tab,
               hipStream_t s) {
  if (!t.records64 && t.node_count > 0) throw std::logic_error("FP64 walk on a tree without FP64 records");
  if (p.n_targets == 0) return;
  WalkData d{};
  d.link = t.link.as<NodeLink>();
  d.geo64 = t.geo64.as<NodeGeo64>();
  d.mom64 = t.mom64.as<NodeMom64>();
  d.quad64 = t.quad64.as<NodeQuad64>();
  d.order = t.order.as<uint32_t>();
  d.node_count_dev = device_node_count(t);
  d.src64 = src;
  d.uniform_mass = uniform_mass;
  d.tpos = target_pos;
  d.tslot = target_slot;
  d.f64x = fx;
  d.f64y = fy;
  d.f64z = fz;
  if (p.target_per_wave) {
    if (p.assignment == 0)
      walk64_target_kernel<0><<<grid_for(p.n_targets, kWalkWaves), kWalkThreads, 0, s>>>(p, d, out);
    else
      walk64_target_kernel<1><<<grid_for(p.n_targets, kWalkWaves), kWalkThreads, 0, s>>>(p, d, out);
    CHUI_LAUNCH_CHECK("walk64_target_kernel");
    return;
  }
  if (p.box64) {
    d.h64tab = h64tab;
    const unsigned grid = grid_for(p.n_targets, kWalkThreads);
    if (p.assignment == 0) walk64_box_kernel<0><<<grid, kWalkThreads, 0, s>>>(p, d, out, counters);
    else walk64_box_kernel<1><<<grid, kWalkThreads, 0, s>>>(p, d, out, counters);
    CHUI_LAUNCH_CHECK("walk64_box_kernel");
    return;
  }
  if (p.assignment == 0) launch64<0>(p, d, out, counters, s);
  else launch64<1>(p, d, out, counters, s);
}

void walk_fp64_adaptive(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass,
                        const double2* gtab, const double* prev, const uint32_t* target_pos,
                        const uint32_t* target_slot, const double* fx, const double* fy, const double* fz,
                        const WalkOutputs& out, unsigned long long* counters, hipStream_t s) {
  if (!t.records64 && t.node_count > 0) throw std::logic_error("FP64 walk on a tree without FP64 records");
  if (p.n_targets == 0) return;
  WalkData d{};
  d.link = t.link.as<NodeLink>();
  d.geo64 = t.geo64.as<NodeGeo64>();
  d.mom64 = t.mom64.as<NodeMom64>();
  d.quad64 = t.quad64.as<NodeQuad64>();
  d.order = t.order.as<uint32_t>();
  d.src64 = src;
  d.uniform_mass = uniform_mass;
  d.tpos = target_pos;
  d.tslot = target_slot;
  d.f64x = fx;
  d.f64y = fy;
  d.f64z = fz;
  d.gtab = gtab;
  d.prev = prev;
  const unsigned grid = grid_for(p.n_targets, kWalkThreads);
  if (p.assignment == 0) walk64_kernel<0, false, true><<<grid, kWalkThreads, 0, s>>>(p, d, out, counters);
  else walk64_kernel<1, false, true><<<grid, kWalkThreads, 0, s>>>(p, d, out, counters);
  CHUI_LAUNCH_CHECK("walk64_kernel (adaptive)");
}

void walk_fp64_group(const Tree& t, const WalkParams& p, const GroupParams& g, const double4* src,
                     int uniform_mass, const double2* gtab, const double4* gcoef, const double* prev,
                     const uint32_t* leaf_of_pos,
                     const uint32_t* target_pos, const uint32_t* target_slot, const double* fx, const double* fy,
                     const double* fz, const WalkOutputs& out, uint32_t* fallback_pos, uint32_t* fallback_slot,
                     uint32_t* fallback_count, unsigned long long* counters, hipStream_t s) {
  if (!t.records64 && t.node_count > 0) throw std::logic_error("FP64 walk on a tree without FP64 records");
  if (p.n_targets == 0) return;
  if (g.block == 0 || g.block > 256) throw std::invalid_argument("group walk residual block must be in [1, 256]");
  WalkData d{};
  d.link = t.link.as<NodeLink>();
  d.geo64 = t.geo64.as<NodeGeo64>();
  d.mom64 = t.mom64.as<NodeMom64>();
  d.quad64 = t.quad64.as<NodeQuad64>();
  d.order = t.order.as<uint32_t>();
  d.src64 = src;
  d.uniform_mass = uniform_mass;
  d.tpos = target_pos;
  d.tslot = target_slot;
  d.f64x = fx;
  d.f64y = fy;
  d.f64z = fz;
  d.gtab = gtab;
  d.gcoef = gcoef;
  d.prev = prev;
  const unsigned blocks = static_cast<unsigned>((static_cast<std::size_t>(p.n_targets) + g.block - 1) / g.block);
  if (p.assignment == 0)
    group_walk64_kernel<0><<<blocks, kWave, 0, s>>>(p, g, d, leaf_of_pos, out, fallback_pos, fallback_slot,
                                                    fallback_count, counters);
  else
    group_walk64_kernel<1><<<blocks, kWave, 0, s>>>(p, g, d, leaf_of_pos, out, fallback_pos, fallback_slot,
                                                    fallback_count, counters);
  CHUI_LAUNCH_CHECK("group_walk64_kernel");
}

void walk_mixed_group(const Tree& t, const WalkParams& p, const GroupParams& g, const double4* src64,
                      const uint4* src32, int uniform_mass, const float4* htab, const double* prev,
                      const uint32_t* leaf_of_pos, const uint32_t* target_pos, const uint32_t* target_slot,
                      const float* fx, const float* fy, const float* fz, const WalkOutputs& out,
                      uint32_t* fallback_pos, uint32_t* fallback_slot, uint32_t* fallback_count,
                      unsigned long long* counters, hipStream_t s) {
  if (!t.records64 && t.node_count > 0) throw std::logic_error("mixed group walk on a tree without FP64 records");
  if (p.n_targets == 0) return;
  if (g.block == 0 || g.block > 256) throw std::invalid_argument("group walk residual block must be in [1, 256]");
  if (!htab || !src32 || !src64) throw std::logic_error("mixed group walk without its source lanes or pair table");
  WalkData d{};
  d.link = t.link.as<NodeLink>();
  d.geo64 = t.geo64.as<NodeGeo64>();
  d.mom64 = t.mom64.as<NodeMom64>();
  d.quad64 = t.quad64.as<NodeQuad64>();
  d.order = t.order.as<uint32_t>();
  d.src64 = src64;
  d.src32 = src32;
  d.htab = htab;
  d.mom32 = t.mom32.as<NodeMom32>();
  d.quad32 = t.quad32.as<NodeQuad32>();
  d.uniform_mass = uniform_mass;
  d.tpos = target_pos;
  d.tslot = target_slot;
  d.f32x = fx;
  d.f32y = fy;
  d.f32z = fz;
  d.prev = prev;
  const unsigned blocks = static_cast<unsigned>((static_cast<std::size_t>(p.n_targets) + g.block - 1) / g.block);
  if (!t.mom32.ptr || !t.quad32.ptr) throw std::logic_error("mixed group walk on a tree without FP32 node records");
  auto go = [&](auto kernel) {
    kernel<<<blocks, kWave, 0, s>>>(p, g, d, leaf_of_pos, out, fallback_pos, fallback_slot, fallback_count, counters);
  };
  if (p.assignment == 0) {
    if (p.radial) go(group_walkmix_kernel<0, true>);
    else go(group_walkmix_kernel<0, false>);
  } else {
    if (p.radial) go(group_walkmix_kernel<1, true>);
    else go(group_walkmix_kernel<1, false>);
  }
  CHUI_LAUNCH_CHECK("group_walkmix_kernel");
}

void adaptive_mixed_finish(const Tree& t, const WalkParams& p, const uint4* src32, const uint32_t* pos,
                           const uint32_t* slot, uint32_t n, int identity, const double* sx, const double* sy,
                           const double* sz, const double* near64, const float* fx, const float* fy,
                           const float* fz, const WalkOutputs& out, hipStream_t s) {
  if (n == 0) return;
  if (out.near2 && !near64) throw std::logic_error("mixed adaptive finish: nearest-pair lane without its FP64 source");
  if (!identity && (!pos || !slot)) throw std::logic_error("mixed adaptive finish: target lists are null");
  WalkData d{};
  d.order = t.order.as<uint32_t>();
  d.src32 = src32;
  d.f32x = fx;
  d.f32y = fy;
  d.f32z = fz;
  const unsigned grid = grid_for(n, 256);
  if (p.assignment == 0)
    adaptive_mixed_finish_kernel<0><<<grid, 256, 0, s>>>(p, d, pos, slot, n, identity, sx, sy, sz, near64, out);
  else
    adaptive_mixed_finish_kernel<1><<<grid, 256, 0, s>>>(p, d, pos, slot, n, identity, sx, sy, sz, near64, out);
  CHUI_LAUNCH_CHECK("adaptive_mixed_finish_kernel");
}

void build_leaf_of_pos(const Tree& t, uint32_t* leaf_of_pos, hipStream_t s) {
  if (t.node_count == 0) return;
  leaf_of_pos_kernel<<<grid_for(t.node_count, 256), 256, 0, s>>>(t.link.as<NodeLink>(), t.node_count, leaf_of_pos);
  CHUI_LAUNCH_CHECK("leaf_of_pos_kernel");
}

std::vector<double> gaussian_hermite_knots() {
  // tree_pm_split_kernel.hpp:159-170: H(t_i) from the closed form (series at
  // t = 0) and step * dH/dt(t_i), t_i = i / 64, interleaved (value, slope).
  constexpr double c = 2.25675833419102514779231780624309034;  // 4 / sqrt(pi)
  constexpr double step = 64.0 / kGaussCells;
  std::vector<double> k(2 * (kGaussCells + 1));
  for (int i = 0; i <= kGaussCells; ++i) {
    const double t = static_cast<double>(i) * step;
    const double q = std::sqrt(t);
    double h, dh;
    if (i == 0) {
      h = c * (1.0 / 3.0);
      dh = -4.0 * kInvSqrtPi / 5.0;
    } else {
      const double qq = (0.5 * q) / 0.5;  // L(q) at split scale 1/2
      const double lf = std::erf(qq) - 2.0 * kInvSqrtPi * qq * std::exp(-qq * qq);
      h = lf / (t * q);
      dh = (2.0 * kInvSqrtPi * std::exp(-t) - 1.5 * h) / t;
    }
    k[2 * i] = h;
    k[2 * i + 1] = step * dh;
  }
  return k;
}

std::vector<double> gaussian_hermite_monomials(const std::vector<double>& knots) {
  // the cubic Hermite cell i on u in [0, 1) as a0 + u (a1 + u (a2 + u a3))
  std::vector<double> a(4 * kGaussCells);
  for (int i = 0; i < kGaussCells; ++i) {
    const double v0 = knots[2 * i], s0 = knots[2 * i + 1], v1 = knots[2 * i + 2], s1 = knots[2 * i + 3];
    a[4 * i] = v0;
    a[4 * i + 1] = s0;
    a[4 * i + 2] = 3.0 * (v1 - v0) - 2.0 * s0 - s1;
    a[4 * i + 3] = 2.0 * (v0 - v1) + s0 + s1;
  }
  return a;
}

void walk_fp64_vjp(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass, const double4* lam,
                   const double* fx, const double* fy, const double* fz, const double* nu, const WalkOutputs& out,
                   unsigned long long* counters, const double2* h64tab, const VjpNodeTerms* nt, hipStream_t s) {
  if (!t.records64 && t.node_count > 0) throw std::logic_error("FP64 walk on a tree without FP64 records");
  if (p.n_targets == 0) return;
  WalkData d{};
  d.link = t.link.as<NodeLink>();
  d.geo64 = t.geo64.as<NodeGeo64>();
  d.mom64 = t.mom64.as<NodeMom64>();
  d.quad64 = t.quad64.as<NodeQuad64>();
  d.order = t.order.as<uint32_t>();
  d.node_count_dev = device_node_count(t);
  d.src64 = src;
  d.uniform_mass = uniform_mass;
  d.f64x = fx;
  d.f64y = fy;
  d.f64z = fz;
  d.lam64 = lam;
  d.nu64 = nu;
  if (p.quad) {
    if (!nt || !p.box64) throw std::logic_error("force VJP with accepted nodes needs the node-term buffers");
    d.vjp_node_acc = nt->node_acc;
    d.vjp_accepted = nt->accepted;
    d.vjp_accepted_count = nt->count;
    d.vjp_accepted_cap = nt->cap;
  }
  if (p.box64) {
    d.h64tab = h64tab;
    const unsigned grid = grid_for(p.n_targets, kWalkThreads);
    if (p.assignment == 0) walk64_box_kernel<0, true><<<grid, kWalkThreads, 0, s>>>(p, d, out, counters);
    else walk64_box_kernel<1, true><<<grid, kWalkThreads, 0, s>>>(p, d, out, counters);
    CHUI_LAUNCH_CHECK("walk64_box_kernel (JVP)");
    return;
  }
  if (p.assignment == 0) launch64<0, true>(p, d, out, counters, s);
  else launch64<1, true>(p, d, out, counters, s);
}

void vjp_node_corrections(const Tree& t, const WalkParams& p, const double4* src, int uniform_mass,
                          double uniform_mass_value, const double4* lam, const VjpNodeTerms& nt, unsigned int count,
                          const WalkOutputs& out, hipStream_t s, const unsigned int* count_dev,
                          unsigned long long* fault) {
  const std::size_t n = t.n;
  if (n == 0) return;
  if (count_dev) {
    vjp_count_check_kernel<<<1, 64, 0, s>>>(count_dev, count, fault);
    CHUI_LAUNCH_CHECK("vjp_count_check_kernel");
  }
  if (count > 0) {
    radix_sort_pairs(reinterpret_cast<uint64_t*>(nt.accepted), nt.vals, reinterpret_cast<uint64_t*>(nt.accepted_alt),
                     nt.vals_alt, count, 0, 64, nt.sort_scratch, s);
    vjp_pair_fix_kernel<<<grid_for(count, 256), 256, 0, s>>>(nt.accepted, count, count_dev, t.link.as<NodeLink>(),
                                                             src, lam, uniform_mass, uniform_mass_value, p, nt.corr);
    CHUI_LAUNCH_CHECK("vjp_pair_fix_kernel");
    vjp_node_push_kernel<<<grid_for(t.node_count, 256), 256, 0, s>>>(
        static_cast<uint32_t>(t.node_count), nt.node_acc, t.link.as<NodeLink>(), t.mom64.as<NodeMom64>(), src,
        uniform_mass, uniform_mass_value, p, nt.corr);
    CHUI_LAUNCH_CHECK("vjp_node_push_kernel");
  }
  vjp_apply_kernel<<<grid_for(n, 256), 256, 0, s>>>(static_cast<uint32_t>(n), t.order.as<uint32_t>(), nt.corr,
                                                    static_cast<double*>(out.ax), static_cast<double*>(out.ay),
                                                    static_cast<double*>(out.az));
  CHUI_LAUNCH_CHECK("vjp_apply_kernel");
}

void walk_mixed(const Tree& t, const WalkParams& p, const uint4* src, int uniform_mass, const float4* htab,
                const float4* ltab, const uint32_t* target_pos, const uint32_t* target_slot, const float* fx,
                const float* fy, const float* fz, const WalkOutputs& out, unsigned long long* counters,
                hipStream_t s) {
  if (p.n_targets == 0) return;
  WalkData d{};
  d.link = t.link.as<NodeLink>();
  d.geo32 = t.geo32.as<NodeGeo32>();
  d.mom32 = t.mom32.as<NodeMom32>();
  d.quad32 = t.quad32.as<NodeQuad32>();
  d.box32 = t.box32.as<NodeBox32>();
  d.order = t.order.as<uint32_t>();
  d.src32 = src;
  d.htab = htab;
  d.ltab = ltab;
  d.child = t.child_index.as<uint32_t>();
  d.uniform_mass = uniform_mass;
  d.tpos = target_pos;
  d.tslot = target_slot;
  d.f32x = fx;
  d.f32y = fy;
  d.f32z = fz;
  d.node_count_dev = device_node_count(t);
  const bool cic = p.assignment == 0;
  if (p.cubic) {
    if (cic) launchmix<0, true>(p, d, out, counters, s);
    else launchmix<1, true>(p, d, out, counters, s);
  } else {
    if (cic) launchmix<0, false>(p, d, out, counters, s);
    else launchmix<1, false>(p, d, out, counters, s);
  }
}

}  // namespace chui

// treepm_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of CHUI's periodic TreePM force evaluation, used as the
// parity checker of the MI355X library and as the timed CPU baseline. Nothing in
// the product links, loads or calls this file; only tests/, the smoke check and
// the CPU-baseline leg of bench.py do.
//
// Each function names the reference lines whose behaviour it restates. Parity of
// this restatement is pinned in tests/ by checks that do not retype its formulas:
//   unwrap_axis      brute-force O(n^2) largest-circular-gap search on tiny inputs
//   morton keys      bit-by-bit interleave loop + std::stable_sort
//   build_tree       invariant validator (cube membership with the >= rule, stable
//                    partition, leaf rule, Morton order inside leaves, preorder)
//   multipoles       direct sums over each node's particles (long double)
//   split kernel     erf/erfc closed forms (math.erf), complementarity, series limit
//   short-range walk brute-force minimum-image direct sum within r_cut (exact for
//                    the strict policy when no multipole is accepted)
//   PM               numpy.fft.rfftn Poisson solve, single-mode analytic force,
//                    CIC mass conservation
//   Ewald            Madelung-type lattice sums and Newtonian limit checks
//   KDK factors      scipy.integrate.quad of 1/(a^2 H)
#include "../include/chui_treepm.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kInvSqrtPi = 0.564189583547756286948079451560772586;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

thread_local std::string g_error;

// ---------------------------------------------------------------------------
// Periodic unwrap: tree_pm_coupling.cpp:141-218 (unwrapPeriodicAxis).
// Wrap into [0,L) (skipped for canonical inputs), sort the IEEE bits, anchor the
// interval right after the largest circular gap (ties -> smaller anchor), and
// add L to every coordinate below the anchor.
double unwrap_axis(const double* in, size_t n, double L, double* out, bool canonical) {
  if (!(std::isfinite(L) && L > 0.0)) throw std::invalid_argument("unwrap requires positive L");
  if (n == 0) return 0.0;
  std::vector<uint64_t> bits(n);
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(in[i])) throw std::invalid_argument("non-finite coordinate");
    double v = canonical ? in[i] : in[i] - L * std::floor(in[i] / L);
    if (v >= L) v = 0.0;
    else if (v < 0.0) v += L;
    if (v == 0.0) v = 0.0;
    out[i] = v;
    std::memcpy(&bits[i], &v, 8);
  }
  std::sort(bits.begin(), bits.end());
  auto as_double = [](uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; };
  double anchor = as_double(bits[0]);
  if (n > 1) {
    double largest = -1.0, best = anchor;
    for (size_t i = 0; i < n; ++i) {
      const size_t nx = (i + 1) % n;
      const double cur = as_double(bits[i]);
      const double cand = as_double(bits[nx]);
      const double next = nx == 0 ? cand + L : cand;
      const double gap = next - cur;
      if (gap > largest || (gap == largest && cand < best)) { largest = gap; best = cand; }
    }
    anchor = best;
  }
  for (size_t i = 0; i < n; ++i)
    if (out[i] < anchor) out[i] += L;
  return anchor;
}

// ---------------------------------------------------------------------------
// Morton ordering: tree_ordering.cpp:73-88 (expandBits21/morton3D), :93-127
// (computeTreeBounds), :128-173 (21-bit llround quantisation), :13-71 (stable sort).
static inline uint64_t expand21(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x1F00000000FFFFULL;
  x = (x | (x << 16)) & 0x1F0000FF0000FFULL;
  x = (x | (x << 8)) & 0x100F00F00F00F00FULL;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}

struct Bounds { double lo[3], hi[3]; };

Bounds bounds_of(const double* x, const double* y, const double* z, size_t n) {
  Bounds b;
  for (int a = 0; a < 3; ++a) { b.lo[a] = INFINITY; b.hi[a] = -INFINITY; }
  const double* p[3] = {x, y, z};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) { b.lo[a] = std::min(b.lo[a], p[a][i]); b.hi[a] = std::max(b.hi[a], p[a][i]); }
  if (n == 0) for (int a = 0; a < 3; ++a) { b.lo[a] = 0.0; b.hi[a] = 0.0; }
  return b;
}

void morton_keys(const double* x, const double* y, const double* z, size_t n, uint64_t* keys) {
  if (n == 0) return;
  const Bounds b = bounds_of(x, y, z, n);
  const double ext = std::max({b.hi[0] - b.lo[0], b.hi[1] - b.lo[1], b.hi[2] - b.lo[2]});
  const double extent = std::max(ext, 1.0e-12);
  const double grid = static_cast<double>((1u << 21) - 1u);
  for (size_t i = 0; i < n; ++i) {
    const double nx = std::clamp((x[i] - b.lo[0]) / extent, 0.0, 1.0);
    const double ny = std::clamp((y[i] - b.lo[1]) / extent, 0.0, 1.0);
    const double nz = std::clamp((z[i] - b.lo[2]) / extent, 0.0, 1.0);
    const uint32_t qx = static_cast<uint32_t>(std::llround(nx * grid));
    const uint32_t qy = static_cast<uint32_t>(std::llround(ny * grid));
    const uint32_t qz = static_cast<uint32_t>(std::llround(nz * grid));
    keys[i] = (expand21(qx) << 2) | (expand21(qy) << 1) | expand21(qz);
  }
}

void morton_order(const double* x, const double* y, const double* z, size_t n,
                  uint64_t* keys_sorted, uint32_t* order) {
  std::vector<uint64_t> keys(n);
  morton_keys(x, y, z, n, keys.data());
  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; ++i) idx[i] = static_cast<uint32_t>(i);
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  for (size_t i = 0; i < n; ++i) { order[i] = idx[i]; keys_sorted[i] = keys[idx[i]]; }
}

// ---------------------------------------------------------------------------
// Octree: tree_gravity.cpp:343-524 (root cube), :1081-1206 (buildNodeRecursive,
// stable octant partition, leaf rule), :1208-1380 (accumulateMultipoles).
struct Tree {
  std::vector<double> cx, cy, cz, half, mass, comx, comy, comz;
  std::vector<double> qxx, qxy, qxz, qyy, qyz, qzz, trace, smin, smax;
  std::vector<uint32_t> child_base, child_index, begin, count, subtree;
  std::vector<uint8_t> child_count;
  std::vector<uint32_t> order;  // sorted_particle_index after build
  uint32_t max_depth = 0;
  size_t size() const { return cx.size(); }
};

static inline uint8_t octant_of(double x, double y, double z, double cx, double cy, double cz) {
  return static_cast<uint8_t>(((x >= cx ? 1 : 0) << 2) | ((y >= cy ? 1 : 0) << 1) | (z >= cz ? 1 : 0));
}

struct Builder {
  const double *x, *y, *z;
  size_t leaf;
  Tree& t;
  std::vector<uint32_t> scratch;
  uint32_t build(uint32_t b, uint32_t e, double cx, double cy, double cz, double h, uint32_t depth) {
    if (depth > 64) throw std::overflow_error("Tree gravity maximum depth contract exceeded");
    t.max_depth = std::max(t.max_depth, depth);
    const uint32_t node = static_cast<uint32_t>(t.size());
    t.cx.push_back(cx); t.cy.push_back(cy); t.cz.push_back(cz); t.half.push_back(h);
    t.mass.push_back(0.0); t.comx.push_back(cx); t.comy.push_back(cy); t.comz.push_back(cz);
    t.qxx.push_back(0); t.qxy.push_back(0); t.qxz.push_back(0); t.qyy.push_back(0); t.qyz.push_back(0);
    t.qzz.push_back(0); t.trace.push_back(0);
    t.smin.push_back(INFINITY); t.smax.push_back(0.0);
    t.child_base.push_back(0); t.child_count.push_back(0);
    for (int i = 0; i < 8; ++i) t.child_index.push_back(kInvalid);
    t.begin.push_back(b); t.count.push_back(e - b); t.subtree.push_back(1);
    if ((e - b) <= leaf || h <= 1.0e-12) return node;
    std::array<uint32_t, 8> cnt{};
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t p = t.order[i];
      ++cnt[octant_of(x[p], y[p], z[p], cx, cy, cz)];
    }
    std::array<uint32_t, 9> off{};
    off[0] = b;
    for (int i = 0; i < 8; ++i) off[i + 1] = off[i] + cnt[i];
    std::array<uint32_t, 8> cur{};
    for (int i = 0; i < 8; ++i) cur[i] = off[i];
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t p = t.order[i];
      scratch[cur[octant_of(x[p], y[p], z[p], cx, cy, cz)]++] = p;
    }
    std::copy(scratch.begin() + b, scratch.begin() + e, t.order.begin() + b);
    const uint32_t base = static_cast<uint32_t>(t.size());
    uint8_t nonempty = 0;
    for (int o = 0; o < 8; ++o) {
      if (off[o] == off[o + 1]) continue;
      const double ch = 0.5 * h;
      const double ccx = cx + ((o & 4) ? ch : -ch);
      const double ccy = cy + ((o & 2) ? ch : -ch);
      const double ccz = cz + ((o & 1) ? ch : -ch);
      const uint32_t c = build(off[o], off[o + 1], ccx, ccy, ccz, ch, depth + 1);
      t.child_index[static_cast<size_t>(node) * 8 + o] = c;
      ++nonempty;
    }
    t.child_base[node] = base;
    t.child_count[node] = nonempty;
    t.subtree[node] = static_cast<uint32_t>(t.size()) - node;
    return node;
  }
};

void accumulate(Tree& t, const double* x, const double* y, const double* z, const double* m,
                uint32_t node, bool quad, double eps) {
  if (t.child_count[node] == 0) {
    const uint32_t b = t.begin[node], e = b + t.count[node];
    double tm = 0, wx = 0, wy = 0, wz = 0;
    t.smin[node] = eps; t.smax[node] = eps;
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t p = t.order[i];
      const double mm = m[p];
      tm += mm; wx += mm * x[p]; wy += mm * y[p]; wz += mm * z[p];
    }
    t.mass[node] = tm;
    if (tm > 0.0) { t.comx[node] = wx / tm; t.comy[node] = wy / tm; t.comz[node] = wz / tm; }
    if (quad && tm > 0.0) {
      const double cx = t.comx[node], cy = t.comy[node], cz = t.comz[node];
      double ixx = 0, ixy = 0, ixz = 0, iyy = 0, iyz = 0, izz = 0;
      for (uint32_t i = b; i < e; ++i) {
        const uint32_t p = t.order[i];
        const double mm = m[p];
        const double rx = x[p] - cx, ry = y[p] - cy, rz = z[p] - cz;
        ixx += mm * rx * rx; ixy += mm * rx * ry; ixz += mm * rx * rz;
        iyy += mm * ry * ry; iyz += mm * ry * rz; izz += mm * rz * rz;
      }
      const double tr = ixx + iyy + izz;
      t.trace[node] = tr;
      t.qxx[node] = 3.0 * ixx - tr; t.qxy[node] = 3.0 * ixy; t.qxz[node] = 3.0 * ixz;
      t.qyy[node] = 3.0 * iyy - tr; t.qyz[node] = 3.0 * iyz; t.qzz[node] = 3.0 * izz - tr;
    }
    return;
  }
  double tm = 0, wx = 0, wy = 0, wz = 0;
  const size_t co = static_cast<size_t>(node) * 8;
  for (int o = 0; o < 8; ++o) {
    const uint32_t c = t.child_index[co + o];
    if (c == kInvalid) continue;
    accumulate(t, x, y, z, m, c, quad, eps);
    t.smin[node] = std::min(t.smin[node], t.smin[c]);
    t.smax[node] = std::max(t.smax[node], t.smax[c]);
    tm += t.mass[c];
    wx += t.mass[c] * t.comx[c]; wy += t.mass[c] * t.comy[c]; wz += t.mass[c] * t.comz[c];
  }
  t.mass[node] = tm;
  if (tm > 0.0) { t.comx[node] = wx / tm; t.comy[node] = wy / tm; t.comz[node] = wz / tm; }
  if (!quad || tm <= 0.0) return;
  const double cx = t.comx[node], cy = t.comy[node], cz = t.comz[node];
  double qxx = 0, qxy = 0, qxz = 0, qyy = 0, qyz = 0, qzz = 0, tr = 0;
  for (int o = 0; o < 8; ++o) {
    const uint32_t c = t.child_index[co + o];
    if (c == kInvalid) continue;
    const double dx = t.comx[c] - cx, dy = t.comy[c] - cy, dz = t.comz[c] - cz;
    const double mm = t.mass[c];
    const double d2 = dx * dx + dy * dy + dz * dz;
    qxx += t.qxx[c] + mm * (3.0 * dx * dx - d2);
    qxy += t.qxy[c] + mm * (3.0 * dx * dy);
    qxz += t.qxz[c] + mm * (3.0 * dx * dz);
    qyy += t.qyy[c] + mm * (3.0 * dy * dy - d2);
    qyz += t.qyz[c] + mm * (3.0 * dy * dz);
    qzz += t.qzz[c] + mm * (3.0 * dz * dz - d2);
    tr += t.trace[c] + mm * d2;
  }
  t.qxx[node] = qxx; t.qxy[node] = qxy; t.qxz[node] = qxz;
  t.qyy[node] = qyy; t.qyz[node] = qyz; t.qzz[node] = qzz; t.trace[node] = tr;
}

void build_tree(const double* x, const double* y, const double* z, const double* m, size_t n,
                size_t leaf, bool quad, double eps, Tree& t) {
  t = Tree{};
  if (n > 0xFFFFFFFEull) throw std::overflow_error("Tree gravity source count exceeds the 32-bit tree-index contract");
  std::vector<uint64_t> keys(n);
  t.order.resize(n);
  morton_order(x, y, z, n, keys.data(), t.order.data());
  if (n == 0) return;
  const Bounds b = bounds_of(x, y, z, n);
  const double ext = std::max({b.hi[0] - b.lo[0], b.hi[1] - b.lo[1], b.hi[2] - b.lo[2]});
  const double max_extent = std::max(ext, 1.0e-10);
  const double cx = 0.5 * (b.lo[0] + b.hi[0]);
  const double cy = 0.5 * (b.lo[1] + b.hi[1]);
  const double cz = 0.5 * (b.lo[2] + b.hi[2]);
  Builder bld{x, y, z, leaf, t, std::vector<uint32_t>(n)};
  bld.build(0, static_cast<uint32_t>(n), cx, cy, cz, 0.5 * max_extent * (1.0 + 1.0e-8), 0);
  accumulate(t, x, y, z, m, 0, quad, eps);
}

// ---------------------------------------------------------------------------
// Split kernel: tree_pm_split_kernel.hpp:66-128 and :204-228, softening
// tree_softening.hpp:243-255.
static inline double series_h(double t) {
  constexpr double c = 2.25675833419102514779231780624309034;
  return c * (1.0 / 3.0 + t * (-1.0 / 5.0 + t * (1.0 / 14.0 + t * (-1.0 / 54.0 + t * (1.0 / 264.0 +
         t * (-1.0 / 1560.0 + t * (1.0 / 10800.0 - t / 85680.0)))))));
}
double short_range_factor(double r, double a) {
  if (r <= 0.0) return 1.0;
  const double q = (0.5 * r) / a;
  if (std::isinf(q)) return 0.0;
  return std::erfc(q) + 2.0 * kInvSqrtPi * q * std::exp(-q * q);
}
double long_range_factor(double r, double a) {
  if (r <= 0.0) return 0.0;
  const double q = (0.5 * r) / a;
  if (std::isinf(q)) return 1.0;
  if (q < 0.125) {
    const double t = q * q;
    const double h = 4.0 * kInvSqrtPi * (1.0 / 3.0 + t * (-1.0 / 5.0 + t * (1.0 / 14.0 + t * (-1.0 / 54.0 +
        t * (1.0 / 264.0 + t * (-1.0 / 1560.0 + t * (1.0 / 10800.0 - t / 85680.0)))))));
    return q * q * q * h;
  }
  return std::erf(q) - 2.0 * kInvSqrtPi * q * std::exp(-q * q);
}
double long_range_inv_r3(double r, double a) {
  const double q = (0.5 * r) / a;
  if (q < 0.125) return ((0.125 * series_h(q * q) / a) / a) / a;
  return ((long_range_factor(r, a) / r) / r) / r;
}
double softened_inv_r3(double r2, double eps) {
  const double s = r2 + eps * eps;
  if (s <= 0.0) return 0.0;
  return 1.0 / (s * std::sqrt(s));
}
double residual_inv_r3(double r2, double eps, double a) {
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
  return softened_inv_r3(r2, eps) - long_range_inv_r3(std::sqrt(r2), a);
}

// ---------------------------------------------------------------------------
// Short-range residual walk: tree_pm_coupling.cpp:4026-4272 (per-target DFS),
// acceptance internal/tree_pm_acceptance.hpp:54-139, MAC
// internal/tree_interaction_common.hpp:36-89, geometry internal/tree_pm_geometry.hpp,
// accepted-node kernel tree_pm_coupling.cpp:743-827.
struct WalkParams {
  double L[3];
  double G, a, rcut, rcut2, theta, eps;
  int criterion;
  bool quad;
};

static inline double min_image(double d, double L) { return d - L * std::nearbyint(d / L); }

static inline bool within_cutoff(double d2, double c, double c2) {
  if (!std::isnormal(c2) || !std::isfinite(d2) ||
      std::abs(d2 - c2) <= 8.0 * std::numeric_limits<double>::epsilon() * c2)
    return std::sqrt(d2) <= c;
  return d2 <= c2;
}

std::array<double, 3> node_accel(const Tree& t, uint32_t n, double dx, double dy, double dz,
                                 const WalkParams& P) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double safe_r2 = std::max(r2, 1.0e-30);
  const double r = std::sqrt(safe_r2);
  const double pe = std::max(t.smax[n], P.eps);
  const double eps2 = pe * pe;
  const double denom = std::max(r2 + eps2, 1.0e-30);
  const double soft = 1.0 / (denom * std::sqrt(denom));
  const double split = short_range_factor(r, P.a);
  const double newton = 1.0 / (safe_r2 * r);
  const double radial = soft + (split - 1.0) * newton;
  const double pref = P.G;
  double ax = pref * t.mass[n] * radial * dx;
  double ay = pref * t.mass[n] * radial * dy;
  double az = pref * t.mass[n] * radial * dz;
  if (!P.quad) return {ax, ay, az};
  const double tr = t.trace[n];
  const double mxx = (t.qxx[n] + tr) / 3.0, mxy = t.qxy[n] / 3.0, mxz = t.qxz[n] / 3.0;
  const double myy = (t.qyy[n] + tr) / 3.0, myz = t.qyz[n] / 3.0, mzz = (t.qzz[n] + tr) / 3.0;
  const double mdx = mxx * dx + mxy * dy + mxz * dz;
  const double mdy = mxy * dx + myy * dy + myz * dz;
  const double mdz = mxz * dx + myz * dy + mzz * dz;
  const double dmd = dx * mdx + dy * mdy + dz * mdz;
  const double a = P.a, a2 = a * a;
  const double ex = std::exp(-r2 / (4.0 * a2));
  const double s1 = -0.5 * kInvSqrtPi * r2 * ex / (a * a2);
  const double s2 = kInvSqrtPi * ex * (-r / (a * a2) + 0.25 * r * r2 / (a * a2 * a2));
  const double f1 = -3.0 * r / (denom * denom * std::sqrt(denom));
  const double f2 = -3.0 / (denom * denom * std::sqrt(denom)) + 15.0 * r2 / (denom * denom * denom * std::sqrt(denom));
  const double n1 = -3.0 / (safe_r2 * safe_r2);
  const double n2 = 12.0 / (safe_r2 * safe_r2 * r);
  const double rf = f1 + s1 * newton + (split - 1.0) * n1;
  const double rs = f2 + s2 * newton + 2.0 * s1 * n1 + (split - 1.0) * n2;
  const double fo = rf / r;
  const double cr = rs / r2 - rf / (r2 * r);
  ax += pref * (mdx * fo + 0.5 * tr * dx * fo + 0.5 * dx * dmd * cr);
  ay += pref * (mdy * fo + 0.5 * tr * dy * fo + 0.5 * dy * dmd * cr);
  az += pref * (mdz * fo + 0.5 * tr * dz * fo + 0.5 * dz * dmd * cr);
  return {ax, ay, az};
}

struct WalkCounters {
  uint64_t visited = 0, pruned = 0, pairs = 0, multipoles = 0, leaves = 0, skipped_pairs = 0;
};

// Strict-reference acceptance of an internal node (tree_pm_acceptance.hpp:54-139
// with full_mac_diagnostics = true, no acceleration history).
static inline bool accept_internal(const Tree& t, uint32_t n, bool inside, bool contained, double r2,
                                   const WalkParams& P) {
  const double h = t.half[n];
  const double r = std::sqrt(std::max(r2, 1.0e-30));
  // uniform softening: smax - smin <= 1e-12 always passes the envelope guard
  bool soft_ok = true;
  if (!(t.smax[n] - t.smin[n] <= 1.0e-12)) soft_ok = r > (2.0 * h + 2.0 * std::max(t.smax[n], P.eps));
  const bool env_ok = P.quad && (2.0 * h / r) < 0.08;
  double offset = 0.0;
  if (P.criterion != CHUI_OPEN_GEOMETRIC) {
    const double ox = t.cx[n] - t.comx[n], oy = t.cy[n] - t.comy[n], oz = t.cz[n] - t.comz[n];
    offset = std::sqrt(ox * ox + oy * oy + oz * oz);
  }
  bool mac = false;
  if (!inside) {
    if (P.criterion == CHUI_OPEN_GEOMETRIC) mac = (2.0 * h / std::sqrt(r2 + 1.0e-30)) < P.theta;
    else mac = ((2.0 * h + offset) / std::sqrt(r2 + 1.0e-30)) < P.theta;
  }
  return mac && env_ok && soft_ok && contained && !inside && std::isfinite(r2);
}

std::array<double, 3> walk_target(const Tree& t, const double* sx, const double* sy, const double* sz,
                                  const double* sm, double px, double py, double pz, uint32_t self,
                                  const WalkParams& P, std::vector<uint32_t>& stack, WalkCounters& C) {
  double ax = 0, ay = 0, az = 0;
  if (t.size() == 0) return {0, 0, 0};
  stack.clear();
  stack.push_back(0);
  while (!stack.empty()) {
    const uint32_t n = stack.back();
    stack.pop_back();
    ++C.visited;
    const double h = t.half[n];
    const double ndx = min_image(t.cx[n] - px, P.L[0]);
    const double ndy = min_image(t.cy[n] - py, P.L[1]);
    const double ndz = min_image(t.cz[n] - pz, P.L[2]);
    const double ax_ = std::abs(ndx), ay_ = std::abs(ndy), az_ = std::abs(ndz);
    const double ex = std::max(0.0, ax_ - h), ey = std::max(0.0, ay_ - h), ez = std::max(0.0, az_ - h);
    if (!within_cutoff(ex * ex + ey * ey + ez * ez, P.rcut, P.rcut2)) { ++C.pruned; continue; }
    const bool leaf = t.child_count[n] == 0;
    double dx = 0, dy = 0, dz = 0;
    bool accept = leaf;
    if (!leaf) {
      const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
      const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
      const bool contained = within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
      dx = min_image(t.comx[n] - px, P.L[0]);
      dy = min_image(t.comy[n] - py, P.L[1]);
      dz = min_image(t.comz[n] - pz, P.L[2]);
      accept = accept_internal(t, n, inside, contained, dx * dx + dy * dy + dz * dz, P);
    }
    if (accept) {
      if (leaf) {
        ++C.leaves;
        const uint32_t b = t.begin[n], e = b + t.count[n];
        for (uint32_t i = b; i < e; ++i) {
          if (t.order[i] == self) continue;
          const double qx = min_image(sx[i] - px, P.L[0]);
          const double qy = min_image(sy[i] - py, P.L[1]);
          const double qz = min_image(sz[i] - pz, P.L[2]);
          const double r2 = qx * qx + qy * qy + qz * qz;
          if (r2 > P.rcut2) { ++C.skipped_pairs; continue; }
          ++C.pairs;
          if (qx == 0.0 && qy == 0.0 && qz == 0.0) continue;
          const double f = residual_inv_r3(r2, P.eps, P.a) * P.G;
          ax += f * sm[i] * qx; ay += f * sm[i] * qy; az += f * sm[i] * qz;
        }
      } else {
        ++C.multipoles;
        const auto c = node_accel(t, n, dx, dy, dz, P);
        ax += c[0]; ay += c[1]; az += c[2];
      }
    } else {
      const size_t co = static_cast<size_t>(n) * 8;
      for (int o = 8; o > 0; --o) {
        const uint32_t c = t.child_index[co + o - 1];
        if (c != kInvalid) stack.push_back(c);
      }
    }
  }
  return {ax, ay, az};
}

// ---------------------------------------------------------------------------
// FFT (FP64, own implementation; FFTW is absent). Radix-2 iterative for powers
// of two, direct DFT otherwise. sign = -1 forward (FFTW_FORWARD), +1 backward.
void fft1d(std::complex<double>* d, size_t n, int sign) {
  if (n <= 1) return;
  if ((n & (n - 1)) == 0) {
    for (size_t i = 1, j = 0; i < n; ++i) {
      size_t bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) std::swap(d[i], d[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
      const size_t h = len / 2;
      for (size_t k = 0; k < h; ++k) {
        const double ang = sign * 2.0 * kPi * static_cast<double>(k) / static_cast<double>(len);
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        for (size_t s = 0; s < n; s += len) {
          const std::complex<double> u = d[s + k], v = d[s + k + h] * w;
          d[s + k] = u + v;
          d[s + k + h] = u - v;
        }
      }
    }
    return;
  }
  std::vector<std::complex<double>> out(n);
  for (size_t k = 0; k < n; ++k) {
    std::complex<double> acc = 0;
    for (size_t j = 0; j < n; ++j) {
      const double ang = sign * 2.0 * kPi * static_cast<double>((k * j) % n) / static_cast<double>(n);
      acc += d[j] * std::complex<double>(std::cos(ang), std::sin(ang));
    }
    out[k] = acc;
  }
  std::copy(out.begin(), out.end(), d);
}

void fft3d(std::vector<std::complex<double>>& a, size_t nx, size_t ny, size_t nz, int sign) {
#pragma omp parallel
  {
    std::vector<std::complex<double>> line(std::max({nx, ny, nz}));
#pragma omp for collapse(2)
    for (size_t ix = 0; ix < nx; ++ix)
      for (size_t iy = 0; iy < ny; ++iy) fft1d(&a[(ix * ny + iy) * nz], nz, sign);
#pragma omp for collapse(2)
    for (size_t ix = 0; ix < nx; ++ix)
      for (size_t iz = 0; iz < nz; ++iz) {
        for (size_t iy = 0; iy < ny; ++iy) line[iy] = a[(ix * ny + iy) * nz + iz];
        fft1d(line.data(), ny, sign);
        for (size_t iy = 0; iy < ny; ++iy) a[(ix * ny + iy) * nz + iz] = line[iy];
      }
#pragma omp for collapse(2)
    for (size_t iy = 0; iy < ny; ++iy)
      for (size_t iz = 0; iz < nz; ++iz) {
        for (size_t ix = 0; ix < nx; ++ix) line[ix] = a[(ix * ny + iy) * nz + iz];
        fft1d(line.data(), nx, sign);
        for (size_t ix = 0; ix < nx; ++ix) a[(ix * ny + iy) * nz + iz] = line[ix];
      }
  }
}

// ---------------------------------------------------------------------------
// PM: pm_solver.cpp:103-115 (wrapIndex/wrapPosition), :925-951 (stencils),
// :4016-4061 + :4544-4553 (deposit, /cell_volume), :4676-4712 (mean, forward FFT),
// :2398-2466 + :2557-2656 (k metadata, window, Gaussian filter, -4 pi G / k^2),
// :2658-2705 (-i k_a, Nyquist zero), :4865-4924 (inverse / M), :5900-5957 (interp).
struct Stencil { long off[3]; double w[3]; int count; };

static inline Stencil stencil_of(double g, int scheme) {
  Stencil s{};
  if (scheme == CHUI_ASSIGN_CIC) {
    const double base = std::floor(g);
    const long i0 = static_cast<long>(base);
    const double t = g - base;
    s.off[0] = i0; s.off[1] = i0 + 1; s.w[0] = 1.0 - t; s.w[1] = t; s.count = 2;
  } else {
    const double c = std::floor(g + 0.5);
    const long ic = static_cast<long>(c);
    const double d = g - c;
    s.off[0] = ic - 1; s.off[1] = ic; s.off[2] = ic + 1;
    s.w[0] = 0.5 * (0.5 - d) * (0.5 - d); s.w[1] = 0.75 - d * d; s.w[2] = 0.5 * (0.5 + d) * (0.5 + d);
    s.count = 3;
  }
  return s;
}
static inline size_t wrap_index(long i, size_t n) {
  long w = i % static_cast<long>(n);
  if (w < 0) w += static_cast<long>(n);
  return static_cast<size_t>(w);
}
static inline double wrap_position(double x, double L) {
  const double w = std::fmod(x, L);
  return w < 0.0 ? w + L : w;
}
static inline double sinc(double x) { return std::abs(x) < 1.0e-12 ? 1.0 : std::sin(x) / x; }

struct PmField {
  size_t nx = 0, ny = 0, nz = 0;
  std::vector<double> fx, fy, fz;  // real-space force meshes, row-major (ix*ny+iy)*nz+iz
};

void pm_mesh_solve(const std::vector<double>& rho, const chui_treepm_options& o, const int mesh[3],
                   double split_scale, PmField& F);

void pm_solve(const double* x, const double* y, const double* z, const double* m, size_t n,
              const chui_treepm_options& o, const int mesh[3], double split_scale, PmField& F,
              std::vector<double>* density_out = nullptr) {
  const size_t nx = mesh[0], ny = mesh[1], nz = mesh[2];
  const size_t M = nx * ny * nz;
  const double Lx = o.box_size[0], Ly = o.box_size[1], Lz = o.box_size[2];
  std::vector<double> rho(M, 0.0);
  const double ix_ = static_cast<double>(nx) / Lx, iy_ = static_cast<double>(ny) / Ly,
               iz_ = static_cast<double>(nz) / Lz;
  for (size_t p = 0; p < n; ++p) {
    const Stencil sx = stencil_of(wrap_position(x[p], Lx) * ix_, o.assignment_scheme);
    const Stencil sy = stencil_of(wrap_position(y[p], Ly) * iy_, o.assignment_scheme);
    const Stencil sz = stencil_of(wrap_position(z[p], Lz) * iz_, o.assignment_scheme);
    for (int a = 0; a < sx.count; ++a) {
      const size_t i = wrap_index(sx.off[a], nx);
      for (int b = 0; b < sy.count; ++b) {
        const size_t j = wrap_index(sy.off[b], ny);
        for (int c = 0; c < sz.count; ++c) {
          const size_t k = wrap_index(sz.off[c], nz);
          const double w = sx.w[a] * sy.w[b] * sz.w[c];
          rho[(i * ny + j) * nz + k] += m[p] * w;
        }
      }
    }
  }
  const double cell_volume = (Lx * Ly * Lz) / static_cast<double>(M);
  for (size_t i = 0; i < M; ++i) rho[i] /= cell_volume;
  if (density_out) *density_out = rho;
  pm_mesh_solve(rho, o, mesh, split_scale, F);
}

// Mesh part of the solve: density (row-major (ix*ny+iy)*nz+iz) -> three real
// force meshes.
void pm_mesh_solve(const std::vector<double>& rho, const chui_treepm_options& o, const int mesh[3],
                   double split_scale, PmField& F) {
  const size_t nx = mesh[0], ny = mesh[1], nz = mesh[2];
  const size_t M = nx * ny * nz;
  const double Lx = o.box_size[0], Ly = o.box_size[1], Lz = o.box_size[2];
  double sum = 0.0;
  for (size_t i = 0; i < M; ++i) sum += rho[i];
  const double mean = sum / static_cast<double>(M);
  std::vector<std::complex<double>> spec(M);
  for (size_t i = 0; i < M; ++i) spec[i] = rho[i] - mean;
  fft3d(spec, nx, ny, nz, -1);
  const int wexp = o.assignment_scheme == CHUI_ASSIGN_CIC ? 2 : 3;
  auto kvals = [&](size_t nn, double L, std::vector<double>& k, std::vector<double>& w) {
    k.resize(nn); w.assign(nn, 1.0);
    const double dk = 2.0 * kPi / L;
    for (size_t i = 0; i < nn; ++i) {
      const long mode = i <= nn / 2 ? static_cast<long>(i) : static_cast<long>(i) - static_cast<long>(nn);
      k[i] = dk * static_cast<double>(mode);
      if (o.window_deconvolution) w[i] = std::pow(sinc(0.5 * k[i] * L / static_cast<double>(nn)), wexp);
    }
  };
  std::vector<double> kx, ky, kz, wx, wy, wz;
  kvals(nx, Lx, kx, wx); kvals(ny, Ly, ky, wy); kvals(nz, Lz, kz, wz);
  const double pref = -4.0 * kPi * o.gravitational_constant;
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j)
      for (size_t k = 0; k < nz; ++k) {
        const size_t idx = (i * ny + j) * nz + k;
        const double k2 = kx[i] * kx[i] + ky[j] * ky[j] + kz[k] * kz[k];
        if (k2 == 0.0) { spec[idx] = 0.0; continue; }
        double corr = 1.0;
        if (o.window_deconvolution) {
          const double tw = wx[i] * wy[j] * wz[k];
          corr = 1.0 / std::max(tw * tw, 1.0e-12);
        }
        const double filt = split_scale > 0.0 ? std::exp(-(std::sqrt(k2) * split_scale) * (std::sqrt(k2) * split_scale)) : 1.0;
        spec[idx] *= pref * corr * filt / k2;
      }
  F.nx = nx; F.ny = ny; F.nz = nz;
  std::vector<double>* outs[3] = {&F.fx, &F.fy, &F.fz};
  for (int axis = 0; axis < 3; ++axis) {
    std::vector<std::complex<double>> g(M);
    for (size_t i = 0; i < nx; ++i)
      for (size_t j = 0; j < ny; ++j)
        for (size_t k = 0; k < nz; ++k) {
          const size_t idx = (i * ny + j) * nz + k;
          double kk;
          if (axis == 0) kk = (nx % 2 == 0 && i == nx / 2) ? 0.0 : kx[i];
          else if (axis == 1) kk = (ny % 2 == 0 && j == ny / 2) ? 0.0 : ky[j];
          else kk = (nz % 2 == 0 && k == nz / 2) ? 0.0 : kz[k];
          g[idx] = std::complex<double>(0.0, -kk) * spec[idx];
        }
    fft3d(g, nx, ny, nz, +1);
    outs[axis]->resize(M);
    for (size_t i = 0; i < M; ++i) (*outs[axis])[i] = g[i].real() / static_cast<double>(M);
  }
}

void pm_interpolate(const PmField& F, const double* x, const double* y, const double* z, size_t n,
                    const chui_treepm_options& o, double* ax, double* ay, double* az) {
  const double Lx = o.box_size[0], Ly = o.box_size[1], Lz = o.box_size[2];
  const double ix_ = static_cast<double>(F.nx) / Lx, iy_ = static_cast<double>(F.ny) / Ly,
               iz_ = static_cast<double>(F.nz) / Lz;
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < n; ++p) {
    const Stencil sx = stencil_of(wrap_position(x[p], Lx) * ix_, o.assignment_scheme);
    const Stencil sy = stencil_of(wrap_position(y[p], Ly) * iy_, o.assignment_scheme);
    const Stencil sz = stencil_of(wrap_position(z[p], Lz) * iz_, o.assignment_scheme);
    double gx = 0, gy = 0, gz = 0;
    for (int a = 0; a < sx.count; ++a) {
      const size_t i = wrap_index(sx.off[a], F.nx);
      for (int b = 0; b < sy.count; ++b) {
        const size_t j = wrap_index(sy.off[b], F.ny);
        for (int c = 0; c < sz.count; ++c) {
          const size_t k = wrap_index(sz.off[c], F.nz);
          const double w = sx.w[a] * sy.w[b] * sz.w[c];
          const size_t idx = (i * F.ny + j) * F.nz + k;
          gx += w * F.fx[idx]; gy += w * F.fy[idx]; gz += w * F.fz[idx];
        }
      }
    }
    ax[p] = gx; ay[p] = gy; az[p] = gz;
  }
}

double mesh_spacing(const chui_treepm_options& o, const int mesh[3]) {
  const double hx = o.box_size[0] / mesh[0], hy = o.box_size[1] / mesh[1], hz = o.box_size[2] / mesh[2];
  return std::cbrt(hx * hy * hz);
}

// ---------------------------------------------------------------------------
// Full solve: tree_pm_coupling.cpp:2282-3519 for one rank, all rows active,
// refresh on: PM accel written first, then the residual walk is added.
struct SolveOut { WalkCounters counters; Tree tree; double anchor[3]; };

void short_range(const double* x, const double* y, const double* z, const double* m, size_t n,
                 const chui_treepm_options& o, const int mesh[3], double* ax, double* ay, double* az,
                 SolveOut& out) {
  const double L[3] = {o.box_size[0], o.box_size[1], o.box_size[2]};
  const double* src[3] = {x, y, z};
  std::vector<double> u[3];
  for (int a = 0; a < 3; ++a) {
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = 0; i < n; ++i) { lo = std::min(lo, src[a][i]); hi = std::max(hi, src[a][i]); }
    u[a].resize(n);
    out.anchor[a] = unwrap_axis(src[a], n, L[a], u[a].data(), lo >= 0.0 && hi < L[a]);
  }
  build_tree(u[0].data(), u[1].data(), u[2].data(), m, n, static_cast<size_t>(o.max_leaf_size),
             o.multipole_order == 2, o.softening_epsilon, out.tree);
  const Tree& t = out.tree;
  std::vector<double> sx(n), sy(n), sz(n), sm(n);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t r = t.order[i];
    sx[i] = x[r]; sy[i] = y[r]; sz[i] = z[r]; sm[i] = m[r];
  }
  const double h = mesh_spacing(o, mesh);
  WalkParams P{};
  for (int a = 0; a < 3; ++a) P.L[a] = L[a];
  P.G = o.gravitational_constant;
  P.a = o.asmth_cells * h;
  P.rcut = o.rcut_cells * h;
  P.rcut2 = P.rcut * P.rcut;
  P.theta = o.opening_theta;
  P.eps = o.softening_epsilon;
  P.criterion = o.opening_criterion;
  P.quad = o.multipole_order == 2;
  std::vector<WalkCounters> per_thread(1);
#pragma omp parallel
  {
#pragma omp single
    {
#ifdef _OPENMP
      per_thread.resize(omp_get_num_threads());
#endif
    }
    std::vector<uint32_t> stack;
    stack.reserve(1 + 7 * 65);
#ifdef _OPENMP
    WalkCounters& C = per_thread[omp_get_thread_num()];
#else
    WalkCounters& C = per_thread[0];
#endif
#pragma omp for schedule(dynamic, 64)
    for (size_t i = 0; i < n; ++i) {
      const auto a = walk_target(t, sx.data(), sy.data(), sz.data(), sm.data(), x[i], y[i], z[i],
                                 static_cast<uint32_t>(i), P, stack, C);
      ax[i] = a[0]; ay[i] = a[1]; az[i] = a[2];
    }
  }
  for (const auto& c : per_thread) {
    out.counters.visited += c.visited; out.counters.pruned += c.pruned; out.counters.pairs += c.pairs;
    out.counters.multipoles += c.multipoles; out.counters.leaves += c.leaves;
    out.counters.skipped_pairs += c.skipped_pairs;
  }
}

// ---------------------------------------------------------------------------
// Ewald: tests/support/periodic_ewald_reference.cpp:202-400 (real + reciprocal
// sums, optional Plummer lattice correction), compensated accumulation.
struct Kahan3 {
  double s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
  void add(double a, double b, double d) {
    const double v[3] = {a, b, d};
    for (int i = 0; i < 3; ++i) {
      const double t = s[i] + v[i];
      if (std::abs(s[i]) >= std::abs(v[i])) c[i] += (s[i] - t) + v[i];
      else c[i] += (v[i] - t) + s[i];
      s[i] = t;
    }
  }
  double get(int i) const { return s[i] + c[i]; }
};

void ewald(const double* x, const double* y, const double* z, const double* m, size_t n,
           const uint32_t* targets, size_t nt, const double L[3], double G, double alpha, int real_images,
           int modes, double eps, double* ax, double* ay, double* az) {
  const double V = L[0] * L[1] * L[2];
  const double rp = 4.0 * kPi * G / V;
  const double k0[3] = {2.0 * kPi / L[0], 2.0 * kPi / L[1], 2.0 * kPi / L[2]};
  const double two_over_sqrt_pi = 2.0 * kInvSqrtPi;
  auto canon = [&](double v, double l) { double w = std::fmod(v, l); if (w < 0) w += l; if (w >= l) w -= l; return w; };
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t ti = 0; ti < nt; ++ti) {
    const uint32_t self = targets[ti];
    const double tx = canon(x[self], L[0]), ty = canon(y[self], L[1]), tz = canon(z[self], L[2]);
    Kahan3 acc;
    for (size_t j = 0; j < n; ++j) {
      double d[3] = {canon(x[j], L[0]) - tx, canon(y[j], L[1]) - ty, canon(z[j], L[2]) - tz};
      for (int a = 0; a < 3; ++a) d[a] -= L[a] * std::nearbyint(d[a] / L[a]);
      const bool is_self = j == self;
      if (is_self) d[0] = d[1] = d[2] = 0.0;
      if (m[j] == 0.0) continue;
      for (int ix = -real_images; ix <= real_images; ++ix)
        for (int iy = -real_images; iy <= real_images; ++iy)
          for (int iz = -real_images; iz <= real_images; ++iz) {
            if (is_self && ix == 0 && iy == 0 && iz == 0) continue;
            const double sx = d[0] + ix * L[0], sy = d[1] + iy * L[1], sz = d[2] + iz * L[2];
            const double r2 = sx * sx + sy * sy + sz * sz;
            if (r2 == 0.0) throw std::invalid_argument("Ewald coincident distinct pair");
            const double r = std::sqrt(r2);
            const double ar = alpha * r;
            const double scr = std::erfc(ar) + two_over_sqrt_pi * ar * std::exp(-ar * ar);
            const double f = G * m[j] * scr / (r2 * r);
            acc.add(f * sx, f * sy, f * sz);
          }
      if (eps > 0.0) {
        const int lim = 5;
        const double e2 = eps * eps;
        for (int ix = -lim; ix <= lim; ++ix)
          for (int iy = -lim; iy <= lim; ++iy)
            for (int iz = -lim; iz <= lim; ++iz) {
              if (is_self && ix == 0 && iy == 0 && iz == 0) continue;
              const double sx = d[0] + ix * L[0], sy = d[1] + iy * L[1], sz = d[2] + iz * L[2];
              const double r2 = sx * sx + sy * sy + sz * sz;
              const double r = std::sqrt(r2);
              const double nw = 1.0 / (r2 * r);
              const double s2 = r2 + e2;
              const double pl = 1.0 / (s2 * std::sqrt(s2));
              const double f = G * m[j] * (pl - nw);
              acc.add(f * sx, f * sy, f * sz);
            }
      }
      if (is_self) continue;
      for (int hx = -modes; hx <= modes; ++hx)
        for (int hy = -modes; hy <= modes; ++hy)
          for (int hz = -modes; hz <= modes; ++hz) {
            if (hx == 0 && hy == 0 && hz == 0) continue;
            const double wx = hx * k0[0], wy = hy * k0[1], wz = hz * k0[2];
            const double k2 = wx * wx + wy * wy + wz * wz;
            const double damp = std::exp(-k2 / (4.0 * alpha * alpha));
            const double ph = wx * d[0] + wy * d[1] + wz * d[2];
            const double f = rp * m[j] * damp * std::sin(ph) / k2;
            acc.add(f * wx, f * wy, f * wz);
          }
    }
    ax[ti] = acc.get(0); ay[ti] = acc.get(1); az[ti] = acc.get(2);
  }
}

}  // namespace oracle

// ===========================================================================
// C API for ctypes (tests, smoke check, CPU baseline).
using namespace oracle;

#define ORACLE_API extern "C" __attribute__((visibility("default")))

#define ORACLE_TRY(...)                                   \
  try { __VA_ARGS__; return 0; }                                 \
  catch (const std::invalid_argument& e) { g_error = e.what(); return 1; } \
  catch (const std::overflow_error& e) { g_error = e.what(); return 3; }   \
  catch (const std::exception& e) { g_error = e.what(); return 2; }

ORACLE_API const char* oracle_last_error() { return g_error.c_str(); }

ORACLE_API int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

ORACLE_API int oracle_unwrap_axis(const double* in, uint64_t n, double L, int canonical, double* out, double* anchor) {
  ORACLE_TRY(*anchor = unwrap_axis(in, n, L, out, canonical != 0))
}

ORACLE_API int oracle_morton_keys(const double* x, const double* y, const double* z, uint64_t n, uint64_t* keys) {
  ORACLE_TRY(morton_keys(x, y, z, n, keys))
}

ORACLE_API int oracle_morton_order(const double* x, const double* y, const double* z, uint64_t n,
                                   uint64_t* keys_sorted, uint32_t* order) {
  ORACLE_TRY(morton_order(x, y, z, n, keys_sorted, order))
}

static Tree g_tree;

// Builds a tree from (already unwrapped) coordinates; node count returned.
ORACLE_API int oracle_build_tree(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                                 int leaf, int quad, double eps, uint64_t* node_count) {
  ORACLE_TRY(build_tree(x, y, z, m, n, static_cast<size_t>(leaf), quad != 0, eps, g_tree);
             *node_count = g_tree.size())
}

// Copies the last built tree (layout identical to chui_treepm_tree_copy).
ORACLE_API int oracle_tree_copy(double* geometry, double* moments, uint32_t* begin, uint32_t* count,
                                uint32_t* child_index, uint32_t* order, uint32_t* subtree) {
  const Tree& t = g_tree;
  const size_t nn = t.size();
  for (size_t i = 0; i < nn; ++i) {
    if (geometry) { geometry[4 * i] = t.cx[i]; geometry[4 * i + 1] = t.cy[i]; geometry[4 * i + 2] = t.cz[i]; geometry[4 * i + 3] = t.half[i]; }
    if (moments) {
      double* q = moments + 11 * i;
      q[0] = t.mass[i]; q[1] = t.comx[i]; q[2] = t.comy[i]; q[3] = t.comz[i];
      q[4] = t.qxx[i]; q[5] = t.qxy[i]; q[6] = t.qxz[i]; q[7] = t.qyy[i]; q[8] = t.qyz[i]; q[9] = t.qzz[i]; q[10] = t.trace[i];
    }
    if (begin) begin[i] = t.begin[i];
    if (count) count[i] = t.count[i];
    if (subtree) subtree[i] = t.subtree[i];
    if (child_index) for (int o = 0; o < 8; ++o) child_index[8 * i + o] = t.child_index[8 * i + o];
  }
  if (order) std::copy(t.order.begin(), t.order.end(), order);
  return 0;
}

ORACLE_API int oracle_split_factors(const double* r, uint64_t n, double a, double* S, double* Lr,
                                    double* residual, double eps) {
  ORACLE_TRY(for (size_t i = 0; i < n; ++i) {
    S[i] = short_range_factor(r[i], a);
    Lr[i] = long_range_factor(r[i], a);
    residual[i] = residual_inv_r3(r[i] * r[i], eps, a);
  })
}

// PM-only long-range accelerations at every particle (density optional, M doubles).
ORACLE_API int oracle_pm_accel(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                               const chui_treepm_options* o, const int32_t mesh[3], double* ax, double* ay,
                               double* az, double* density_out) {
  ORACLE_TRY({
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    const double a = o->asmth_cells * mesh_spacing(*o, msh);
    PmField F;
    std::vector<double> rho;
    pm_solve(x, y, z, m, n, *o, msh, a, F, density_out ? &rho : nullptr);
    if (density_out) std::copy(rho.begin(), rho.end(), density_out);
    pm_interpolate(F, x, y, z, n, *o, ax, ay, az);
  })
}

// Mesh-level long-range solve of a given density (M doubles, row-major) with
// split scale a (0 = unfiltered Newtonian PM); writes three M-double meshes.
ORACLE_API int oracle_pm_mesh(const double* density, const chui_treepm_options* o, const int32_t mesh[3],
                              double split_scale, double* fx, double* fy, double* fz) {
  ORACLE_TRY({
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    const size_t M = static_cast<size_t>(msh[0]) * msh[1] * msh[2];
    std::vector<double> rho(density, density + M);
    PmField F;
    pm_mesh_solve(rho, *o, msh, split_scale, F);
    std::copy(F.fx.begin(), F.fx.end(), fx);
    std::copy(F.fy.begin(), F.fy.end(), fy);
    std::copy(F.fz.begin(), F.fz.end(), fz);
  })
}

// Interpolation of three given meshes at particle positions (assignment from o).
ORACLE_API int oracle_pm_interpolate(const double* fx, const double* fy, const double* fz, const int32_t mesh[3],
                                     const double* x, const double* y, const double* z, uint64_t n,
                                     const chui_treepm_options* o, double* ax, double* ay, double* az) {
  ORACLE_TRY({
    const size_t M = static_cast<size_t>(mesh[0]) * mesh[1] * mesh[2];
    PmField F;
    F.nx = mesh[0]; F.ny = mesh[1]; F.nz = mesh[2];
    F.fx.assign(fx, fx + M); F.fy.assign(fy, fy + M); F.fz.assign(fz, fz + M);
    pm_interpolate(F, x, y, z, n, *o, ax, ay, az);
  })
}

// Split-kernel coefficients: L(r)/r^3 and the softened residual at r^2.
ORACLE_API int oracle_long_range_inv_r3(const double* r, uint64_t n, double a, double* out) {
  ORACLE_TRY(for (size_t i = 0; i < n; ++i) out[i] = long_range_inv_r3(r[i], a))
}
ORACLE_API int oracle_residual_inv_r3(const double* r2, uint64_t n, double eps, double a, double* out) {
  ORACLE_TRY(for (size_t i = 0; i < n; ++i) out[i] = residual_inv_r3(r2[i], eps, a))
}

// Residual acceleration at target t of n sources taken as ONE accepted node
// (monopole, plus quadrupole when quad != 0), non-periodic: the accepted-node
// kernel on its own.
ORACLE_API int oracle_cluster_node_accel(const double* x, const double* y, const double* z, const double* m,
                                         uint64_t n, const double t[3], double G, double eps, double a, int quad,
                                         double out[3]) {
  ORACLE_TRY({
    if (n == 0) throw std::invalid_argument("empty cluster");
    Tree tr;
    build_tree(x, y, z, m, n, n, quad != 0, eps, tr);
    WalkParams P{};
    P.G = G; P.a = a; P.eps = eps; P.quad = quad != 0;
    const auto c = node_accel(tr, 0, tr.comx[0] - t[0], tr.comy[0] - t[1], tr.comz[0] - t[2], P);
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
  })
}

// Short-range residual only; counters = {visited, pruned, pairs, multipoles, leaves, skipped}.
ORACLE_API int oracle_short_range(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                                  const chui_treepm_options* o, const int32_t mesh[3], double* ax, double* ay,
                                  double* az, uint64_t* counters, double* anchors) {
  ORACLE_TRY({
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    SolveOut out;
    short_range(x, y, z, m, n, *o, msh, ax, ay, az, out);
    g_tree = std::move(out.tree);
    if (counters) {
      counters[0] = out.counters.visited; counters[1] = out.counters.pruned; counters[2] = out.counters.pairs;
      counters[3] = out.counters.multipoles; counters[4] = out.counters.leaves; counters[5] = out.counters.skipped_pairs;
    }
    if (anchors) for (int a = 0; a < 3; ++a) anchors[a] = out.anchor[a];
  })
}

// Full TreePM force evaluation (PM + residual), all rows active. timings_ms = {pm, tree+walk}.
ORACLE_API int oracle_solve(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                            const chui_treepm_options* o, const int32_t mesh[3], double* ax, double* ay, double* az,
                            uint64_t* counters) {
  ORACLE_TRY({
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    const double a = o->asmth_cells * mesh_spacing(*o, msh);
    PmField F;
    pm_solve(x, y, z, m, n, *o, msh, a, F);
    pm_interpolate(F, x, y, z, n, *o, ax, ay, az);
    std::vector<double> sx(n), sy(n), sz(n);
    SolveOut out;
    short_range(x, y, z, m, n, *o, msh, sx.data(), sy.data(), sz.data(), out);
    for (size_t i = 0; i < n; ++i) { ax[i] += sx[i]; ay[i] += sy[i]; az[i] += sz[i]; }
    if (counters) {
      counters[0] = out.counters.visited; counters[1] = out.counters.pruned; counters[2] = out.counters.pairs;
      counters[3] = out.counters.multipoles; counters[4] = out.counters.leaves; counters[5] = out.counters.skipped_pairs;
    }
  })
}

ORACLE_API int oracle_ewald(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                            const uint32_t* targets, uint64_t nt, const double L[3], double G, double alpha,
                            int real_images, int modes, double eps, double* ax, double* ay, double* az) {
  ORACLE_TRY(ewald(x, y, z, m, n, targets, nt, L, G, alpha, real_images, modes, eps, ax, ay, az))
}

// KDK pieces: gravity_runtime.cpp:3045-3058 (kick), reference_runtime_composition.cpp:59-108 (drift).
ORACLE_API int oracle_kick(double* vx, double* vy, double* vz, const double* ax, const double* ay,
                           const double* az, uint64_t n, double kick, double drag) {
  for (size_t i = 0; i < n; ++i) {
    vx[i] = drag * vx[i] + kick * ax[i];
    vy[i] = drag * vy[i] + kick * ay[i];
    vz[i] = drag * vz[i] + kick * az[i];
  }
  return 0;
}

// Canonical periodic wrap: include/cosmosim/core/periodic_domain.hpp:8-16.
static inline double canonical_position(double x, double L) {
  double w = std::fmod(x, L);
  if (w < 0.0) w += L;
  if (!(w < L)) w = 0.0;
  return w;
}

ORACLE_API int oracle_drift(double* x, double* y, double* z, const double* vx, const double* vy, const double* vz,
                            uint64_t n, double drift, const double L[3]) {
  for (size_t i = 0; i < n; ++i) {
    x[i] += vx[i] * drift; y[i] += vy[i] * drift; z[i] += vz[i] * drift;
    x[i] = canonical_position(x[i], L[0]); y[i] = canonical_position(y[i], L[1]); z[i] = canonical_position(z[i], L[2]);
  }
  return 0;
}

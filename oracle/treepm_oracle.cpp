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

// Optional record of one target's walk, in visit order: internal nodes it
// accepted as multipoles, internal nodes it opened, leaves it summed and nodes
// the cutoff pruned (test infrastructure: the node-by-node validator).
struct WalkTrace {
  std::vector<uint32_t> accepted, opened, leaves, pruned;
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
                                  const WalkParams& P, std::vector<uint32_t>& stack, WalkCounters& C,
                                  WalkTrace* trace = nullptr) {
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
    if (!within_cutoff(ex * ex + ey * ey + ez * ez, P.rcut, P.rcut2)) {
      ++C.pruned;
      if (trace) trace->pruned.push_back(n);
      continue;
    }
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
    if (trace) (leaf ? trace->leaves : accept ? trace->accepted : trace->opened).push_back(n);
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
// Gaussian coefficient table: tree_pm_split_kernel.hpp:130-183. H(t) =
// 8 a^3 L(r)/r^3 at t = r^2/(4a^2) as a cubic Hermite interpolant on 4096 equal
// cells of [0, 64]; knot values from the closed form (series at t = 0), knot
// slopes dH/dt = (2 e^{-t}/sqrt(pi) - 1.5 H)/t pre-multiplied by the cell width.
struct GaussTable {
  static constexpr size_t kCells = 4096;
  static constexpr double kStep = 64.0 / 4096.0;
  std::vector<double> value, slope;
  GaussTable() : value(kCells + 1), slope(kCells + 1) {
    for (size_t i = 0; i <= kCells; ++i) {
      const double t = static_cast<double>(i) * kStep;
      const double q = std::sqrt(t);
      value[i] = i == 0 ? series_h(0.0) : long_range_factor(q, 0.5) / (t * q);
      const double dh = i == 0 ? -4.0 * kInvSqrtPi / 5.0 : (2.0 * kInvSqrtPi * std::exp(-t) - 1.5 * value[i]) / t;
      slope[i] = kStep * dh;
    }
  }
  double hermite(double t) const {
    const double cell = t / kStep;
    const size_t i = static_cast<size_t>(cell);
    const double u = cell - static_cast<double>(i), u2 = u * u, u3 = u2 * u;
    return (2 * u3 - 3 * u2 + 1) * value[i] + (u3 - 2 * u2 + u) * slope[i] + (-2 * u3 + 3 * u2) * value[i + 1] +
           (u3 - u2) * slope[i + 1];
  }
  // dimensionlessCoefficient: series below t = 1/64, analytic from t = 64
  double coefficient(double t) const {
    if (t < 0.015625) return series_h(t);
    if (!(t < 64.0)) return long_range_inv_r3(std::sqrt(t), 0.5);
    return hermite(t);
  }
  // longRangeInvR3
  double long_inv_r3(double r2, double a) const {
    const double t = 0.25 * (r2 / a) / a;
    if (!(t >= 0.0) || t >= 64.0) return long_range_inv_r3(std::sqrt(r2), a);
    return ((0.125 * coefficient(t) / a) / a) / a;
  }
};

const GaussTable& gauss_table() {
  static const GaussTable table;
  return table;
}

// treePmResidualRadialCoefficients (tree_pm_split_kernel.hpp:245-282): for the
// residual f(x) = (x + eps^2)^{-3/2} - H(t)/(8a^3), x = r^2, t = x/(4a^2):
// value f, f'(r)/r = 2 df/dx and (f'' - f'/r)/r^2 = 4 d^2f/dx^2.
struct Radial { double value, first_over_r, contracted_second; };
Radial residual_radial(double x, double eps, double a) {
  const double d = x + eps * eps;
  const double soft = 1.0 / (d * std::sqrt(d));
  const double t = (0.25 * x / a) / a;
  constexpr double c = 2.25675833419102514779231780624309034;
  double h, dh, ddh;
  if (t < 0.015625) {
    static const double coef[8] = {1.0 / 3.0, -1.0 / 5.0, 1.0 / 14.0, -1.0 / 54.0,
                                   1.0 / 264.0, -1.0 / 1560.0, 1.0 / 10800.0, -1.0 / 85680.0};
    h = coef[7]; dh = 0.0; ddh = 0.0;
    for (int i = 7; i > 0; --i) {
      ddh = ddh * t + 2.0 * dh;
      dh = dh * t + h;
      h = h * t + coef[i - 1];
    }
    h *= c; dh *= c; ddh *= c;
  } else {
    const double q = std::sqrt(t);
    const double e = std::exp(-t);
    h = (std::erf(q) - 0.5 * c * q * e) / (t * q);
    dh = (0.5 * c * e - 1.5 * h) / t;
    ddh = (-0.5 * c * e - 2.5 * dh) / t;
  }
  const double scale = ((0.125 / a) / a) / a;
  const double dt_dx = (0.25 / a) / a;
  return {soft - scale * h, -3.0 * soft / d - 2.0 * scale * dt_dx * dh,
          15.0 * soft / (d * d) - 4.0 * scale * dt_dx * dt_dx * ddh};
}

// monopolePlusQuadrupoleAccelPeriodic (tree_pm_coupling.cpp:829-894): the same
// expansion with the shared radial coefficients; non-finite coefficients use
// node_accel.
std::array<double, 3> node_accel_radial(const Tree& t, uint32_t n, double dx, double dy, double dz,
                                        const WalkParams& P) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double safe_r2 = std::max(r2, 1.0e-30);
  const double pe = std::max(t.smax[n], P.eps);
  const Radial R = residual_radial(safe_r2, pe, P.a);
  if (!std::isfinite(R.value) || !std::isfinite(R.first_over_r) || !std::isfinite(R.contracted_second))
    return node_accel(t, n, dx, dy, dz, P);
  const double pref = P.G;
  double ax = pref * t.mass[n] * R.value * dx;
  double ay = pref * t.mass[n] * R.value * dy;
  double az = pref * t.mass[n] * R.value * dz;
  if (!P.quad) return {ax, ay, az};
  const double tr = t.trace[n];
  const double mxx = (t.qxx[n] + tr) / 3.0, mxy = t.qxy[n] / 3.0, mxz = t.qxz[n] / 3.0;
  const double myy = (t.qyy[n] + tr) / 3.0, myz = t.qyz[n] / 3.0, mzz = (t.qzz[n] + tr) / 3.0;
  const double mdx = mxx * dx + mxy * dy + mxz * dz;
  const double mdy = mxy * dx + myy * dy + myz * dz;
  const double mdz = mxz * dx + myz * dy + mzz * dz;
  const double dmd = dx * mdx + dy * mdy + dz * mdz;
  const double fo = R.first_over_r, cr = R.contracted_second;
  ax += pref * (mdx * fo + 0.5 * tr * dx * fo + 0.5 * dx * dmd * cr);
  ay += pref * (mdy * fo + 0.5 * tr * dy * fo + 0.5 * dy * dmd * cr);
  az += pref * (mdz * fo + 0.5 * tr * dz * fo + 0.5 * dz * dmd * cr);
  return {ax, ay, az};
}

// ---------------------------------------------------------------------------
// Adaptive acceptance (TreePmAcceptancePolicy::kAdaptiveRelative).
struct AdaptiveParams {
  WalkParams P;
  double amax;       // adaptive_maximum_opening_angle
  double alpha;      // relative_force_tolerance
  double floor;      // relative_force_acceleration_floor_code
  int estimator;     // chui_adaptive_estimator
  bool com_cut;      // cutoff_node_policy == kComDistance
  bool lookup;       // gaussian_pair_lookup_enabled
  bool radial;       // radial_coefficients_enabled
  double inv2a2, inv8a3;
};

// residualAccuracyProxy (tree_pm_acceptance.hpp:14-51).
double residual_proxy(double h, double offset, double M, double r2, double eps, double a, double G) {
  if (!(r2 > 0.0) || !std::isfinite(r2)) return INFINITY;
  const double r = std::sqrt(r2);
  const double d = r2 + eps * eps;
  const double inv_d5 = 1.0 / (d * d * std::sqrt(d));
  const double inv_r3 = 1.0 / (r2 * r);
  const double inv_r4 = 1.0 / (r2 * r2);
  const double lf = long_range_factor(r, a);
  const double q = r / (2.0 * a);
  const double e = std::exp(-q * q);
  const double a3 = a * a * a;
  const double s1 = -0.5 * kInvSqrtPi * r2 * e / a3;
  const double s2 = kInvSqrtPi * e * (-r / a3 + 0.25 * r * r2 / (a3 * a * a));
  const double first = -3.0 * r * inv_d5 + s1 * inv_r3 - 3.0 * (-lf) * inv_r4;
  const double second = -3.0 * inv_d5 + 15.0 * r2 * inv_d5 / d + s2 * inv_r3 - 6.0 * s1 * inv_r4 +
                         12.0 * (-lf) * inv_r4 / r;
  const double width = 2.0 * h;
  const double rho = std::sqrt(3.0) * h + offset;
  const double proxy = G * M * std::max(width * width * inv_r4,
                                        0.5 * rho * rho * (3.0 * std::abs(first) + std::abs(r * second - first)));
  return std::isfinite(proxy) && proxy >= 0.0 ? proxy : INFINITY;
}

// acceptTreePmNode (tree_pm_acceptance.hpp:54-139) under the adaptive policy for
// a node that is neither around the target nor outside the cutoff container
// (those are opened before the MAC, tree_pm_coupling.cpp:4094-4104).
static inline bool accept_adaptive(const Tree& t, uint32_t n, double r2, bool history, double prev,
                                   const AdaptiveParams& A) {
  const double h = t.half[n];
  const double r = std::sqrt(std::max(r2, 1.0e-30));
  bool soft_ok = true;
  if (!(t.smax[n] - t.smin[n] <= 1.0e-12)) soft_ok = r > (2.0 * h + 2.0 * std::max(t.smax[n], A.P.eps));
  const bool env_ok = (2.0 * h / r) < A.amax;
  const double ox = t.cx[n] - t.comx[n], oy = t.cy[n] - t.comy[n], oz = t.cz[n] - t.comz[n];
  const double offset = std::sqrt(ox * ox + oy * oy + oz * oz);
  bool mac;
  if (history && std::isfinite(prev) && prev > A.floor) {
    const double allowance = A.alpha * prev;
    const double width = 2.0 * h;
    const double inv_r4 = 1.0 / (r2 * r2);
    const double lower = A.P.G * t.mass[n] * (width * width * inv_r4);
    if (std::isfinite(allowance) && std::isfinite(lower) && lower > allowance) {
      mac = false;
    } else if (A.estimator == CHUI_ESTIMATOR_RELATIVE_MONOPOLE) {
      mac = std::isfinite(allowance) && std::isfinite(lower);
    } else {
      const double est = residual_proxy(h, offset, t.mass[n], r2, std::max(t.smax[n], A.P.eps), A.P.a, A.P.G);
      mac = std::isfinite(est) && std::isfinite(allowance) && est <= allowance;
    }
  } else {
    mac = ((2.0 * h + offset) / std::sqrt(r2 + 1.0e-30)) < A.P.theta;
  }
  return mac && env_ok && soft_ok && std::isfinite(r2);
}

// Leaf-pair residual coefficient of the per-target walk
// (tree_pm_coupling.cpp:4216-4229): Gaussian table when enabled, else analytic.
static inline double pair_residual(double r2, const AdaptiveParams& A) {
  if (A.lookup && r2 > 0.0) {
    const double t = r2 * A.inv2a2;
    const GaussTable& T = gauss_table();
    const double lr = t < 64.0 ? T.coefficient(t) * A.inv8a3 : T.long_inv_r3(r2, A.P.a);
    return softened_inv_r3(r2, A.P.eps) - lr;
  }
  return residual_inv_r3(r2, A.P.eps, A.P.a);
}

// Per-target walk under the adaptive policy (tree_pm_coupling.cpp:4026-4272).
std::array<double, 3> walk_target_adaptive(const Tree& t, const double* sx, const double* sy, const double* sz,
                                           const double* sm, double px, double py, double pz, uint32_t self,
                                           bool history, double prev, const AdaptiveParams& A,
                                           std::vector<uint32_t>& stack, WalkCounters& C) {
  const WalkParams& P = A.P;
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
    if (t.child_count[n] == 0) {
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
        const double f = pair_residual(r2, A) * P.G;
        ax += f * sm[i] * qx; ay += f * sm[i] * qy; az += f * sm[i] * qz;
      }
      continue;
    }
    const bool inside = ax_ <= h && ay_ <= h && az_ <= h;
    const double mx = ax_ + h, my = ay_ + h, mz = az_ + h;
    const bool contained = A.com_cut || within_cutoff(mx * mx + my * my + mz * mz, P.rcut, P.rcut2);
    bool accept = false;
    double dx = 0, dy = 0, dz = 0, r2 = 0;
    if (!inside && contained) {
      dx = min_image(t.comx[n] - px, P.L[0]);
      dy = min_image(t.comy[n] - py, P.L[1]);
      dz = min_image(t.comz[n] - pz, P.L[2]);
      r2 = dx * dx + dy * dy + dz * dz;
      accept = accept_adaptive(t, n, r2, history, prev, A);
    }
    if (accept && A.com_cut && r2 > P.rcut2) { ++C.pruned; continue; }
    if (accept) {
      ++C.multipoles;
      const auto c = A.radial ? node_accel_radial(t, n, dx, dy, dz, P) : node_accel(t, n, dx, dy, dz, P);
      ax += c[0]; ay += c[1]; az += c[2];
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
// Group walk: tree_pm_group_walk.hpp:78-341 and its driver
// tree_pm_coupling.cpp:4318-4450.
struct GroupList {
  std::vector<uint32_t> nodes, begin, end;
  std::vector<double> ox, oy, oz, half;
};

struct GroupCounters {
  uint64_t groups = 0, grouped = 0, size_max = 0, fallback = 0, fallback_history = 0, fallback_geometry = 0;
  uint64_t candidate_leaves = 0, culled = 0, candidate_particles = 0, entries_max = 0;
};

// buildTreePmGroupInteractionList: box-box prune, leaves listed with their cell
// geometry relative to the group centre, internal nodes accepted by the
// group-conservative relative-monopole criterion.
void group_list(const Tree& t, const double c[3], const double hg[3], double amin, double tolg, double amax2,
                bool contain, const WalkParams& P, std::vector<uint32_t>& stack, GroupList& out, WalkCounters& C) {
  out = GroupList{};
  if (t.size() == 0) return;
  stack.clear();
  stack.push_back(0);
  while (!stack.empty()) {
    const uint32_t n = stack.back();
    stack.pop_back();
    ++C.visited;
    const double h = t.half[n];
    const double dc[3] = {min_image(t.cx[n] - c[0], P.L[0]), min_image(t.cy[n] - c[1], P.L[1]),
                          min_image(t.cz[n] - c[2], P.L[2])};
    double gap2 = 0.0, far2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double v = std::abs(dc[a]);
      const double gap = std::max(0.0, v - h - hg[a]);
      const double far = v + h + hg[a];
      gap2 += gap * gap;
      far2 += far * far;
    }
    if (gap2 > P.rcut2) { ++C.pruned; continue; }
    if (t.child_count[n] == 0) {
      ++C.leaves;
      out.begin.push_back(t.begin[n]);
      out.end.push_back(t.begin[n] + t.count[n]);
      out.ox.push_back(dc[0]); out.oy.push_back(dc[1]); out.oz.push_back(dc[2]);
      out.half.push_back(h);
      continue;
    }
    bool accept = gap2 > 0.0 && (!contain || far2 <= P.rcut2);
    if (accept) {
      const double dm[3] = {min_image(t.comx[n] - c[0], P.L[0]), min_image(t.comy[n] - c[1], P.L[1]),
                            min_image(t.comz[n] - c[2], P.L[2])};
      double rmin2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double gap = std::max(0.0, std::abs(dm[a]) - hg[a]);
        rmin2 += gap * gap;
      }
      const double width2 = 4.0 * h * h;
      accept = rmin2 > 0.0 && width2 < amax2 * rmin2 && t.mass[n] * width2 <= tolg * amin * rmin2 * rmin2;
    }
    if (accept) {
      ++C.multipoles;
      out.nodes.push_back(n);
    } else {
      const size_t co = static_cast<size_t>(n) * 8;
      for (int o = 8; o > 0; --o) {
        const uint32_t ch = t.child_index[co + o - 1];
        if (ch != kInvalid) stack.push_back(ch);
      }
    }
  }
}

// accumulateTreePmGroupPairsImpl: Hermite-table residual over tree-order
// sources [b, e), one conditional periodic shift per axis, the target's own
// source excluded by identity (distinct coincident sources count as pairs).
static void group_pairs(const double* sx, const double* sy, const double* sz, const double* sm,
                        const uint32_t* order, uint32_t self, double px, double py, double pz, uint32_t b,
                        uint32_t e, const AdaptiveParams& A, double& ax, double& ay, double& az, uint64_t& pairs) {
  const WalkParams& P = A.P;
  const GaussTable& T = gauss_table();
  const double tlim = std::nextafter(64.0, 0.0);
  const double eps2 = P.eps * P.eps;
  double s[3] = {0, 0, 0};
  for (uint32_t j = b; j < e; ++j) {
    double d[3] = {sx[j] - px, sy[j] - py, sz[j] - pz};
    for (int a = 0; a < 3; ++a) {
      const double hl = 0.5 * P.L[a];
      d[a] = d[a] > hl ? d[a] - P.L[a] : (d[a] < -hl ? d[a] + P.L[a] : d[a]);
    }
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const bool inside = order[j] != self && r2 <= P.rcut2;
    const double tt = std::min(r2 * A.inv2a2, tlim);
    const double lr = T.hermite(tt) * A.inv8a3;
    const double dd = r2 + eps2;
    const double soft = 1.0 / (dd * std::sqrt(dd));
    const double f = inside && r2 > 0.0 ? (soft - lr) * P.G * sm[j] : 0.0;
    for (int a = 0; a < 3; ++a) s[a] += f * d[a];
    pairs += inside ? 1 : 0;
  }
  ax += s[0]; ay += s[1]; az += s[2];
}

// accumulateTreePmGroupCulledLeaves + the accepted multipoles of the list for one
// target: leaves whose cell lies beyond r_cut (1 + 1e-9) of the target are
// skipped, contiguous survivors are summed as one range.
std::array<double, 3> group_target(const Tree& t, const GroupList& L, const double c[3], const double* sx,
                                   const double* sy, const double* sz, const double* sm, uint32_t self, double px,
                                   double py, double pz, const AdaptiveParams& A, WalkCounters& C,
                                   GroupCounters& G) {
  const WalkParams& P = A.P;
  const double to[3] = {min_image(px - c[0], P.L[0]), min_image(py - c[1], P.L[1]), min_image(pz - c[2], P.L[2])};
  const double cull2 = P.rcut2 * (1.0 + 1.0e-9);
  auto wrap = [&](double d, int a) {
    const double hl = 0.5 * P.L[a];
    return d > hl ? d - P.L[a] : (d < -hl ? d + P.L[a] : d);
  };
  double ax = 0, ay = 0, az = 0;
  uint64_t pairs = 0;
  bool open = false;
  uint32_t rb = 0, re = 0;
  for (size_t l = 0; l < L.begin.size(); ++l) {
    const double h = L.half[l];
    const double gx = std::max(0.0, std::abs(wrap(L.ox[l] - to[0], 0)) - h);
    const double gy = std::max(0.0, std::abs(wrap(L.oy[l] - to[1], 1)) - h);
    const double gz = std::max(0.0, std::abs(wrap(L.oz[l] - to[2], 2)) - h);
    if (gx * gx + gy * gy + gz * gz > cull2) { ++G.culled; continue; }
    G.candidate_particles += L.end[l] - L.begin[l];
    if (open && re == L.begin[l]) { re = L.end[l]; continue; }
    if (open) group_pairs(sx, sy, sz, sm, t.order.data(), self, px, py, pz, rb, re, A, ax, ay, az, pairs);
    rb = L.begin[l]; re = L.end[l]; open = true;
  }
  if (open) group_pairs(sx, sy, sz, sm, t.order.data(), self, px, py, pz, rb, re, A, ax, ay, az, pairs);
  G.candidate_leaves += L.begin.size();
  C.pairs += pairs;
  for (const uint32_t n : L.nodes) {
    const double dx = min_image(t.comx[n] - px, P.L[0]);
    const double dy = min_image(t.comy[n] - py, P.L[1]);
    const double dz = min_image(t.comz[n] - pz, P.L[2]);
    if (A.com_cut && dx * dx + dy * dy + dz * dz > P.rcut2) { ++C.pruned; continue; }
    const auto q = A.radial ? node_accel_radial(t, n, dx, dy, dz, P) : node_accel(t, n, dx, dy, dz, P);
    ax += q[0]; ay += q[1]; az += q[2];
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

// Adaptive-policy residual over active slots (tree_pm_coupling.cpp:3839-3869
// eligibility, :4451-4546 residual blocks and target groups, :4333-4450 group
// evaluation with per-target fallback). prev = previous total-acceleration
// magnitude per active slot (nullable). Outputs per active slot.
struct AdaptiveOut { WalkCounters counters; GroupCounters group; uint32_t status = 0; };

void adaptive_short_range(const double* x, const double* y, const double* z, const double* m, size_t n,
                          const uint32_t* active, size_t na, const double* prev, const chui_treepm_options& o,
                          const chui_treepm_adaptive_options& ad, const int mesh[3], double* ax, double* ay,
                          double* az, AdaptiveOut& out) {
  const double L[3] = {o.box_size[0], o.box_size[1], o.box_size[2]};
  const double* src[3] = {x, y, z};
  std::vector<double> u[3];
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY; hi[a] = -INFINITY;
    for (size_t i = 0; i < n; ++i) { lo[a] = std::min(lo[a], src[a][i]); hi[a] = std::max(hi[a], src[a][i]); }
    u[a].resize(n);
    unwrap_axis(src[a], n, L[a], u[a].data(), lo[a] >= 0.0 && hi[a] < L[a]);
  }
  Tree t;
  build_tree(u[0].data(), u[1].data(), u[2].data(), m, n, static_cast<size_t>(o.max_leaf_size),
             o.multipole_order == 2, o.softening_epsilon, t);
  std::vector<double> sx(n), sy(n), sz(n), sm(n);
  std::vector<uint32_t> pos_of_row(n), leaf_of_pos(n, kInvalid);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t r = t.order[i];
    sx[i] = x[r]; sy[i] = y[r]; sz[i] = z[r]; sm[i] = m[r];
    pos_of_row[r] = static_cast<uint32_t>(i);
  }
  for (size_t nd = 0; nd < t.size(); ++nd)
    if (t.child_count[nd] == 0)
      for (uint32_t p = t.begin[nd]; p < t.begin[nd] + t.count[nd]; ++p) leaf_of_pos[p] = static_cast<uint32_t>(nd);
  const double h = mesh_spacing(o, mesh);
  AdaptiveParams A{};
  for (int a = 0; a < 3; ++a) A.P.L[a] = L[a];
  A.P.G = o.gravitational_constant;
  A.P.a = o.asmth_cells * h;
  A.P.rcut = o.rcut_cells * h;
  A.P.rcut2 = A.P.rcut * A.P.rcut;
  A.P.theta = o.opening_theta;
  A.P.eps = o.softening_epsilon;
  A.P.criterion = o.opening_criterion;
  A.P.quad = o.multipole_order == 2;
  A.amax = ad.maximum_opening_angle;
  A.alpha = o.relative_force_tolerance;
  A.floor = o.relative_force_acceleration_floor;
  A.estimator = ad.error_estimator;
  A.com_cut = ad.cutoff_node_policy == CHUI_CUTOFF_COM_DISTANCE;
  A.lookup = ad.gaussian_pair_lookup != 0;
  A.radial = ad.radial_coefficients != 0;
  A.inv2a2 = (0.25 / A.P.a) / A.P.a;
  A.inv8a3 = ((0.125 / A.P.a) / A.P.a) / A.P.a;
  // group walk eligibility
  const double t_cut = (0.25 * A.P.rcut2 / A.P.a) / A.P.a;
  uint32_t status = CHUI_GROUP_WALK_ACTIVE;
  if (!ad.group_walk) status = CHUI_GROUP_WALK_NOT_REQUESTED;
  else if (t.size() == 0) status = CHUI_GROUP_WALK_NO_LOCAL_TREE;
  else if (ad.error_estimator != CHUI_ESTIMATOR_RELATIVE_MONOPOLE || !A.lookup)
    status = CHUI_GROUP_WALK_INCOMPATIBLE_POLICY;
  else if (!(std::isfinite(t_cut) && t_cut < 63.0)) status = CHUI_GROUP_WALK_CUTOFF_OUTSIDE_TABLE;
  else if (!(std::isnormal(A.inv2a2) && std::isnormal(A.inv8a3))) status = CHUI_GROUP_WALK_LOOKUP_SCALING_UNREPRESENTABLE;
  out.status = status;
  const bool grouped = status == CHUI_GROUP_WALK_ACTIVE;
  // visit order: active slots by tree position
  std::vector<uint32_t> row(na), visit(na);
  for (size_t s = 0; s < na; ++s) {
    row[s] = active ? active[s] : static_cast<uint32_t>(s);
    if (row[s] >= n) throw std::invalid_argument("TreePM active particle index out of range");
    visit[s] = static_cast<uint32_t>(s);
  }
  std::stable_sort(visit.begin(), visit.end(),
                   [&](uint32_t a, uint32_t b) { return pos_of_row[row[a]] < pos_of_row[row[b]]; });
  const size_t B = static_cast<size_t>(o.residual_block_size);
  const double tolg = o.relative_force_tolerance / o.gravitational_constant;
  const double amax2 = ad.maximum_opening_angle * ad.maximum_opening_angle;
  const double max_extent = ad.group_max_extent_cutoff_fraction * std::sqrt(A.P.rcut2);
  const size_t nblocks = (na + B - 1) / B;
  std::vector<WalkCounters> wc(nblocks);
  std::vector<GroupCounters> gc(nblocks);
#pragma omp parallel
  {
    std::vector<uint32_t> stack;
    stack.reserve(1 + 7 * 65);
    GroupList list;
#pragma omp for schedule(dynamic, 1)
    for (size_t blk = 0; blk < nblocks; ++blk) {
      WalkCounters& C = wc[blk];
      GroupCounters& G = gc[blk];
      const size_t vb = blk * B, ve = std::min(vb + B, na);
      auto per_target = [&](size_t v) {
        const uint32_t s = visit[v];
        const bool hist = prev != nullptr && std::isfinite(prev[s]);
        const auto a = walk_target_adaptive(t, sx.data(), sy.data(), sz.data(), sm.data(), x[row[s]], y[row[s]],
                                            z[row[s]], row[s], hist, hist ? prev[s] : 0.0, A, stack, C);
        ax[s] = a[0]; ay[s] = a[1]; az[s] = a[2];
      };
      if (!grouped) {
        for (size_t v = vb; v < ve; ++v) per_target(v);
        continue;
      }
      size_t v = vb;
      while (v < ve) {
        const uint32_t leaf = leaf_of_pos[pos_of_row[row[visit[v]]]];
        size_t ge = v + 1;
        if (ad.group_target_size == 0) {
          while (ge < ve && leaf_of_pos[pos_of_row[row[visit[ge]]]] == leaf) ++ge;
        } else {
          const uint32_t r0 = row[visit[v]];
          const double ref[3] = {x[r0], y[r0], z[r0]};
          double glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
          while (ge < ve && ge - v < ad.group_target_size) {
            const uint32_t r = row[visit[ge]];
            const double p[3] = {x[r], y[r], z[r]};
            double nlo[3], nhi[3];
            bool within = true;
            for (int a = 0; a < 3; ++a) {
              const double d = min_image(p[a] - ref[a], L[a]);
              nlo[a] = std::min(glo[a], d);
              nhi[a] = std::max(ghi[a], d);
              within = within && nhi[a] - nlo[a] <= max_extent;
            }
            if (!within) break;
            for (int a = 0; a < 3; ++a) { glo[a] = nlo[a]; ghi[a] = nhi[a]; }
            ++ge;
          }
        }
        const size_t count = ge - v;
        // evaluate_target_group: history and bounded-metric checks, group box
        int fall = 0;  // 1 history, 2 geometry
        double amin = INFINITY, ref[3] = {0, 0, 0}, glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
        for (size_t i = 0; i < count && !fall; ++i) {
          const uint32_t s = visit[v + i];
          const double pa = prev ? prev[s] : NAN;
          if (!(std::isfinite(pa) && pa > o.relative_force_acceleration_floor)) { fall = 1; break; }
          amin = std::min(amin, pa);
          const double p[3] = {x[row[s]], y[row[s]], z[row[s]]};
          for (int a = 0; a < 3; ++a) {
            if (!(lo[a] - p[a] >= -L[a] && hi[a] - p[a] <= L[a])) { fall = 2; break; }
            if (i == 0) { ref[a] = p[a]; continue; }
            const double d = min_image(p[a] - ref[a], L[a]);
            glo[a] = std::min(glo[a], d);
            ghi[a] = std::max(ghi[a], d);
          }
        }
        if (fall) {
          G.fallback += count;
          (fall == 1 ? G.fallback_history : G.fallback_geometry) += count;
          for (size_t i = v; i < ge; ++i) per_target(i);
          v = ge;
          continue;
        }
        double c[3], hg[3];
        for (int a = 0; a < 3; ++a) {
          c[a] = ref[a] + 0.5 * (glo[a] + ghi[a]);
          hg[a] = 0.5 * (ghi[a] - glo[a]);
        }
        group_list(t, c, hg, amin, tolg, amax2, !A.com_cut, A.P, stack, list, C);
        ++G.groups;
        G.grouped += count;
        G.size_max = std::max<uint64_t>(G.size_max, count);
        G.entries_max = std::max<uint64_t>(G.entries_max, list.begin.size() + list.nodes.size());
        for (size_t i = v; i < ge; ++i) {
          const uint32_t s = visit[i];
          const auto a = group_target(t, list, c, sx.data(), sy.data(), sz.data(), sm.data(), row[s], x[row[s]],
                                      y[row[s]], z[row[s]], A, C, G);
          ax[s] = a[0]; ay[s] = a[1]; az[s] = a[2];
        }
        v = ge;
      }
    }
  }
  for (size_t b = 0; b < nblocks; ++b) {
    const WalkCounters& c = wc[b];
    out.counters.visited += c.visited; out.counters.pruned += c.pruned; out.counters.pairs += c.pairs;
    out.counters.multipoles += c.multipoles; out.counters.leaves += c.leaves;
    out.counters.skipped_pairs += c.skipped_pairs;
    const GroupCounters& g = gc[b];
    out.group.groups += g.groups; out.group.grouped += g.grouped;
    out.group.size_max = std::max(out.group.size_max, g.size_max);
    out.group.fallback += g.fallback; out.group.fallback_history += g.fallback_history;
    out.group.fallback_geometry += g.fallback_geometry;
    out.group.candidate_leaves += g.candidate_leaves; out.group.culled += g.culled;
    out.group.candidate_particles += g.candidate_particles;
    out.group.entries_max = std::max(out.group.entries_max, g.entries_max);
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
ORACLE_API int oracle_last_tree_size(uint64_t* nodes) {
  ORACLE_TRY(*nodes = g_tree.size())
}

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

// Adaptive-policy short-range residual per active slot (active may be NULL for
// all rows, prev may be NULL); counters in the chui_treepm_walk_counters layout.
// One walk of the strict policy for an arbitrary target point (px, py, pz),
// not a source (no self exclusion), over the tree of the given sources:
// short_range()'s tree, source copy and parameters. Writes the residual
// acceleration and the trace lists (node numbers in preorder numbering, each
// list truncated to cap entries; sizes[] receives the full lengths in the
// order accepted, opened, leaves, pruned).
ORACLE_API int oracle_probe_walk(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                                 const chui_treepm_options* o, const int32_t mesh[3], double px, double py,
                                 double pz, double* acc3, uint32_t* accepted, uint32_t* opened, uint32_t* leaves,
                                 uint32_t* pruned, uint64_t cap, uint64_t* sizes) {
  ORACLE_TRY({
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    const double L[3] = {o->box_size[0], o->box_size[1], o->box_size[2]};
    const double* src[3] = {x, y, z};
    std::vector<double> u[3];
    for (int a = 0; a < 3; ++a) {
      double lo = INFINITY, hi = -INFINITY;
      for (size_t i = 0; i < n; ++i) { lo = std::min(lo, src[a][i]); hi = std::max(hi, src[a][i]); }
      u[a].resize(n);
      unwrap_axis(src[a], n, L[a], u[a].data(), lo >= 0.0 && hi < L[a]);
    }
    build_tree(u[0].data(), u[1].data(), u[2].data(), m, n, static_cast<size_t>(o->max_leaf_size),
               o->multipole_order == 2, o->softening_epsilon, g_tree);
    const Tree& t = g_tree;
    std::vector<double> sx(n), sy(n), sz(n), sm(n);
    for (size_t i = 0; i < n; ++i) {
      const uint32_t r = t.order[i];
      sx[i] = x[r]; sy[i] = y[r]; sz[i] = z[r]; sm[i] = m[r];
    }
    const double h = mesh_spacing(*o, msh);
    WalkParams P{};
    for (int a = 0; a < 3; ++a) P.L[a] = L[a];
    P.G = o->gravitational_constant;
    P.a = o->asmth_cells * h;
    P.rcut = o->rcut_cells * h;
    P.rcut2 = P.rcut * P.rcut;
    P.theta = o->opening_theta;
    P.eps = o->softening_epsilon;
    P.criterion = o->opening_criterion;
    P.quad = o->multipole_order == 2;
    std::vector<uint32_t> stack;
    WalkCounters C;
    WalkTrace T;
    const auto a = walk_target(t, sx.data(), sy.data(), sz.data(), sm.data(), px, py, pz, kInvalid, P, stack, C, &T);
    acc3[0] = a[0]; acc3[1] = a[1]; acc3[2] = a[2];
    const std::vector<uint32_t>* lists[4] = {&T.accepted, &T.opened, &T.leaves, &T.pruned};
    uint32_t* outs[4] = {accepted, opened, leaves, pruned};
    for (int k = 0; k < 4; ++k) {
      sizes[k] = lists[k]->size();
      if (outs[k]) std::copy_n(lists[k]->begin(), std::min<size_t>(cap, lists[k]->size()), outs[k]);
    }
  });
}

// residualAccuracyProxy of the adaptive walk (residual_proxy) on arrays of
// (h, offset, M, r^2) with one softening, split scale and G.
ORACLE_API int oracle_residual_proxy(const double* h, const double* offset, const double* mass, const double* r2,
                                     uint64_t n, double eps, double a, double G, double* out) {
  ORACLE_TRY(for (size_t i = 0; i < n; ++i) out[i] = residual_proxy(h[i], offset[i], mass[i], r2[i], eps, a, G));
}

ORACLE_API int oracle_short_range_adaptive(const double* x, const double* y, const double* z, const double* m,
                                           uint64_t n, const uint32_t* active, uint64_t n_active, const double* prev,
                                           const chui_treepm_options* o, const chui_treepm_adaptive_options* ad,
                                           const int32_t mesh[3], double* ax, double* ay, double* az,
                                           chui_treepm_walk_counters* counters) {
  ORACLE_TRY({
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    AdaptiveOut out;
    adaptive_short_range(x, y, z, m, n, active, active ? n_active : n, prev, *o, *ad, msh, ax, ay, az, out);
    if (counters) {
      chui_treepm_walk_counters c{};
      c.struct_size = sizeof(c);
      c.group_walk_status = out.status;
      c.visited_nodes = out.counters.visited;
      c.cutoff_pruned_nodes = out.counters.pruned;
      c.direct_pair_evaluations = out.counters.pairs;
      c.accepted_internal_multipoles = out.counters.multipoles;
      c.accepted_leaves = out.counters.leaves;
      c.group_walk_groups = out.group.groups;
      c.group_walk_grouped_targets = out.group.grouped;
      c.group_walk_target_group_size_max = out.group.size_max;
      c.group_walk_fallback_targets = out.group.fallback;
      c.group_walk_fallback_history = out.group.fallback_history;
      c.group_walk_fallback_geometry = out.group.fallback_geometry;
      c.group_walk_candidate_leaves = out.group.candidate_leaves;
      c.group_walk_culled_leaves = out.group.culled;
      c.group_walk_candidate_particles = out.group.candidate_particles;
      c.group_walk_list_entries_max = out.group.entries_max;
      *counters = c;
    }
  })
}

// Interaction list of one group of source rows (all with history prev[i]) in
// the tree of all n sources: accepted nodes (geometry/moments via
// oracle_tree_copy) and listed leaves. Returns the group box in box[6]
// (centre, half) and the list sizes in sizes[2]; nodes / leaves get at most
// cap entries each.
ORACLE_API int oracle_group_list(const double* x, const double* y, const double* z, const double* m, uint64_t n,
                                 const uint32_t* rows, uint64_t k, const double* prev,
                                 const chui_treepm_options* o, const chui_treepm_adaptive_options* ad,
                                 const int32_t mesh[3], uint64_t cap, uint32_t* nodes, uint32_t* leaves,
                                 uint64_t* sizes, double* box) {
  ORACLE_TRY({
    if (k == 0) throw std::invalid_argument("empty group");
    const int msh[3] = {mesh[0], mesh[1], mesh[2]};
    const double L[3] = {o->box_size[0], o->box_size[1], o->box_size[2]};
    const double* src[3] = {x, y, z};
    std::vector<double> u[3];
    for (int a = 0; a < 3; ++a) {
      double lo = INFINITY, hi = -INFINITY;
      for (size_t i = 0; i < n; ++i) { lo = std::min(lo, src[a][i]); hi = std::max(hi, src[a][i]); }
      u[a].resize(n);
      unwrap_axis(src[a], n, L[a], u[a].data(), lo >= 0.0 && hi < L[a]);
    }
    g_tree = Tree{};
    build_tree(u[0].data(), u[1].data(), u[2].data(), m, n, static_cast<size_t>(o->max_leaf_size),
               o->multipole_order == 2, o->softening_epsilon, g_tree);
    double amin = INFINITY, ref[3] = {0, 0, 0}, glo[3] = {0, 0, 0}, ghi[3] = {0, 0, 0};
    for (uint64_t i = 0; i < k; ++i) {
      amin = std::min(amin, prev[i]);
      const double p[3] = {x[rows[i]], y[rows[i]], z[rows[i]]};
      for (int a = 0; a < 3; ++a) {
        if (i == 0) { ref[a] = p[a]; continue; }
        const double d = min_image(p[a] - ref[a], L[a]);
        glo[a] = std::min(glo[a], d);
        ghi[a] = std::max(ghi[a], d);
      }
    }
    double c[3], hg[3];
    for (int a = 0; a < 3; ++a) {
      c[a] = ref[a] + 0.5 * (glo[a] + ghi[a]);
      hg[a] = 0.5 * (ghi[a] - glo[a]);
      box[a] = c[a];
      box[3 + a] = hg[a];
    }
    WalkParams P{};
    for (int a = 0; a < 3; ++a) P.L[a] = L[a];
    const double h = mesh_spacing(*o, msh);
    P.rcut = o->rcut_cells * h;
    P.rcut2 = P.rcut * P.rcut;
    std::vector<uint32_t> stack;
    GroupList list;
    WalkCounters C;
    group_list(g_tree, c, hg, amin, o->relative_force_tolerance / o->gravitational_constant,
               ad->maximum_opening_angle * ad->maximum_opening_angle, ad->cutoff_node_policy != CHUI_CUTOFF_COM_DISTANCE,
               P, stack, list, C);
    sizes[0] = list.nodes.size();
    sizes[1] = list.begin.size();
    for (size_t i = 0; i < list.nodes.size() && i < cap; ++i) nodes[i] = list.nodes[i];
    // listed leaves as node indices: the leaf whose particle range starts at begin
    size_t j = 0;
    for (size_t nd = 0; nd < g_tree.size() && j < list.begin.size(); ++nd)
      if (g_tree.child_count[nd] == 0 && g_tree.begin[nd] == list.begin[j]) {
        if (j < cap) leaves[j] = static_cast<uint32_t>(nd);
        ++j;
      }
  })
}

// Gaussian table H(t) (dimensionlessCoefficient) and the group kernel's
// clamped Hermite form at n points; the knots when knots != NULL (2 x 4097).
ORACLE_API int oracle_gauss_table(const double* t, uint64_t n, double* coefficient, double* hermite,
                                  double* knots) {
  ORACLE_TRY({
    const GaussTable& T = gauss_table();
    const double tlim = std::nextafter(64.0, 0.0);
    for (uint64_t i = 0; i < n; ++i) {
      coefficient[i] = T.coefficient(t[i]);
      hermite[i] = T.hermite(std::min(t[i], tlim));
    }
    if (knots)
      for (size_t i = 0; i <= GaussTable::kCells; ++i) {
        knots[i] = T.value[i];
        knots[GaussTable::kCells + 1 + i] = T.slope[i];
      }
  })
}

// Radial residual coefficients (value, f'/r, (f'' - f'/r)/r^2) at x = r^2.
ORACLE_API int oracle_residual_radial(const double* x, uint64_t n, double eps, double a, double* out3) {
  ORACLE_TRY({
    for (uint64_t i = 0; i < n; ++i) {
      const Radial R = residual_radial(x[i], eps, a);
      out3[3 * i] = R.value; out3[3 * i + 1] = R.first_over_r; out3[3 * i + 2] = R.contracted_second;
    }
  })
}

// One source cluster evaluated as one accepted node with the radial-coefficient
// expansion (radial = 1) or the reference expansion (radial = 0).
ORACLE_API int oracle_cluster_node_accel_radial(const double* x, const double* y, const double* z,
                                                const double* m, uint64_t n, const double t[3], double G,
                                                double eps, double a, int quad, int radial, double out[3]) {
  ORACLE_TRY({
    if (n == 0) throw std::invalid_argument("empty cluster");
    Tree tr;
    build_tree(x, y, z, m, n, n, quad != 0, eps, tr);
    WalkParams P{};
    P.G = G; P.a = a; P.eps = eps; P.quad = quad != 0;
    const double dx = tr.comx[0] - t[0], dy = tr.comy[0] - t[1], dz = tr.comz[0] - t[2];
    const auto c = radial ? node_accel_radial(tr, 0, dx, dy, dz, P) : node_accel(tr, 0, dx, dy, dz, P);
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
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

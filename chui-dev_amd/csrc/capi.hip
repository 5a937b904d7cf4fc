// capi.hip — TreePM coordinator on the device and the C ABI around it.
//
// One chui_treepm_context plays the role of gravity::TreePmCoordinator
// (tree_pm_coupling.hpp:527-814): it owns the mesh workspaces, the tree and all
// scratch across calls (grow-only device buffers, so steady-state solves do not
// allocate), and evaluates PM + short-range residual accelerations for an
// active set. Every exported function catches all exceptions and returns a
// chui_status; the message is kept per thread for chui_treepm_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/chui_treepm.h"
#include "common.hpp"
#include "kdk.hpp"
#include "pm.hpp"
#include "primitives.hpp"
#include "tree.hpp"
#include "walk.hpp"

using chui::DeviceBuffer;

namespace {

thread_local std::string g_last_error;

enum Stage { kEvStart, kEvTree, kEvMultipoles, kEvPm, kEvTargets, kEvWalk, kEvCount };

}  // namespace

struct chui_treepm_context {
  int device = 0;
  int mesh[3] = {0, 0, 0};
  int mesh_lg[3] = {0, 0, 0};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  chui::Tree tree;
  chui::PmWorkspace pm64, pm32;
  chui::PmWorkspace vjp_fwd, vjp_adj;      // force VJP: forward field and adjoint potential
  DeviceBuffer vjp_lam, vjp_comp[3];       // force VJP: tree-ordered lam and deposit lanes
  // force VJP with accepted quadrupole nodes: node accumulators, the accepted
  // list (and sort scratch), its count, per-source corrections
  DeviceBuffer vjp_nacc, vjp_rec, vjp_rec_alt, vjp_vals, vjp_vals_alt, vjp_sort, vjp_count, vjp_corr;
  std::size_t vjp_accepted_cap = 0;
  std::size_t vjp_safe_cap = 0;  // accepted-list slots a recorded VJP sorts (learned by eager VJPs)
  DeviceBuffer src64, src32, htab;  // tree-ordered packed sources; H(t) table
  double htab_key[2] = {0.0, 0.0};
  DeviceBuffer h64tab;              // FP64 box walk: degree-5 cells of H(t)
  double h64_key[2] = {0.0, 0.0};
  // mixed timed walk: bit-indexed residual cells, keyed by (scale, eps, r_s)
  DeviceBuffer ltab;
  std::vector<float> ltab_host;
  double ltab_key[3] = {0.0, -1.0, 0.0};
  uint32_t ltab_cells = 0;
  int32_t ltab_first_octave = 0;
  DeviceBuffer hx, hy, hz, hm, hactive, hout[9];
  DeviceBuffer tpos, tslot, tkeys_alt, tvals_alt, tsort_scratch, flag;
  DeviceBuffer counters, red_scratch, ewald_scratch;
  // adaptive policy: Gaussian Hermite knots, leaf of each tree position,
  // group-walk fallback targets and counters, source bounds
  DeviceBuffer gtab, gcoef, leaf_of_pos, fb_pos, fb_slot, gcounters, bounds;
  // mixed adaptive solves: FP64 short-range sums (3 lanes by output slot) and
  // nearest-pair distances (by source row) of the FP64 per-target walk
  DeviceBuffer mix_fb, mix_near;
  hipEvent_t ev[kEvCount] = {};
  bool events = false;
  // profiling mode: events recorded every solve, harvested at the next solve's
  // first synchronisation point or by chui_treepm_profile_read
  bool profiling = false;
  bool pending = false;
  double prof_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t prof_solves = 0;
  std::vector<float> host_float;
  // Walk traversal fault: the walks set counters[kFaultSlot] (sticky: the
  // per-solve reset leaves it) when a wave's lanes all left the tree early.
  // Every solve copies it to pinned host memory behind fault_ev; it is
  // checked at the next synchronisation through the library or the next solve.
  uint64_t* fault_host = nullptr;
  hipEvent_t fault_ev = nullptr;
  bool fault_pending = false;
  // Capture-safe solves (chui_treepm_set_capture_safe): strict solves (mixed
  // or FP64) and FP64 force VJPs that read nothing back and allocate nothing
  // once a solve of the same sources has sized the context, so that they can
  // be recorded into a HIP graph; `capturing` while the context's stream
  // records one through chui_treepm_graph_begin.
  bool capture_safe = false;
  bool capturing = false;
  // launch shapes (chui_treepm_set_launch_shape, the autotuner's knobs):
  // lines per block of the PM line passes and waves per block of the timed
  // mixed walk; 0 = the built-in choice
  int fft_lines = 0;
  int walk_waves = 0;
};

namespace {

using Ctx = chui_treepm_context;

// ---------------------------------------------------------------------------
chui_status fail(chui_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <class F>
chui_status guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return CHUI_OK;
  } catch (const chui::UnsupportedError& e) {
    return fail(CHUI_ERR_UNSUPPORTED, e.what());
  } catch (const chui::OutOfMemoryError& e) {
    return fail(CHUI_ERR_OUT_OF_MEMORY, e.what());
  } catch (const chui::DeviceError& e) {
    return fail(CHUI_ERR_DEVICE, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(CHUI_ERR_INVALID_ARGUMENT, e.what());
  } catch (const std::overflow_error& e) {
    return fail(CHUI_ERR_OVERFLOW, e.what());
  } catch (const std::logic_error& e) {
    return fail(CHUI_ERR_LOGIC, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CHUI_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(CHUI_ERR_RUNTIME, e.what());
  } catch (...) {
    return fail(CHUI_ERR_RUNTIME, "unknown failure");
  }
}

int ilog2_exact(int n) {
  if (n < 2 || (n & (n - 1)) != 0) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

struct Split {
  double h, a, rcut;
};

Split split_of(const chui_treepm_options& o, const int mesh[3]) {
  // gravity_runtime.cpp:574-578 (cube root of the three mesh spacings) and
  // makeTreePmSplitPolicyFromMeshSpacing (tree_pm_split_kernel.hpp:50-63).
  const double hx = o.box_size[0] / mesh[0], hy = o.box_size[1] / mesh[1], hz = o.box_size[2] / mesh[2];
  Split s;
  s.h = std::cbrt(hx * hy * hz);
  s.a = o.asmth_cells * s.h;
  s.rcut = o.rcut_cells * s.h;
  return s;
}

void validate_options(const chui_treepm_options* o, const int mesh[3]) {
  if (!o) throw std::invalid_argument("TreePM options pointer is null");
  if (o->struct_size != sizeof(chui_treepm_options))
    throw std::invalid_argument("chui_treepm_options.struct_size does not match this library");
  for (int a = 0; a < 3; ++a) {
    if (!(std::isfinite(o->box_size[a]) && o->box_size[a] > 0.0))
      throw std::invalid_argument("Periodic TreePM tree geometry requires finite positive axis lengths");
    if (mesh[a] <= 0) throw std::invalid_argument("numerics.treepm_pm_grid_n{xyz} must all be > 0");
  }
  if (!(std::isfinite(o->gravitational_constant) && o->gravitational_constant > 0.0))
    throw std::overflow_error("TreePM code-unit gravitational prefactor is non-finite or non-positive");
  if (!(std::isfinite(o->opening_theta) && o->opening_theta > 0.0))
    throw std::invalid_argument("numerics.treepm_tree_opening_theta must be finite and > 0");
  if (o->max_leaf_size < 1 || o->max_leaf_size > 1024)
    throw std::invalid_argument("numerics.treepm_tree_max_leaf_size must be in [1,1024]");
  const int b = o->residual_block_size;
  if (b != 16 && b != 32 && b != 64 && b != 128 && b != 256)
    throw std::invalid_argument("numerics.treepm_residual_block_size must be 16, 32, 64, 128 or 256");
  if (!(std::isfinite(o->softening_epsilon) && o->softening_epsilon >= 0.0))
    throw std::invalid_argument("Invalid tree gravity options");
  if (!(std::isfinite(o->asmth_cells) && o->asmth_cells > 0.0))
    throw std::invalid_argument("TreePM asmth_cells must be > 0");
  if (!(std::isfinite(o->rcut_cells) && o->rcut_cells > 0.0))
    throw std::invalid_argument("TreePM rcut_cells must be > 0");
  if (o->assignment_scheme != CHUI_ASSIGN_CIC && o->assignment_scheme != CHUI_ASSIGN_TSC)
    throw std::invalid_argument("Unknown PM assignment scheme");
  if (o->opening_criterion < 0 || o->opening_criterion > 2)
    throw std::invalid_argument("Invalid tree gravity opening criterion");
  if (o->multipole_order != 0 && o->multipole_order != 2)
    throw std::invalid_argument("Invalid tree gravity multipole order");
  if (o->precision != CHUI_PRECISION_FP64 && o->precision != CHUI_PRECISION_MIXED)
    throw std::invalid_argument("Unknown precision");
  const Split s = split_of(*o, mesh);
  // config.cpp:3669 and :3682-3692
  if (o->softening_epsilon > 0.20 * s.a)
    throw std::invalid_argument("periodic TreePM Plummer softening exceeds the certified epsilon/r_s <= 0.20 envelope");
  const double lmin = std::min({o->box_size[0], o->box_size[1], o->box_size[2]});
  if (s.rcut >= std::nextafter(0.5 * lmin, 0.0))
    throw std::invalid_argument("periodic TreePM requires r_cut < half the shortest box axis");
  if (o->acceptance_policy != CHUI_ACCEPT_STRICT && o->acceptance_policy != CHUI_ACCEPT_ADAPTIVE)
    throw std::invalid_argument("Unknown TreePM acceptance policy");
  if (o->target_mapping < CHUI_TARGETS_AUTO || o->target_mapping > CHUI_TARGETS_PER_WAVE)
    throw std::invalid_argument("Unknown target mapping");
  if (o->acceptance_policy == CHUI_ACCEPT_ADAPTIVE &&
      !(std::isfinite(o->relative_force_tolerance) && o->relative_force_tolerance > 0.0))
    throw std::invalid_argument("numerics.treepm_tree_relative_force_tolerance must be finite and > 0");
  if (o->opening_criterion == CHUI_OPEN_RELATIVE && o->multipole_order == 2) {
    // Without an acceleration history the relative criterion falls back to the
    // COM-distance MAC (tree_interaction_common.hpp:60-62); handled in the walk.
  }
}

// ---- gather kernels ---------------------------------------------------------
// Tree-order copies of the caller's sources (tree_pm_coupling.cpp:3732-3800):
// FP64 keeps the caller's values verbatim; mixed stores 32-bit fixed-point
// periodic coordinates and the FP32 mass.
__global__ void gather64_kernel(const uint32_t* __restrict__ order, std::size_t n, const double* __restrict__ x,
                                const double* __restrict__ y, const double* __restrict__ z,
                                const double* __restrict__ m, double4* __restrict__ out) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = order[i];
  out[i] = make_double4(x[r], y[r], z[r], m[r]);
}

__device__ __forceinline__ uint32_t fixed_of(double x, double L) {
  // canonical wrap (periodic_domain.hpp:8-16), then 32-bit fixed point
  double w = fmod(x, L);
  if (w < 0.0) w += L;
  if (!(w < L)) w = 0.0;
  const double f = floor(w * (4294967296.0 / L));
  return f >= 4294967295.0 ? 0xFFFFFFFFu : static_cast<uint32_t>(f);
}

// Force VJP: tree-ordered lam (x, y, z, 0) and the three deposit lanes
// (position, lam_a) of the adjoint PM deposit.
__global__ void gather_lam_kernel(const uint32_t* __restrict__ order, std::size_t n, const double4* __restrict__ src,
                                  const double* __restrict__ lx, const double* __restrict__ ly,
                                  const double* __restrict__ lz, double4* __restrict__ lam, double4* __restrict__ c0,
                                  double4* __restrict__ c1, double4* __restrict__ c2) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = order[i];
  const double4 p = src[i];
  const double a = lx[r], b = ly[r], c = lz[r];
  lam[i] = make_double4(a, b, c, 0.0);
  c0[i] = make_double4(p.x, p.y, p.z, a);
  c1[i] = make_double4(p.x, p.y, p.z, b);
  c2[i] = make_double4(p.x, p.y, p.z, c);
}

// from the tree's own tree-order copy (canonical coordinates and mass)
__global__ void gather_fixed_kernel(const double4* __restrict__ pk, std::size_t n, double Lx, double Ly, double Lz,
                                    uint4* __restrict__ out) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 q = pk[i];
  out[i] = make_uint4(fixed_of(q.x, Lx), fixed_of(q.y, Ly), fixed_of(q.z, Lz),
                      __float_as_uint(static_cast<float>(q.w)));
}

__global__ void target_keys_kernel(const uint32_t* __restrict__ active, std::size_t na,
                                   const uint32_t* __restrict__ pos_of_row, std::size_t n, uint32_t* __restrict__ keys,
                                   uint32_t* __restrict__ vals, uint32_t* __restrict__ bad) {
  const std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= na) return;
  const uint32_t r = active[s];
  if (r >= n) {
    atomicOr(bad, 1u);
    keys[s] = 0;
  } else {
    keys[s] = pos_of_row[r];
  }
  vals[s] = static_cast<uint32_t>(s);
}

// A capture-safe solve turns a device flag into fault bits instead of reading
// it back.
__global__ void flag_fault_kernel(const uint32_t* __restrict__ flag, unsigned long long* __restrict__ fault,
                                  unsigned long long bits) {
  if (threadIdx.x == 0 && *flag) atomicOr(fault, bits);
}

template <class T>
__global__ void to_double_kernel(const T* __restrict__ in, double* __restrict__ out, std::size_t n) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<double>(in[i]);
}

double read_device_double(const double* p, hipStream_t s) {
  double v = 0.0;
  CHUI_HIP(hipMemcpyAsync(&v, p, sizeof(double), hipMemcpyDeviceToHost, s));
  CHUI_HIP(hipStreamSynchronize(s));
  return v;
}

void record(Ctx* c, Stage st) {
  if (c->events) CHUI_HIP(hipEventRecord(c->ev[st], c->stream));
}

// Whether the context's stream records a graph: one begun through this library
// (chui_treepm_graph_begin) or by the caller on the stream it handed over
// (torch.cuda.graph, hipStreamBeginCapture).
bool recording(const Ctx* c) {
  if (c->capturing) return true;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CHUI_HIP(hipStreamIsCapturing(c->stream, &cs));
  return cs != hipStreamCaptureStatusNone;
}

// Walk counters with the sticky traversal-fault word after them; allocated and
// zeroed once, so the fault word survives the per-solve counter resets.
unsigned long long* walk_counters(Ctx* c, hipStream_t s) {
  if (!c->counters.ptr) {
    c->counters.reserve(chui::kWalkCounterWords * sizeof(unsigned long long));
    chui::fill_u64(c->counters.as<uint64_t>(), 0ull, chui::kWalkCounterWords, s);
  }
  chui::fill_u64(c->counters.as<uint64_t>(), 0ull, chui::kWalkCounters, s);
  chui::fill_u64(c->counters.as<uint64_t>() + chui::kPairChecksumSlot, 0ull, 1, s);
  return reinterpret_cast<unsigned long long*>(c->counters.ptr);
}

// Copies the fault word to pinned host memory after the stream's prior work.
void post_fault_word(Ctx* c, hipStream_t s) {
  if (!c->fault_host) {
    void* p = nullptr;
    CHUI_HIP(hipHostMalloc(&p, 64, hipHostMallocDefault));
    c->fault_host = static_cast<uint64_t*>(p);
    *c->fault_host = 0;
  }
  if (!c->fault_ev) CHUI_HIP(hipEventCreateWithFlags(&c->fault_ev, hipEventDisableTiming));
  CHUI_HIP(hipMemcpyAsync(c->fault_host, c->counters.as<unsigned long long>() + chui::kFaultSlot,
                          sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  CHUI_HIP(hipEventRecord(c->fault_ev, s));
  c->fault_pending = true;
}

// Reports a posted traversal fault as std::logic_error (CHUI_ERR_LOGIC). With
// wait = false a copy still in flight stays pending for a later check.
void check_fault(Ctx* c, bool wait) {
  if (!c->fault_pending) return;
  if (wait) {
    CHUI_HIP(hipEventSynchronize(c->fault_ev));
  } else {
    const hipError_t q = hipEventQuery(c->fault_ev);
    if (q == hipErrorNotReady) return;
    CHUI_HIP(q);
  }
  c->fault_pending = false;
  const uint64_t f = *c->fault_host;
  if (f != 0) {
    *c->fault_host = 0;
    chui::fill_u64(c->counters.as<uint64_t>() + chui::kFaultSlot, 0ull, 1, c->stream);
    // capture-safe solves report here what an ordinary solve throws at once
    if (f & chui::kFaultCoordinates)
      throw std::invalid_argument("Periodic TreePM tree geometry requires finite source coordinates");
    if (f & chui::kFaultMasses)
      throw std::invalid_argument("Tree gravity requires finite coordinates and finite non-negative masses");
    if (f & chui::kFaultActive) throw std::invalid_argument("TreePM active particle index out of range");
    if (f & chui::kFaultTooDeep)
      throw chui::UnsupportedError(
          "octree deeper than 21 levels (more than max_leaf_size coincident or near-coincident particles) "
          "is not built by the GPU path yet");
    if (f & chui::kFaultPrepared)
      throw std::runtime_error(
          "capture-safe solve: the sources outgrew the prepared context (node capacity, tree depth or mass "
          "uniformity); run an ordinary solve of the new sources and capture again");
    throw std::logic_error("TreePM walk lost progress (wave reduction selected no pending node)");
  }
}

float elapsed(Ctx* c, Stage a, Stage b) {
  float ms = 0.0f;
  CHUI_HIP(hipEventElapsedTime(&ms, c->ev[a], c->ev[b]));
  return ms;
}

void harvest(Ctx* c) {
  if (!c->pending) return;
  CHUI_HIP(hipEventSynchronize(c->ev[kEvWalk]));
  c->prof_ms[0] += elapsed(c, kEvStart, kEvWalk);
  c->prof_ms[1] += elapsed(c, kEvStart, kEvTree);
  c->prof_ms[2] += elapsed(c, kEvTree, kEvMultipoles);
  c->prof_ms[3] += elapsed(c, kEvMultipoles, kEvPm);
  c->prof_ms[4] += elapsed(c, kEvPm, kEvTargets);
  c->prof_ms[5] += elapsed(c, kEvTargets, kEvWalk);
  c->prof_solves += 1;
  c->pending = false;
}

// Walk parameters shared by the force walk and the force VJP walk (FP64 and
// mixed fields that depend only on the options, the split and the mesh).
chui::WalkParams base_walk_params(const Ctx* c, const chui_treepm_options* o, const Split& sp, std::size_t mesh_row) {
  chui::WalkParams wp{};
  wp.walk_waves = c->walk_waves > 0 ? c->walk_waves : 4;
  for (int a = 0; a < 3; ++a) {
    wp.L[a] = o->box_size[a];
    wp.mesh[a] = c->mesh[a];
    wp.mesh_log2[a] = c->mesh_lg[a];
    wp.fx_to_len[a] = static_cast<float>(o->box_size[a] / 4294967296.0);
  }
  wp.G = o->gravitational_constant;
  wp.a = sp.a;
  wp.rcut = sp.rcut;
  wp.rcut2 = sp.rcut * sp.rcut;
  wp.theta = o->opening_theta;
  wp.eps = o->softening_epsilon;
  wp.criterion = o->opening_criterion == CHUI_OPEN_GEOMETRIC ? 0 : 1;
  wp.quad = o->multipole_order == 2;
  wp.assignment = o->assignment_scheme;
  wp.c4 = static_cast<float>(0.25 / (sp.a * sp.a));
  wp.c8 = static_cast<float>(0.125 / (sp.a * sp.a * sp.a));
  wp.cubic = (o->box_size[0] == o->box_size[1] && o->box_size[0] == o->box_size[2]) ? 1 : 0;
  {
    // cutoff radii in separation units: fixed-point units for a cubic box
    const double sep = wp.cubic ? o->box_size[0] / 4294967296.0 : 1.0;
    const double r2 = wp.rcut2 / (sep * sep);
    wp.rcut2_sep = static_cast<float>(r2);
    wp.prune2_sep = static_cast<float>(r2 * (1.0 + 1.0e-5));
    wp.contain2_sep = static_cast<float>(r2 * (1.0 - 1.0e-5));
  }
  for (int a = 0; a < 3; ++a) {
    // box-classified walk for waves whose targets lie within +-2 r_cut of the
    // box centre per axis. At 256^3 a bound of +-r_cut leaves 79% of the
    // waves to the lane-by-lane walk; +-2 r_cut measured 13% faster in total,
    // +-3 and +-4 r_cut slower again (more sources per wider box).
    const double sep = wp.cubic ? o->box_size[a] / 4294967296.0 : 1.0;
    wp.box_half_max[a] = static_cast<float>(2.0 * sp.rcut / sep);
  }
  for (int a = 0; a < 3; ++a) {
    // float offsets from a wave origin: targets within L/16 (precision) and
    // within L/2 - r_cut (no wrap of the origin's frame inside a pair), with a
    // margin of 2^24 units for the r_cut conversion
    const double rcut_units = sp.rcut / o->box_size[a] * 4294967296.0;
    const double lim = std::min(268435456.0, 2147483648.0 - rcut_units - 16777216.0);
    wp.fast_half_extent[a] = lim > 0.0 ? static_cast<uint32_t>(lim) : 0u;
  }
  wp.node_count = static_cast<uint32_t>(c->tree.node_count);
  wp.mesh_row = mesh_row;
  return wp;
}

// The FP64 box walk's table of H on [0, t_max] (t_max just above r_cut^2 / 4a^2)
// and the parameters that go with it.
void prepare_h64(Ctx* c, chui::WalkParams& wp, const Split& sp, hipStream_t s) {
  const double t_max = wp.rcut2 * (0.25 / (sp.a * sp.a)) * (1.0 + 1.0e-3);
  const int cells = static_cast<int>(std::min<double>(chui::kH64MaxCells, std::max(16.0, std::ceil(t_max / 0.04))));
  if (c->h64_key[0] != t_max || c->h64_key[1] != cells || !c->h64tab.ptr) {
    const std::vector<double> tab = chui::gaussian_quintic_table(t_max, cells);
    chui::require_not_capturing(s, "the FP64 residual table");
    c->h64tab.reserve(tab.size() * sizeof(double));
    CHUI_HIP(hipMemcpyAsync(c->h64tab.ptr, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s));
    CHUI_HIP(hipStreamSynchronize(s));
    c->h64_key[0] = t_max;
    c->h64_key[1] = cells;
  }
  wp.h64_cells = cells;
  wp.h64_scale = (0.25 / (sp.a * sp.a)) / (t_max / cells);
  wp.eps2_64 = wp.eps * wp.eps;
  wp.c8_64 = ((0.125 / sp.a) / sp.a) / sp.a;
}

// The mixed walks' pair arithmetic in H-table units (lengths times
// sqrt(t_to_cell), so r^2 is the coordinate of the 128-cell table of H) and
// the table itself, uploaded when its range or split scale changes.
void prepare_mixed_pairs(Ctx* c, chui::WalkParams& wp, const Split& sp, const chui_treepm_options* o,
                         bool uniform, double um, hipStream_t s) {
  const double t_max = wp.rcut2 * (0.25 / (sp.a * sp.a)) * (1.0 + 1.0e-3);
  const std::vector<float> tab = chui::gaussian_coefficient_table(wp, t_max);  // also sets wp.t_to_cell
  const double t2c = 0.25 / (sp.a * sp.a) / (t_max / chui::kHTableCells);
  const double sq = std::sqrt(t2c);
  for (int a = 0; a < 3; ++a) wp.tab_scale[a] = static_cast<float>(o->box_size[a] / 4294967296.0 * sq);
  wp.rcut2_tab = static_cast<float>(wp.rcut2 * t2c);
  // floor 1e-12 (table units, r^2 <= 128 for every pair inside r_cut): a
  // coincident pair (r^2 = 0) keeps a finite softened term (< 1e18 times
  // its mass), so its zero separation gives a zero force
  wp.eps2_tab = static_cast<float>(std::max(wp.eps * wp.eps * t2c, 1.0e-12));
  wp.c8_tab = static_cast<float>(0.125 / (sp.a * sp.a * sp.a) / (t2c * sq));
  // with uniform masses the pair kernel leaves the mass to this factor
  wp.force_scale = static_cast<float>(wp.G * t2c * (uniform ? um : 1.0));
  wp.tab_to_len2 = static_cast<float>(1.0 / t2c);
  wp.accept_half_max = static_cast<float>(0.041 * sp.rcut);
  if (c->htab_key[0] != t_max || c->htab_key[1] != sp.a || !c->htab.ptr) {
    chui::require_not_capturing(s, "the pair table");
    c->htab.reserve(tab.size() * sizeof(float));
    CHUI_HIP(hipMemcpyAsync(c->htab.ptr, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, s));
    CHUI_HIP(hipStreamSynchronize(s));
    c->htab_key[0] = t_max;
    c->htab_key[1] = sp.a;
  }
}

// TreePmOptions defaults of the adaptive/group-walk fields (tree_pm_coupling.hpp:248-270).
chui_treepm_adaptive_options adaptive_defaults() {
  chui_treepm_adaptive_options a{};
  a.struct_size = sizeof(a);
  a.error_estimator = CHUI_ESTIMATOR_RESIDUAL_DERIVATIVE;
  a.maximum_opening_angle = 0.25;
  a.cutoff_node_policy = CHUI_CUTOFF_CONTAIN_NODE;
  a.group_max_extent_cutoff_fraction = 0.5;
  return a;
}

void validate_adaptive(const chui_treepm_adaptive_options& a) {
  if (a.struct_size != sizeof(chui_treepm_adaptive_options))
    throw std::invalid_argument("chui_treepm_adaptive_options.struct_size does not match this library");
  if (!(std::isfinite(a.maximum_opening_angle) && a.maximum_opening_angle > 0.0))
    throw std::invalid_argument("numerics.treepm_adaptive_maximum_opening_angle must be finite and > 0");
  if (a.error_estimator != CHUI_ESTIMATOR_RESIDUAL_DERIVATIVE && a.error_estimator != CHUI_ESTIMATOR_RELATIVE_MONOPOLE)
    throw std::invalid_argument("Unknown TreePM adaptive error estimator");
  if (a.cutoff_node_policy != CHUI_CUTOFF_CONTAIN_NODE && a.cutoff_node_policy != CHUI_CUTOFF_COM_DISTANCE)
    throw std::invalid_argument("Unknown TreePM cutoff node policy");
  if (!(std::isfinite(a.group_max_extent_cutoff_fraction) && a.group_max_extent_cutoff_fraction > 0.0))
    throw std::invalid_argument("numerics.treepm_group_walk_max_extent_cutoff_fraction must be finite and > 0");
}

// Per-axis min/max of the caller's source coordinates (the bounds of the
// bounded periodic metric, tree_pm_geometry.hpp:27-32), as order keys.
__global__ void source_bounds_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                     const double* __restrict__ z, std::size_t n,
                                     unsigned long long* __restrict__ keys) {
  const double* src[3] = {x, y, z};
  uint64_t lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0, 0, 0};
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    for (int a = 0; a < 3; ++a) {
      const uint64_t k = chui::order_key(src[a][i]);
      lo[a] = k < lo[a] ? k : lo[a];
      hi[a] = k > hi[a] ? k : hi[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    const uint64_t l = chui::wave_min(lo[a]), h = chui::wave_max(hi[a]);
    if (chui::lane_id() == 0) {
      atomicMin(&keys[a], static_cast<unsigned long long>(l));
      atomicMax(&keys[3 + a], static_cast<unsigned long long>(h));
    }
  }
}

// Core force evaluation on device arrays (tree_pm_coupling.cpp:2282-3519 for
// one rank: PM refresh/reuse, PM interpolation, unwrap + build, residual walk).
void solve_device_impl(Ctx* c, const double* x, const double* y, const double* z, const double* m, uint64_t n,
                       const uint32_t* active, uint64_t n_active, void* ax, void* ay, void* az, void* pmx,
                       void* pmy, void* pmz, void* srx, void* sry, void* srz, const chui_treepm_options* o,
                       int refresh, chui_treepm_diagnostics* diag, const double* prev = nullptr,
                       const chui_treepm_adaptive_options* adaptive = nullptr,
                       chui_treepm_walk_counters* wcount = nullptr, void* near2 = nullptr) {
  validate_options(o, c->mesh);
  const bool adaptive_policy = o->acceptance_policy == CHUI_ACCEPT_ADAPTIVE;
  const chui_treepm_adaptive_options ad = adaptive ? *adaptive : adaptive_defaults();
  validate_adaptive(ad);
  if (wcount && wcount->struct_size != sizeof(chui_treepm_walk_counters))
    throw std::invalid_argument("chui_treepm_walk_counters.struct_size does not match this library");
  if (n > 0xFFFFFFFEull) throw std::overflow_error("Tree gravity source count exceeds the 32-bit tree-index contract");
  if (active == nullptr) n_active = n;
  if (n_active > 0 && (!ax || !ay || !az)) throw std::invalid_argument("TreePM accumulator output lanes are null");
  if (n > 0 && (!x || !y || !z || !m)) throw std::invalid_argument("TreePM source lanes are null");
  if ((pmx || pmy || pmz) && !(pmx && pmy && pmz))
    throw std::invalid_argument("TreePM long-range capture requires all three lanes");
  if ((srx || sry || srz) && !(srx && sry && srz))
    throw std::invalid_argument("TreePM short-range lanes require all three lanes");
  const bool mixed = o->precision == CHUI_PRECISION_MIXED;
  const bool sr_only = o->short_range_only != 0;
  // tree_pm_coupling.cpp:2508-2511
  if (sr_only && (refresh || pmx || pmy || pmz))
    throw std::invalid_argument("short-range split surface cannot request a PM refresh/capture or zoom correction");
  // Capture-safe mode: the first solve of a source count is an ordinary one
  // that sizes the context for it (and prepares the node capacity, depth and
  // mass uniformity); the solves after it read nothing back. Only those can be
  // recorded into a graph.
  const bool safe_mode = c->capture_safe;
  const bool safe = safe_mode && n > 0 && c->tree.prepared_n == n && c->tree.sized_for_capture;
  const bool rec = recording(c);
  if (rec && !safe)
    throw std::runtime_error(
        "graph capture of a solve on a context not prepared for these sources: run one capture-safe solve of "
        "them before capturing");
  if (safe_mode && (adaptive_policy || diag || wcount || near2 || prev))
    throw chui::UnsupportedError(
        "capture-safe solves are strict solves without diagnostics, counters, the nearest-pair lane or an "
        "acceleration history");
  hipStream_t s = c->stream;
  if (!rec) {
    harvest(c);
    check_fault(c, false);
  }
  c->events = !rec && (diag != nullptr || c->profiling);
  if (c->events && !c->ev[0])
    for (auto& e : c->ev) CHUI_HIP(hipEventCreate(&e));
  record(c, kEvStart);
  const Split sp = split_of(*o, c->mesh);
  // the counters' fault word takes a capture-safe build's fault bits
  unsigned long long* const walk_cnt = walk_counters(c, s);

  // --- tree ---
  chui::TreeParams tp;
  tp.n = n;
  for (int a = 0; a < 3; ++a) tp.L[a] = o->box_size[a];
  tp.leaf_size = static_cast<uint32_t>(o->max_leaf_size);
  tp.quadrupole = o->multipole_order == 2;
  tp.mixed = mixed;
  tp.capture_safe = safe;
  tp.size_for_capture = safe_mode && !safe;
  tp.fault = walk_cnt + chui::kFaultSlot;
  chui::build_tree(c->tree, x, y, z, m, tp, o->softening_epsilon, s);
  record(c, kEvTree);
  const bool uniform = n > 0 && c->tree.mass_uniform;
  const double um = uniform ? c->tree.uniform_mass : 0.0;

  // --- tree-ordered source lanes ---
  if (n > 0) {
    if (!mixed) {
      c->src64.reserve(n * sizeof(double4));
      gather64_kernel<<<chui::grid_for(n, 256), 256, 0, s>>>(c->tree.order.as<uint32_t>(), n, x, y, z, m,
                                                             c->src64.as<double4>());
      CHUI_LAUNCH_CHECK("gather64_kernel");
    } else {
      c->src32.reserve(n * sizeof(uint4));
      gather_fixed_kernel<<<chui::grid_for(n, 256), 256, 0, s>>>(c->tree.pk.as<double4>(), n, o->box_size[0],
                                                                 o->box_size[1], o->box_size[2],
                                                                 c->src32.as<uint4>());
      CHUI_LAUNCH_CHECK("gather_fixed_kernel");
      if (adaptive_policy) {
        // the adaptive decisions (groups, histories, acceptance) are made in
        // FP64 also in mixed precision: FP64 sources and node records
        c->src64.reserve(n * sizeof(double4));
        gather64_kernel<<<chui::grid_for(n, 256), 256, 0, s>>>(c->tree.order.as<uint32_t>(), n, x, y, z, m,
                                                               c->src64.as<double4>());
        CHUI_LAUNCH_CHECK("gather64_kernel");
        chui::ensure_records64(c->tree, s);
      }
    }
  }
  record(c, kEvMultipoles);

  // --- PM (refresh or cadence reuse) ---
  chui::PmGeometry g{};
  for (int a = 0; a < 3; ++a) {
    g.n[a] = c->mesh[a];
    g.lg[a] = c->mesh_lg[a];
    g.L[a] = o->box_size[a];
  }
  g.nzc = static_cast<std::size_t>(c->mesh[2] / 2 + 1);
  g.row = 2 * g.nzc;
  chui::PmOptions po;
  po.G = o->gravitational_constant;
  po.split_scale = sp.a;
  po.assignment = o->assignment_scheme;
  po.deconvolution = o->window_deconvolution;
  po.fft_lines = c->fft_lines;
  chui::PmWorkspace& pw = mixed ? c->pm32 : c->pm64;
  const double key[8] = {o->box_size[0], o->box_size[1], o->box_size[2], o->gravitational_constant, sp.a,
                         static_cast<double>(o->assignment_scheme), static_cast<double>(o->window_deconvolution),
                         o->scale_factor};
  // tree_pm_coupling.cpp:2483-2506: exact reuse needs the same source state
  // (generation and force epoch); the stale window admits any earlier one.
  // The field version and the mesh-level options must match either way.
  const bool source_admitted =
      o->allow_stale_long_range_field != 0 ||
      (pw.source_generation == o->source_generation && pw.epoch_sequence == o->force_epoch_sequence &&
       pw.epoch_scale_factor == o->force_epoch_scale_factor);
  const bool key_match =
      pw.valid && source_admitted && pw.field_version == o->pm_field_version && std::equal(key, key + 8, pw.key);
  if (!sr_only && !refresh && !key_match) {
    if (!pw.valid) throw std::runtime_error("TreePM long-range mesh field is unavailable for reuse");
    throw std::runtime_error("TreePM long-range reuse requested without a compatible PM field");
  }
  if (!sr_only && refresh) {
    chui::DepositTiling dt{};
    for (int a = 0; a < 3; ++a) {
      dt.n[a] = g.n[a];
      dt.inv_h[a] = static_cast<double>(g.n[a]) / g.L[a];
    }
    dt.scheme = po.assignment;
    dt.max_cells = chui::kDepositTileBytes / (mixed ? sizeof(float) : sizeof(double));
    dt.chunk = chui::kDepositChunk;
    const chui::DepositUnits units = chui::build_deposit_units(c->tree, dt, s);
    if (!mixed) {
      chui::pm_solve_fp64(pw, g, po, c->src64.as<double4>(), uniform ? 1 : 0, um, n, units, s);
    } else {
      chui::pm_solve_mixed(pw, g, po, c->src32.as<uint4>(), uniform ? 1 : 0, static_cast<float>(um), n, units, s);
    }
    std::copy(key, key + 8, pw.key);
    pw.source_generation = o->source_generation;
    pw.field_version = o->pm_field_version;
    pw.epoch_sequence = o->force_epoch_sequence;
    pw.epoch_scale_factor = o->force_epoch_scale_factor;
  }
  record(c, kEvPm);

  // --- targets ---
  const uint32_t* target_pos = nullptr;
  const uint32_t* target_slot = nullptr;
  int identity = 1;
  if (active != nullptr && n_active > 0) {
    chui::build_pos_of_row(c->tree, s);
    c->tpos.reserve(n_active * 4);
    c->tslot.reserve(n_active * 4);
    c->tkeys_alt.reserve(n_active * 4);
    c->tvals_alt.reserve(n_active * 4);
    c->tsort_scratch.reserve(chui::radix_sort_scratch_bytes(n_active));
    c->flag.reserve(16);
    chui::fill_u32(c->flag.as<uint32_t>(), 0u, 1, s);
    target_keys_kernel<<<chui::grid_for(n_active, 256), 256, 0, s>>>(active, n_active,
                                                                     c->tree.pos_of_row.as<uint32_t>(), n,
                                                                     c->tpos.as<uint32_t>(), c->tslot.as<uint32_t>(),
                                                                     c->flag.as<uint32_t>());
    CHUI_LAUNCH_CHECK("target_keys_kernel");
    if (safe) {
      flag_fault_kernel<<<1, 64, 0, s>>>(c->flag.as<uint32_t>(), walk_cnt + chui::kFaultSlot, chui::kFaultActive);
      CHUI_LAUNCH_CHECK("flag_fault_kernel");
    } else {
      uint32_t bad = 0;
      c->tree.readback.read(&bad, c->flag.ptr, 4, s);
      if (bad) throw std::invalid_argument("TreePM active particle index out of range");
    }
    // Visit targets in tree order (tree_pm_coupling.cpp:3801-3838).
    chui::radix_sort_pairs32(c->tpos.as<uint32_t>(), c->tslot.as<uint32_t>(), c->tkeys_alt.as<uint32_t>(),
                             c->tvals_alt.as<uint32_t>(), n_active, 0, 32, c->tsort_scratch.ptr, s);
    target_pos = c->tpos.as<uint32_t>();
    target_slot = c->tslot.as<uint32_t>();
    identity = 0;
  }
  record(c, kEvTargets);

  // --- walk (+ fused PM interpolation) ---
  chui::WalkParams wp = base_walk_params(c, o, sp, g.row);
  wp.n_targets = static_cast<uint32_t>(n_active);
  wp.identity_targets = identity;
  // the mixed walk's counting variant also produces the nearest-pair lane
  wp.count = (diag != nullptr || wcount != nullptr || (mixed && near2 != nullptr)) ? 1 : 0;
  // one target per wave when asked, or when few sources are active: 64
  // visit-ordered targets then lie far apart and a shared walk covers the
  // union of their neighbourhoods, costing about a full walk whatever the
  // active count. The measured short-range-only crossovers fall with the
  // source count: mixed 2-3% active at 256^3 (clustered, 1065 pairs per
  // target), 12% at 128^3 and 35% at 64^3; FP64 15% at 128^3. Auto switches
  // at f(n) = min(0.25, f0 (2^24 / n)^0.75), f0 = 0.025 (mixed) or 0.03 (FP64).
  const double f0 = mixed ? 0.025 : 0.03;
  const double f_switch = std::min(0.25, f0 * std::pow(16777216.0 / static_cast<double>(std::max<uint64_t>(n, 1)), 0.75));
  const bool sparse = static_cast<double>(n_active) <= f_switch * static_cast<double>(n);
  // (strict walks without counters; FP64 also without the nearest-pair lane)
  wp.target_per_wave = !wp.count && (mixed || near2 == nullptr) &&
                       (o->target_mapping == CHUI_TARGETS_PER_WAVE ||
                        (o->target_mapping == CHUI_TARGETS_AUTO && sparse)) ? 1 : 0;
  wp.uniform_mass = um;
  wp.uniform_mass_f = static_cast<float>(um);
  wp.pm_enabled = sr_only ? 0 : 1;
  chui::WalkOutputs out{ax, ay, az, pmx, pmy, pmz, srx, sry, srz, near2};
  uint32_t group_status = ad.group_walk ? CHUI_GROUP_WALK_INCOMPATIBLE_POLICY : CHUI_GROUP_WALK_NOT_REQUESTED;
  const bool group_counts = adaptive_policy || wcount != nullptr;
  if (group_counts) {
    c->gcounters.reserve(chui::kGroupCounters * sizeof(unsigned long long));
    chui::fill_u64(c->gcounters.as<uint64_t>(), 0ull, chui::kGroupCounters, s);
  }
  if (n_active > 0 && n > 0 && adaptive_policy) {
    // Adaptive acceptance (tree_pm_acceptance.hpp:54-139) and, when eligible,
    // the group walk (eligibility tree_pm_coupling.cpp:3839-3869, :4282-4286).
    wp.estimator = ad.error_estimator;
    wp.com_cut = ad.cutoff_node_policy == CHUI_CUTOFF_COM_DISTANCE ? 1 : 0;
    wp.lookup = ad.gaussian_pair_lookup ? 1 : 0;
    wp.radial = ad.radial_coefficients ? 1 : 0;
    wp.amax = ad.maximum_opening_angle;
    wp.alpha = o->relative_force_tolerance;
    wp.accel_floor = o->relative_force_acceleration_floor;
    wp.inv2a2 = (0.25 / sp.a) / sp.a;
    wp.inv8a3 = ((0.125 / sp.a) / sp.a) / sp.a;
    if (!c->gtab.ptr) {
      const std::vector<double> knots = chui::gaussian_hermite_knots();
      const std::vector<double> mono = chui::gaussian_hermite_monomials(knots);
      c->gtab.reserve(knots.size() * sizeof(double));
      c->gcoef.reserve(mono.size() * sizeof(double));
      CHUI_HIP(hipMemcpyAsync(c->gtab.ptr, knots.data(), knots.size() * sizeof(double), hipMemcpyHostToDevice, s));
      CHUI_HIP(hipMemcpyAsync(c->gcoef.ptr, mono.data(), mono.size() * sizeof(double), hipMemcpyHostToDevice, s));
      CHUI_HIP(hipStreamSynchronize(s));
    }
    const double t_cut = (0.25 * wp.rcut2 / sp.a) / sp.a;
    if (!ad.group_walk) group_status = CHUI_GROUP_WALK_NOT_REQUESTED;
    else if (c->tree.node_count == 0) group_status = CHUI_GROUP_WALK_NO_LOCAL_TREE;
    else if (ad.error_estimator != CHUI_ESTIMATOR_RELATIVE_MONOPOLE || !ad.gaussian_pair_lookup)
      group_status = CHUI_GROUP_WALK_INCOMPATIBLE_POLICY;
    else if (!(std::isfinite(t_cut) && t_cut < 63.0)) group_status = CHUI_GROUP_WALK_CUTOFF_OUTSIDE_TABLE;
    else if (!(std::isnormal(wp.inv2a2) && std::isnormal(wp.inv8a3)))
      group_status = CHUI_GROUP_WALK_LOOKUP_SCALING_UNREPRESENTABLE;
    else group_status = CHUI_GROUP_WALK_ACTIVE;
    const double2* gtab = c->gtab.as<double2>();
    const double* f0 = mixed ? nullptr : static_cast<const double*>(pw.force[0]);
    const double* f1 = mixed ? nullptr : static_cast<const double*>(pw.force[1]);
    const double* f2 = mixed ? nullptr : static_cast<const double*>(pw.force[2]);
    const float* g0 = mixed ? static_cast<const float*>(pw.force[0]) : nullptr;
    const float* g1 = mixed ? static_cast<const float*>(pw.force[1]) : nullptr;
    const float* g2 = mixed ? static_cast<const float*>(pw.force[2]) : nullptr;
    if (mixed) prepare_mixed_pairs(c, wp, sp, o, uniform, um, s);
    auto* wcnt = walk_cnt;
    // The per-target adaptive walk (targets the group walk hands back, or all
    // of them). In mixed precision it runs in FP64 without PM into FP64 lanes,
    // and adaptive_mixed_finish adds the FP32 PM part and writes FP32 outputs.
    auto per_target = [&](const chui::WalkParams& fw, const uint32_t* pos, const uint32_t* slot) {
      if (!mixed) {
        chui::walk_fp64_adaptive(c->tree, fw, c->src64.as<double4>(), uniform ? 1 : 0, gtab, prev, pos, slot, f0, f1,
                                 f2, out, wcnt, s);
        return;
      }
      const std::size_t slots = n_active;  // output slots are below n_active
      c->mix_fb.reserve(3 * slots * sizeof(double));
      double* sx = c->mix_fb.as<double>();
      double* sy = sx + slots;
      double* sz = sy + slots;
      double* nr = nullptr;
      if (near2) {
        c->mix_near.reserve(n * sizeof(double));
        nr = c->mix_near.as<double>();
      }
      chui::WalkParams sw = fw;
      sw.pm_enabled = 0;
      const chui::WalkOutputs so{sx, sy, sz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nr};
      chui::walk_fp64_adaptive(c->tree, sw, c->src64.as<double4>(), uniform ? 1 : 0, gtab, prev, pos, slot, nullptr,
                               nullptr, nullptr, so, wcnt, s);
      chui::adaptive_mixed_finish(c->tree, fw, c->src32.as<uint4>(), pos, slot, fw.n_targets, fw.identity_targets, sx,
                                  sy, sz, nr, g0, g1, g2, out, s);
    };
    if (group_status == CHUI_GROUP_WALK_ACTIVE) {
      c->leaf_of_pos.reserve(n * sizeof(uint32_t));
      chui::build_leaf_of_pos(c->tree, c->leaf_of_pos.as<uint32_t>(), s);
      c->fb_pos.reserve(n_active * sizeof(uint32_t));
      c->fb_slot.reserve(n_active * sizeof(uint32_t));
      c->bounds.reserve(8 * sizeof(unsigned long long));
      const unsigned long long init[8] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
      CHUI_HIP(hipMemcpyAsync(c->bounds.ptr, init, sizeof(init), hipMemcpyHostToDevice, s));
      auto* bkeys = c->bounds.as<unsigned long long>();
      source_bounds_kernel<<<std::min<unsigned>(chui::grid_for(n, 256), 1024u), 256, 0, s>>>(x, y, z, n, bkeys);
      CHUI_LAUNCH_CHECK("source_bounds_kernel");
      unsigned long long keys[8];
      c->tree.readback.read(keys, bkeys, sizeof(keys), s);
      chui::GroupParams gp{};
      gp.block = static_cast<uint32_t>(o->residual_block_size);
      gp.group_size = ad.group_target_size;
      gp.max_extent = ad.group_max_extent_cutoff_fraction * std::sqrt(wp.rcut2);
      gp.tolg = o->relative_force_tolerance / o->gravitational_constant;
      gp.amax2 = ad.maximum_opening_angle * ad.maximum_opening_angle;
      gp.contain = wp.com_cut ? 0 : 1;
      for (int a = 0; a < 3; ++a) {
        uint64_t lo = keys[a], hi = keys[3 + a];
        auto decode = [](uint64_t k) {
          const uint64_t b = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
          double d;
          std::memcpy(&d, &b, sizeof(d));
          return d;
        };
        gp.lo[a] = decode(lo);
        gp.hi[a] = decode(hi);
      }
      auto* fb_count = reinterpret_cast<uint32_t*>(bkeys + 6);  // zeroed above
      if (mixed)
        chui::walk_mixed_group(c->tree, wp, gp, c->src64.as<double4>(), c->src32.as<uint4>(), uniform ? 1 : 0,
                               c->htab.as<float4>(), prev, c->leaf_of_pos.as<uint32_t>(), target_pos, target_slot,
                               g0, g1, g2, out, c->fb_pos.as<uint32_t>(), c->fb_slot.as<uint32_t>(), fb_count,
                               reinterpret_cast<unsigned long long*>(c->gcounters.ptr), s);
      else
        chui::walk_fp64_group(c->tree, wp, gp, c->src64.as<double4>(), uniform ? 1 : 0, gtab,
                              c->gcoef.as<double4>(), prev,
                              c->leaf_of_pos.as<uint32_t>(), target_pos, target_slot, f0, f1, f2, out,
                              c->fb_pos.as<uint32_t>(), c->fb_slot.as<uint32_t>(), fb_count,
                              reinterpret_cast<unsigned long long*>(c->gcounters.ptr), s);
      uint32_t nfb = 0;
      c->tree.readback.read(&nfb, fb_count, sizeof(nfb), s);
      if (nfb > n_active) throw std::logic_error("TreePM group walk fallback list overflow");
      if (nfb > 0) {
        chui::WalkParams fw = wp;
        fw.n_targets = nfb;
        fw.identity_targets = 0;
        per_target(fw, c->fb_pos.as<uint32_t>(), c->fb_slot.as<uint32_t>());
      }
    } else {
      per_target(wp, target_pos, target_slot);
    }
  } else if (n_active > 0 && n > 0) {
    if (!mixed) {
      // solves without counters or the nearest-pair lane take the
      // box-classified walk with the FP64 polynomial table of H
      wp.box64 = (!wp.count && !wp.target_per_wave && near2 == nullptr) ? 1 : 0;
      if (wp.box64) prepare_h64(c, wp, sp, s);
      chui::walk_fp64(c->tree, wp, c->src64.as<double4>(), uniform ? 1 : 0, target_pos, target_slot,
                      static_cast<const double*>(pw.force[0]), static_cast<const double*>(pw.force[1]),
                      static_cast<const double*>(pw.force[2]), out, walk_cnt, c->h64tab.as<double2>(), s);
    } else {
      prepare_mixed_pairs(c, wp, sp, o, uniform, um, s);
      // The timed walk (no counters, 64 targets per wave) takes the residual
      // table indexed by the float bits of s = r^2 + eps^2, in units where
      // r_cut^2 + eps^2 = 1; the counting and per-wave walks keep the
      // reciprocal square root and the Gaussian table above.
      if (!wp.count && !wp.target_per_wave) {
        const double scale = 1.0 / (wp.rcut2 + wp.eps * wp.eps);
        const double key[3] = {scale, wp.eps, sp.a};
        if (!(c->ltab_key[0] == key[0] && c->ltab_key[1] == key[1] && c->ltab_key[2] == key[2])) {
          int32_t e0 = 0;
          uint32_t cells = 0;
          c->ltab_host = chui::log_residual_table(scale, wp.eps, sp.a, e0, cells);
          c->ltab_cells = cells;
          c->ltab_first_octave = e0;
          if (cells > 0) {
            chui::require_not_capturing(s, "the residual table");
            c->ltab.reserve(c->ltab_host.size() * sizeof(float));
            CHUI_HIP(hipMemcpyAsync(c->ltab.ptr, c->ltab_host.data(), c->ltab_host.size() * sizeof(float),
                                    hipMemcpyHostToDevice, s));
            CHUI_HIP(hipStreamSynchronize(s));
          }
          std::copy(key, key + 3, c->ltab_key);
        }
        if (c->ltab_cells > 0) {
          const double sq = std::sqrt(scale);
          for (int a = 0; a < 3; ++a) wp.tab_scale[a] = static_cast<float>(o->box_size[a] / 4294967296.0 * sq);
          wp.rcut2_tab = static_cast<float>(wp.rcut2 * scale);
          wp.eps2_tab = static_cast<float>(wp.eps * wp.eps * scale);
          wp.force_scale = static_cast<float>(wp.G * scale * (uniform ? um : 1.0));
          wp.log_cells = c->ltab_cells;
          wp.log_base = static_cast<uint32_t>(c->ltab_first_octave + 127) << chui::kLogCellBits;
        }
      }
      chui::walk_mixed(c->tree, wp, c->src32.as<uint4>(), uniform ? 1 : 0, c->htab.as<float4>(),
                       c->ltab.as<float4>(), target_pos, target_slot, static_cast<const float*>(pw.force[0]),
                       static_cast<const float*>(pw.force[1]), static_cast<const float*>(pw.force[2]), out, walk_cnt,
                       s);
    }
  }
  record(c, kEvWalk);
  if (!rec) post_fault_word(c, s);  // a graph launch (or chui_treepm_fault_check) posts it after the replay
  if (wcount || diag) check_fault(c, true);
  if (wcount) {
    unsigned long long wc[chui::kWalkCounters + 1] = {};
    unsigned long long gc[chui::kGroupCounters] = {};
    CHUI_HIP(hipMemcpyAsync(wc, c->counters.ptr, sizeof(wc), hipMemcpyDeviceToHost, s));
    CHUI_HIP(hipMemcpyAsync(gc, c->gcounters.ptr, sizeof(gc), hipMemcpyDeviceToHost, s));
    CHUI_HIP(hipStreamSynchronize(s));
    chui_treepm_walk_counters r{};
    r.struct_size = sizeof(r);
    r.group_walk_status = group_status;
    r.visited_nodes = wc[0] + gc[0];
    r.cutoff_pruned_nodes = wc[1] + gc[1];
    r.direct_pair_evaluations = wc[2] + gc[2];
    r.accepted_internal_multipoles = wc[3] + gc[3];
    r.accepted_leaves = wc[4] + gc[4];
    r.group_walk_groups = gc[5];
    r.group_walk_grouped_targets = gc[6];
    r.group_walk_target_group_size_max = gc[7];
    r.group_walk_fallback_targets = gc[8];
    r.group_walk_fallback_history = gc[9];
    r.group_walk_fallback_geometry = gc[10];
    r.group_walk_candidate_leaves = gc[11];
    r.group_walk_culled_leaves = gc[12];
    r.group_walk_candidate_particles = gc[13];
    r.group_walk_list_entries_max = gc[14];
    *wcount = r;
  }
  if (diag) {
    if (diag->struct_size != sizeof(chui_treepm_diagnostics))
      throw std::invalid_argument("chui_treepm_diagnostics.struct_size does not match this library");
    unsigned long long cnt[chui::kWalkCounterWords] = {};
    CHUI_HIP(hipMemcpyAsync(cnt, c->counters.ptr, sizeof(cnt), hipMemcpyDeviceToHost, s));
    CHUI_HIP(hipStreamSynchronize(s));
    diag->node_count = c->tree.node_count;
    diag->leaf_count = c->tree.leaf_count;
    diag->max_depth = c->tree.max_depth;
    diag->visited_nodes = cnt[0];
    diag->cutoff_pruned_nodes = cnt[1];
    diag->direct_pair_evaluations = cnt[2];
    diag->accepted_internal_multipoles = cnt[3];
    diag->accepted_leaves = cnt[4];
    diag->wave_node_visits = cnt[5];
    diag->wave_leaf_source_slots = cnt[6];
    diag->wave_sum_flushes = cnt[7];
    diag->pair_set_checksum = cnt[chui::kPairChecksumSlot];
    diag->split_scale_comoving = sp.a;
    diag->cutoff_radius_comoving = sp.rcut;
    diag->mesh_spacing_comoving = sp.h;
    diag->tree_root_half_size_comoving = c->tree.root_half;
    for (int a = 0; a < 3; ++a) diag->unwrap_anchor[a] = c->tree.unwrap.anchor[a];
    diag->ms_total = elapsed(c, kEvStart, kEvWalk);
    diag->ms_unwrap_keys_sort = 0.0;
    diag->ms_tree_build = elapsed(c, kEvStart, kEvTree);
    diag->ms_multipoles = elapsed(c, kEvTree, kEvMultipoles);
    diag->ms_pm_deposit = 0.0;
    diag->ms_pm_fft_poisson = elapsed(c, kEvMultipoles, kEvPm);
    diag->ms_pm_interpolate = 0.0;
    diag->ms_walk = elapsed(c, kEvTargets, kEvWalk);
  }
  if (c->profiling) c->pending = true;
  c->events = false;
}

void require_ctx(const Ctx* c) {
  if (!c) throw std::invalid_argument("TreePM context is null");
}

// Only when another device is current: a stream that records a graph under
// a global capture mode admits as few runtime calls as possible.
void set_device(const Ctx* c) {
  int cur = -1;
  CHUI_HIP(hipGetDevice(&cur));
  if (cur != c->device) CHUI_HIP(hipSetDevice(c->device));
}

}  // namespace

// ===========================================================================
extern "C" {

chui_status chui_treepm_default_options(chui_treepm_options* o) {
  return guarded([&] {
    if (!o) throw std::invalid_argument("options pointer is null");
    std::memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(chui_treepm_options);
    o->box_size[0] = o->box_size[1] = o->box_size[2] = 0.0;
    o->scale_factor = 1.0;
    o->gravitational_constant = 1.0;
    o->assignment_scheme = CHUI_ASSIGN_CIC;       // PmSolveOptions default (pm_solver.hpp:231)
    o->window_deconvolution = 0;                  // PmSolveOptions default (pm_solver.hpp:232)
    o->asmth_cells = 1.25;                        // numerics.treepm_asmth_cells
    o->rcut_cells = 6.25;                         // numerics.treepm_rcut_cells
    o->opening_theta = 0.7;                       // TreeGravityOptions default
    o->opening_criterion = CHUI_OPEN_COM_DISTANCE;
    o->multipole_order = 2;
    o->max_leaf_size = 16;
    o->softening_epsilon = 0.0;
    o->relative_force_tolerance = 0.005;
    o->relative_force_acceleration_floor = 1.0e-30;
    o->acceptance_policy = CHUI_ACCEPT_STRICT;
    o->residual_block_size = 64;
    o->precision = CHUI_PRECISION_FP64;
    o->force_epoch_scale_factor = 1.0;            // ForceEvaluationEpoch default
  });
}

chui_status chui_treepm_options_set(chui_treepm_options* o, const char* key_c, const char* value_c) {
  return guarded([&] {
    if (!o || !key_c || !value_c) throw std::invalid_argument("null argument");
    std::string key(key_c), value(value_c);
    for (const char* prefix : {"numerics.", "cosmology.", "units."})
      if (key.rfind(prefix, 0) == 0) key = key.substr(std::strlen(prefix));
    auto num = [&]() {
      char* end = nullptr;
      const double v = std::strtod(value.c_str(), &end);
      if (end == value.c_str()) throw std::invalid_argument("value of " + key + " is not a number: " + value);
      return v;
    };
    auto boolean = [&]() {
      if (value == "true" || value == "1" || value == "yes" || value == "on") return 1;
      if (value == "false" || value == "0" || value == "no" || value == "off") return 0;
      throw std::invalid_argument("value of " + key + " is not a boolean: " + value);
    };
    // "<magnitude> [mpc|kpc]" (config.cpp:300-362). A bare magnitude is in
    // Mpc, the reference's default units.length_unit; decks with another
    // length unit go through the deck reader, which resolves it.
    auto length_mpc = [&]() {
      std::istringstream in(value);
      double mag = 0.0;
      in >> mag;
      if (in.fail() || !std::isfinite(mag))
        throw std::invalid_argument("invalid value '" + value + "': expected finite numeric magnitude");
      std::string unit;
      in >> unit;
      in >> std::ws;
      if (!in.eof()) throw std::invalid_argument("invalid value '" + value + "': unexpected trailing token(s)");
      for (char& ch : unit) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
      if (unit.empty() || unit == "mpc") return mag;
      if (unit == "kpc") return mag * 1.0e-3;
      throw std::invalid_argument("key '" + key + "': unsupported unit in value '" + value + "'");
    };
    if (key == "box_size" || key == "box_size_mpc_comoving") {
      o->box_size[0] = o->box_size[1] = o->box_size[2] = length_mpc();
    } else if (key == "box_size_x") {
      o->box_size[0] = length_mpc();
    } else if (key == "box_size_y") {
      o->box_size[1] = length_mpc();
    } else if (key == "box_size_z") {
      o->box_size[2] = length_mpc();
    } else if (key == "treepm_asmth_cells") {
      o->asmth_cells = num();
    } else if (key == "treepm_rcut_cells") {
      o->rcut_cells = num();
    } else if (key == "treepm_tree_opening_theta") {
      o->opening_theta = num();
    } else if (key == "treepm_tree_opening_criterion") {
      if (value == "geometric" || value == "barnes_hut_geometric") o->opening_criterion = CHUI_OPEN_GEOMETRIC;
      else if (value == "com_distance") o->opening_criterion = CHUI_OPEN_COM_DISTANCE;
      else if (value == "relative" || value == "relative_force_error") o->opening_criterion = CHUI_OPEN_RELATIVE;
      else throw std::invalid_argument("unknown treepm_tree_opening_criterion: " + value);
    } else if (key == "treepm_tree_max_leaf_size") {
      o->max_leaf_size = static_cast<int32_t>(num());
    } else if (key == "treepm_tree_relative_force_tolerance") {
      o->relative_force_tolerance = num();
    } else if (key == "treepm_tree_relative_force_acceleration_floor") {
      o->relative_force_acceleration_floor = num();
    } else if (key == "treepm_assignment_scheme") {
      if (value == "cic") o->assignment_scheme = CHUI_ASSIGN_CIC;
      else if (value == "tsc") o->assignment_scheme = CHUI_ASSIGN_TSC;
      else throw std::invalid_argument("unknown treepm_assignment_scheme: " + value);
    } else if (key == "treepm_enable_window_deconvolution") {
      o->window_deconvolution = boolean();
    } else if (key == "treepm_adaptive_acceptance_enabled") {
      o->acceptance_policy = boolean() ? CHUI_ACCEPT_ADAPTIVE : CHUI_ACCEPT_STRICT;
    } else if (key == "treepm_residual_block_size") {
      o->residual_block_size = static_cast<int32_t>(num());
    } else if (key == "gravity_softening") {
      // comoving length, handed to the force path in Mpc (gravity_runtime.cpp:566-572)
      o->softening_epsilon = length_mpc();
    } else if (key == "gravity_softening_mpc") {
      o->softening_epsilon = num();
    } else if (key == "gravitational_constant") {
      o->gravitational_constant = num();
    } else if (key == "treepm_precision") {
      if (value == "fp64") o->precision = CHUI_PRECISION_FP64;
      else if (value == "mixed") o->precision = CHUI_PRECISION_MIXED;
      else throw std::invalid_argument("unknown treepm_precision: " + value);
    } else if (key == "treepm_pm_reuse_policy") {
      // config.cpp:987-996; gravity_runtime.cpp:564-565 admits stale fields
      // only under stale_within_cadence
      if (value == "exact" || value == "multirate_outer_kick") o->allow_stale_long_range_field = 0;
      else if (value == "stale_within_cadence") o->allow_stale_long_range_field = 1;
      else
        throw std::invalid_argument("key 'numerics.treepm_pm_reuse_policy': invalid value '" + value +
                                    "' (supported: exact, stale_within_cadence, multirate_outer_kick)");
    } else if (key == "treepm_pm_force_gradient") {
      if (value != "spectral") throw chui::UnsupportedError("only the spectral PM gradient is built");
    } else if (key == "treepm_pm_grid" || key == "treepm_pm_grid_nx" || key == "treepm_pm_grid_ny" ||
               key == "treepm_pm_grid_nz") {
      throw std::invalid_argument(key + " is a context (mesh) property: pass it to chui_treepm_create");
    } else {
      throw std::invalid_argument("unknown TreePM option key: " + key);
    }
  });
}

chui_status chui_treepm_options_validate(const chui_treepm_options* o, const int32_t mesh[3]) {
  return guarded([&] {
    if (!mesh) throw std::invalid_argument("mesh pointer is null");
    const int m[3] = {mesh[0], mesh[1], mesh[2]};
    validate_options(o, m);
  });
}

chui_status chui_treepm_create(const int32_t mesh[3], int32_t device, chui_treepm_context** out) {
  return guarded([&] {
    if (!mesh || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    for (int a = 0; a < 3; ++a) {
      if (mesh[a] <= 0) throw std::invalid_argument("numerics.treepm_pm_grid_n{xyz} must all be > 0");
      int m = mesh[a];
      while (m % 3 == 0) m /= 3;
      if (mesh[a] % 2 != 0 || ilog2_exact(m) < 1)
        throw chui::UnsupportedError("PM mesh axes must be even and of the form 2^a 3^b on the GPU path");
    }
    int count = 0;
    CHUI_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw std::invalid_argument("HIP device ordinal out of range");
    auto* c = new chui_treepm_context();
    c->device = device;
    for (int a = 0; a < 3; ++a) {
      c->mesh[a] = mesh[a];
      c->mesh_lg[a] = ilog2_exact(mesh[a]);  // -1: not a power of two (mixed-radix line passes)
    }
    try {
      CHUI_HIP(hipSetDevice(device));
      CHUI_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->own_stream = true;
      chui::noalloc_scratch_init(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

chui_status chui_treepm_destroy(chui_treepm_context* c) {
  return guarded([&] {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev)
      if (e) (void)hipEventDestroy(e);
    if (c->fault_ev) (void)hipEventDestroy(c->fault_ev);
    if (c->fault_host) (void)hipHostFree(c->fault_host);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
  });
}

// ---- ensembles ------------------------------------------------------------------
// Independent periodic boxes, one context (own mesh, tree, stream) per box.
// A batched solve hands the boxes, largest cost first, to host worker threads;
// each worker runs whole solves on its boxes' streams, so small boxes overlap
// on the device instead of queueing behind one another.
struct chui_treepm_ensemble {
  int device = 0;
  int workers = 1;
  std::vector<chui_treepm_context*> boxes;
  // graph capture: box 0's stream records, the other boxes' streams join it
  bool capturing = false;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
};

// A recorded graph of capture-safe solves (and whatever else the caller issued
// on the member contexts' streams while it was recorded). Launches replay it on
// the stream it was recorded on, after the members' prior work, and order the
// members' later work after it.
struct chui_treepm_graph {
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t origin = nullptr;
  std::vector<chui_treepm_context*> members;
  std::vector<hipEvent_t> in_ev;  // one per member stream other than origin
  hipEvent_t out_ev = nullptr;
  ~chui_treepm_graph() {
    for (auto e : in_ev)
      if (e) (void)hipEventDestroy(e);
    if (out_ev) (void)hipEventDestroy(out_ev);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

// Waits for the context's work, reports its faults and checks that a recorded
// solve would allocate nothing: the walk counters and the fault word's pinned
// copy exist, and the context is in capture-safe mode.
void ready_for_capture(Ctx* c) {
  if (!c->capture_safe)
    throw std::invalid_argument("graph capture needs capture-safe solves (chui_treepm_set_capture_safe)");
  if (c->capturing) throw std::logic_error("a graph is already being recorded on this context");
  CHUI_HIP(hipStreamSynchronize(c->stream));
  check_fault(c, true);
  (void)walk_counters(c, c->stream);
  post_fault_word(c, c->stream);
  CHUI_HIP(hipStreamSynchronize(c->stream));
  check_fault(c, true);
}

chui_treepm_graph* finish_capture(hipStream_t origin, int device, std::vector<Ctx*> members) {
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(origin, &g);
  for (auto* c : members) c->capturing = false;
  CHUI_HIP(e);
  auto* gr = new chui_treepm_graph();
  gr->graph = g;
  gr->device = device;
  gr->origin = origin;
  gr->members = std::move(members);
  try {
    CHUI_HIP(hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0));
    for (auto* c : gr->members)
      if (c->stream != origin) {
        hipEvent_t ev = nullptr;
        CHUI_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        gr->in_ev.push_back(ev);
      }
    CHUI_HIP(hipEventCreateWithFlags(&gr->out_ev, hipEventDisableTiming));
  } catch (...) {
    delete gr;
    throw;
  }
  return gr;
}

}  // namespace

chui_status chui_treepm_ensemble_create(int32_t box_count, const int32_t* meshes, int32_t device, int32_t workers,
                                        chui_treepm_ensemble** out) {
  return guarded([&] {
    if (!meshes || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    if (box_count <= 0) throw std::invalid_argument("an ensemble needs at least one box");
    if (workers < 1 || workers > 64) throw std::invalid_argument("ensemble workers must be in [1, 64]");
    auto* e = new chui_treepm_ensemble();
    e->device = device;
    e->workers = workers;
    try {
      for (int32_t b = 0; b < box_count; ++b) {
        chui_treepm_context* c = nullptr;
        const chui_status st = chui_treepm_create(meshes + 3 * b, device, &c);
        if (st != CHUI_OK) {
          const std::string msg = "box " + std::to_string(b) + ": " + g_last_error;
          if (st == CHUI_ERR_UNSUPPORTED) throw chui::UnsupportedError(msg);
          if (st == CHUI_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
          throw std::runtime_error(msg);
        }
        e->boxes.push_back(c);
      }
    } catch (...) {
      for (auto* c : e->boxes) chui_treepm_destroy(c);
      delete e;
      throw;
    }
    *out = e;
  });
}

chui_status chui_treepm_ensemble_destroy(chui_treepm_ensemble* e) {
  return guarded([&] {
    if (!e) return;
    for (auto* c : e->boxes) chui_treepm_destroy(c);
    delete e;
  });
}

chui_status chui_treepm_ensemble_box(chui_treepm_ensemble* e, int32_t index, chui_treepm_context** out) {
  return guarded([&] {
    if (!e || !out) throw std::invalid_argument("null argument");
    if (index < 0 || static_cast<std::size_t>(index) >= e->boxes.size())
      throw std::invalid_argument("ensemble box index out of range");
    *out = e->boxes[static_cast<std::size_t>(index)];
  });
}

chui_status chui_treepm_ensemble_set_capture_safe(chui_treepm_ensemble* e, int32_t enable) {
  return guarded([&] {
    if (!e) throw std::invalid_argument("null argument");
    if (e->capturing) throw std::logic_error("capture-safe mode cannot change while a graph is recorded");
    for (auto* c : e->boxes) c->capture_safe = enable != 0;
  });
}

// Records on box 0's stream; every other box's stream joins the capture by
// waiting on an event of it, so each box's work forms its own branch of the
// graph and the boxes still overlap on the device.
chui_status chui_treepm_ensemble_graph_begin(chui_treepm_ensemble* e) {
  return guarded([&] {
    if (!e) throw std::invalid_argument("null argument");
    if (e->capturing) throw std::logic_error("a graph is already being recorded on this ensemble");
    CHUI_HIP(hipSetDevice(e->device));
    for (auto* c : e->boxes) ready_for_capture(c);
    if (!e->fork_ev) CHUI_HIP(hipEventCreateWithFlags(&e->fork_ev, hipEventDisableTiming));
    while (e->join_ev.size() < e->boxes.size()) {
      hipEvent_t ev = nullptr;
      CHUI_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      e->join_ev.push_back(ev);
    }
    hipStream_t origin = e->boxes[0]->stream;
    CHUI_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
    e->capturing = true;
    for (auto* c : e->boxes) c->capturing = true;
    try {
      CHUI_HIP(hipEventRecord(e->fork_ev, origin));
      for (auto* c : e->boxes)
        if (c->stream != origin) CHUI_HIP(hipStreamWaitEvent(c->stream, e->fork_ev, 0));
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(origin, &g);
      if (g) (void)hipGraphDestroy(g);
      e->capturing = false;
      for (auto* c : e->boxes) c->capturing = false;
      throw;
    }
  });
}

chui_status chui_treepm_ensemble_graph_end(chui_treepm_ensemble* e, chui_treepm_graph** out) {
  return guarded([&] {
    if (!e || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    if (!e->capturing) throw std::logic_error("no graph is being recorded on this ensemble");
    CHUI_HIP(hipSetDevice(e->device));
    hipStream_t origin = e->boxes[0]->stream;
    e->capturing = false;
    std::size_t k = 0;
    hipError_t joined = hipSuccess;
    for (auto* c : e->boxes) {
      if (c->stream == origin) continue;
      if (joined == hipSuccess) joined = hipEventRecord(e->join_ev[k], c->stream);
      if (joined == hipSuccess) joined = hipStreamWaitEvent(origin, e->join_ev[k], 0);
      ++k;
    }
    chui_treepm_graph* g = finish_capture(origin, e->device, e->boxes);
    if (joined != hipSuccess) {
      delete g;
      CHUI_HIP(joined);
    }
    *out = g;
  });
}

chui_status chui_treepm_ensemble_solve_device(chui_treepm_ensemble* e, const chui_treepm_box* boxes,
                                              int32_t box_count, int32_t refresh_long_range) {
  return guarded([&] {
    if (!e || !boxes) throw std::invalid_argument("null argument");
    if (box_count != static_cast<int32_t>(e->boxes.size()))
      throw std::invalid_argument("ensemble solve needs one chui_treepm_box per ensemble box");
    // largest cost first (N log N + M log M): a greedy list schedule over the workers
    std::vector<int> order(static_cast<std::size_t>(box_count));
    std::vector<double> cost(order.size());
    for (int b = 0; b < box_count; ++b) {
      order[b] = b;
      const chui_treepm_context* c = e->boxes[b];
      const double n = static_cast<double>(boxes[b].n) + 1.0;
      const double m = static_cast<double>(c->mesh[0]) * c->mesh[1] * c->mesh[2];
      cost[b] = n * std::log2(n) + m * std::log2(m);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::vector<chui_status> status(order.size(), CHUI_OK);
    std::vector<std::string> message(order.size());
    if (e->capturing) {
      // recording: issue every box from this thread, nothing to wait for
      for (int b : order) {
        const chui_treepm_box& bx = boxes[b];
        if (!bx.options) throw std::invalid_argument("ensemble box options are null");
        solve_device_impl(e->boxes[b], bx.pos_x, bx.pos_y, bx.pos_z, bx.mass, bx.n, nullptr, bx.n, bx.accel_x,
                          bx.accel_y, bx.accel_z, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, bx.options,
                          refresh_long_range, nullptr);
      }
      return;
    }
    auto work = [&]() {
      (void)hipSetDevice(e->device);
      for (;;) {
        const int k = next.fetch_add(1);
        if (k >= box_count || failed.load()) return;
        const int b = order[k];
        const chui_treepm_box& bx = boxes[b];
        chui_treepm_context* c = e->boxes[b];
        status[b] = guarded([&] {
          if (!bx.options) throw std::invalid_argument("ensemble box options are null");
          solve_device_impl(c, bx.pos_x, bx.pos_y, bx.pos_z, bx.mass, bx.n, nullptr, bx.n, bx.accel_x, bx.accel_y,
                            bx.accel_z, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, bx.options,
                            refresh_long_range, nullptr);
          CHUI_HIP(hipStreamSynchronize(c->stream));
          check_fault(c, true);
        });
        if (status[b] != CHUI_OK) {
          message[b] = g_last_error;
          failed.store(true);
        }
      }
    };
    const int nw = std::min<int>(e->workers, box_count);
    std::vector<std::thread> pool;
    for (int w = 1; w < nw; ++w) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (int b = 0; b < box_count; ++b) {
      if (status[b] == CHUI_OK) continue;
      const std::string msg = "ensemble box " + std::to_string(b) + ": " + message[b];
      switch (status[b]) {
        case CHUI_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case CHUI_ERR_UNSUPPORTED: throw chui::UnsupportedError(msg);
        case CHUI_ERR_OVERFLOW: throw std::overflow_error(msg);
        case CHUI_ERR_OUT_OF_MEMORY: throw chui::OutOfMemoryError(msg);
        case CHUI_ERR_DEVICE: throw chui::DeviceError(msg);
        case CHUI_ERR_LOGIC: throw std::logic_error(msg);
        default: throw std::runtime_error(msg);
      }
    }
  });
}

chui_status chui_treepm_set_stream(chui_treepm_context* c, void* stream) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    CHUI_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream && c->stream) CHUI_HIP(hipStreamDestroy(c->stream));
    if (stream) {
      c->stream = static_cast<hipStream_t>(stream);
      c->own_stream = false;
    } else {
      CHUI_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->own_stream = true;
    }
  });
}

chui_status chui_treepm_synchronize(chui_treepm_context* c) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    CHUI_HIP(hipStreamSynchronize(c->stream));
    check_fault(c, true);
  });
}

chui_status chui_treepm_set_launch_shape(chui_treepm_context* c, int32_t fft_lines_per_block,
                                        int32_t walk_waves_per_block) {
  return guarded([&] {
    require_ctx(c);
    if (c->capturing) throw std::logic_error("launch shapes cannot change while a graph is recorded");
    const int fl = fft_lines_per_block, ww = walk_waves_per_block;
    if (!(fl == 0 || fl == 1 || fl == 2 || fl == 4 || fl == 8 || fl == 16))
      throw std::invalid_argument("fft_lines_per_block must be 0 or a power of two up to 16");
    if (!(ww == 0 || ww == 2 || ww == 4 || ww == 8))
      throw std::invalid_argument("walk_waves_per_block must be 0, 2, 4 or 8");
    c->fft_lines = fl;
    c->walk_waves = ww;
  });
}

chui_status chui_treepm_fault_check(chui_treepm_context* c) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (recording(c)) throw std::logic_error("fault check while the context's stream records a graph");
    if (!c->counters.ptr) return;
    post_fault_word(c, c->stream);
    check_fault(c, true);
  });
}

// ---- capture-safe solves and HIP graphs ------------------------------------
chui_status chui_treepm_set_capture_safe(chui_treepm_context* c, int32_t enable) {
  return guarded([&] {
    require_ctx(c);
    if (c->capturing) throw std::logic_error("capture-safe mode cannot change while a graph is recorded");
    c->capture_safe = enable != 0;
  });
}


chui_status chui_treepm_graph_begin(chui_treepm_context* c) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    ready_for_capture(c);
    CHUI_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
  });
}

chui_status chui_treepm_graph_end(chui_treepm_context* c, chui_treepm_graph** out) {
  return guarded([&] {
    require_ctx(c);
    if (!out) throw std::invalid_argument("null argument");
    *out = nullptr;
    if (!c->capturing) throw std::logic_error("no graph is being recorded on this context");
    set_device(c);
    *out = finish_capture(c->stream, c->device, {c});
  });
}

chui_status chui_treepm_graph_launch(chui_treepm_graph* g) {
  return guarded([&] {
    if (!g || !g->exec) throw std::invalid_argument("null graph");
    CHUI_HIP(hipSetDevice(g->device));
    std::size_t k = 0;
    for (auto* c : g->members) {
      if (c->capturing) throw std::logic_error("graph launch while a member context records a graph");
      if (c->stream == g->origin) continue;
      CHUI_HIP(hipEventRecord(g->in_ev[k], c->stream));
      CHUI_HIP(hipStreamWaitEvent(g->origin, g->in_ev[k], 0));
      ++k;
    }
    CHUI_HIP(hipGraphLaunch(g->exec, g->origin));
    // each member's fault word after the replay, checked at its next synchronisation
    for (auto* c : g->members) post_fault_word(c, g->origin);
    CHUI_HIP(hipEventRecord(g->out_ev, g->origin));
    for (auto* c : g->members)
      if (c->stream != g->origin) CHUI_HIP(hipStreamWaitEvent(c->stream, g->out_ev, 0));
  });
}

chui_status chui_treepm_graph_destroy(chui_treepm_graph* g) {
  return guarded([&] {
    if (!g) return;
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->origin);
    delete g;
  });
}

// Transpose of the Jacobian of the total acceleration with respect to the
// source positions, applied to lam: out_k = sum_i (d a_i / d x_k)^T lam_i over
// all N sources (all active). FP64 only, for the strict policy with monopole
// (multipole_order 0: no node is accepted, tree_pm_acceptance.hpp:62-66, and
// the short-range force is the exact pair sum within r_cut) or quadrupole
// moments (multipole_order 2: the walk takes the forward's acceptance
// decisions, and each accepted node's expansion, tree_pm_coupling.cpp:743-827,
// is differentiated with respect to the target and to the node's particles
// through its centre of mass and quadrupole). The tree and both PM fields are
// rebuilt at x.
void force_vjp_impl(Ctx* c, const double* x, const double* y, const double* z, const double* m, uint64_t n,
                    const double* lx, const double* ly, const double* lz, double* ox, double* oy, double* oz,
                    const chui_treepm_options* o) {
  validate_options(o, c->mesh);
  if (o->precision != CHUI_PRECISION_FP64)
    throw chui::UnsupportedError("the force VJP is built for FP64 precision only");
  if (o->acceptance_policy != CHUI_ACCEPT_STRICT)
    throw chui::UnsupportedError("the force VJP differentiates the strict acceptance policy only");
  const bool quad = o->multipole_order == 2;
  if (n > 0xFFFFFFFEull) throw std::overflow_error("Tree gravity source count exceeds the 32-bit tree-index contract");
  if (n > 0 && (!x || !y || !z || !m || !lx || !ly || !lz || !ox || !oy || !oz))
    throw std::invalid_argument("TreePM force VJP lanes are null");
  if (n == 0) return;
  hipStream_t s = c->stream;
  // capture-safe mode as in solve_device_impl; a recorded VJP also takes its
  // accepted-node count from the device (the list sorted over the capacity an
  // eager VJP learned)
  const bool safe_mode = c->capture_safe;
  const bool safe = safe_mode && c->tree.prepared_n == n && c->tree.sized_for_capture;
  const bool rec = recording(c);
  if (rec && (!safe || (quad && c->vjp_safe_cap == 0)))
    throw std::runtime_error(
        "graph capture of a force VJP on a context not prepared for these sources: run one capture-safe VJP of "
        "them before capturing");
  if (!rec) {
    harvest(c);
    check_fault(c, false);
  }
  unsigned long long* const walk_cnt = walk_counters(c, s);
  const Split sp = split_of(*o, c->mesh);
  chui::TreeParams tp;
  tp.n = n;
  for (int a = 0; a < 3; ++a) tp.L[a] = o->box_size[a];
  tp.leaf_size = static_cast<uint32_t>(o->max_leaf_size);
  tp.quadrupole = quad;
  tp.mixed = false;
  tp.capture_safe = safe;
  tp.size_for_capture = safe_mode && !safe;
  tp.fault = walk_cnt + chui::kFaultSlot;
  chui::build_tree(c->tree, x, y, z, m, tp, o->softening_epsilon, s);
  const bool uniform = c->tree.mass_uniform;
  const double um = uniform ? c->tree.uniform_mass : 0.0;
  c->src64.reserve(n * sizeof(double4));
  gather64_kernel<<<chui::grid_for(n, 256), 256, 0, s>>>(c->tree.order.as<uint32_t>(), n, x, y, z, m,
                                                         c->src64.as<double4>());
  CHUI_LAUNCH_CHECK("gather64_kernel");
  c->vjp_lam.reserve(n * sizeof(double4));
  for (auto& b : c->vjp_comp) b.reserve(n * sizeof(double4));
  gather_lam_kernel<<<chui::grid_for(n, 256), 256, 0, s>>>(
      c->tree.order.as<uint32_t>(), n, c->src64.as<double4>(), lx, ly, lz, c->vjp_lam.as<double4>(),
      c->vjp_comp[0].as<double4>(), c->vjp_comp[1].as<double4>(), c->vjp_comp[2].as<double4>());
  CHUI_LAUNCH_CHECK("gather_lam_kernel");

  chui::PmGeometry g{};
  for (int a = 0; a < 3; ++a) {
    g.n[a] = c->mesh[a];
    g.lg[a] = c->mesh_lg[a];
    g.L[a] = o->box_size[a];
  }
  g.nzc = static_cast<std::size_t>(c->mesh[2] / 2 + 1);
  g.row = 2 * g.nzc;
  chui::PmOptions po;
  po.G = o->gravitational_constant;
  po.split_scale = sp.a;
  po.assignment = o->assignment_scheme;
  po.deconvolution = o->window_deconvolution;
  po.fft_lines = c->fft_lines;
  chui::DepositTiling dt{};
  for (int a = 0; a < 3; ++a) {
    dt.n[a] = g.n[a];
    dt.inv_h[a] = static_cast<double>(g.n[a]) / g.L[a];
  }
  dt.scheme = po.assignment;
  dt.max_cells = chui::kDepositTileBytes / sizeof(double);
  dt.chunk = chui::kDepositChunk;
  const chui::DepositUnits units = chui::build_deposit_units(c->tree, dt, s);
  chui::pm_solve_fp64(c->vjp_fwd, g, po, c->src64.as<double4>(), uniform ? 1 : 0, um, n, units, s);
  c->vjp_fwd.valid = false;  // a field at the VJP's positions, never reused by a solve
  const double4* comp[3] = {c->vjp_comp[0].as<double4>(), c->vjp_comp[1].as<double4>(),
                            c->vjp_comp[2].as<double4>()};
  chui::pm_adjoint_fp64(c->vjp_adj, g, po, comp, n, units, s);

  chui::WalkParams wp = base_walk_params(c, o, sp, g.row);
  wp.quad = quad ? 1 : 0;
  wp.n_targets = static_cast<uint32_t>(n);
  wp.identity_targets = 1;
  wp.count = 0;
  wp.uniform_mass = um;
  wp.uniform_mass_f = static_cast<float>(um);
  wp.pm_enabled = 1;
  chui::WalkOutputs out{ox, oy, oz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  wp.box64 = 1;
  prepare_h64(c, wp, sp, s);
  // accepted-node terms: a list sized 4 per source, or the last VJP's count
  // if larger; a walk that accepts more runs again with the list it needs
  chui::VjpNodeTerms nt;
  std::size_t cap = rec ? c->vjp_safe_cap : std::max<std::size_t>(4 * n + 1024, c->vjp_accepted_cap);
  for (int attempt = 0;; ++attempt) {
    if (quad) {
      const std::size_t nodes = c->tree.node_count;  // the capacity after a capture-safe build
      const std::size_t nodes_sz =
          c->tree.sized_for_capture ? std::max(nodes, c->tree.bfs_capacity) : nodes;
      cap = std::min<std::size_t>(cap, 0xFFFFFFFFull);
      c->vjp_nacc.reserve(std::max<std::size_t>(nodes_sz, 1) * 10 * sizeof(double));
      c->vjp_rec.reserve(cap * sizeof(uint64_t));
      c->vjp_rec_alt.reserve(cap * sizeof(uint64_t));
      c->vjp_vals.reserve(cap * sizeof(uint32_t));
      c->vjp_vals_alt.reserve(cap * sizeof(uint32_t));
      c->vjp_sort.reserve(chui::radix_sort_scratch_bytes(cap));
      c->vjp_count.reserve(sizeof(uint32_t));
      c->vjp_corr.reserve(n * 3 * sizeof(double));
      chui::fill_f64(c->vjp_nacc.as<double>(), 0.0, nodes * 10, s);
      chui::fill_u32(c->vjp_count.as<uint32_t>(), 0u, 1, s);
      chui::fill_f64(c->vjp_corr.as<double>(), 0.0, n * 3, s);
      if (rec) chui::fill_u64(c->vjp_rec.as<uint64_t>(), ~0ull, cap, s);  // unused slots sort last
      nt.node_acc = c->vjp_nacc.as<double>();
      nt.accepted = c->vjp_rec.as<unsigned long long>();
      nt.accepted_alt = c->vjp_rec_alt.as<unsigned long long>();
      nt.vals = c->vjp_vals.as<uint32_t>();
      nt.vals_alt = c->vjp_vals_alt.as<uint32_t>();
      nt.sort_scratch = c->vjp_sort.ptr;
      nt.count = c->vjp_count.as<unsigned int>();
      nt.cap = static_cast<unsigned int>(cap);
      nt.corr = c->vjp_corr.as<double>();
    }
    chui::walk_fp64_vjp(c->tree, wp, c->src64.as<double4>(), uniform ? 1 : 0, c->vjp_lam.as<double4>(),
                        static_cast<const double*>(c->vjp_fwd.force[0]),
                        static_cast<const double*>(c->vjp_fwd.force[1]),
                        static_cast<const double*>(c->vjp_fwd.force[2]), c->vjp_adj.buf[0].as<double>(), out,
                        walk_cnt, c->h64tab.as<double2>(), quad ? &nt : nullptr, s);
    if (!quad) break;
    if (rec) {
      chui::vjp_node_corrections(c->tree, wp, c->src64.as<double4>(), uniform ? 1 : 0, um, c->vjp_lam.as<double4>(),
                                 nt, static_cast<unsigned int>(cap), out, s, nt.count, walk_cnt + chui::kFaultSlot);
      break;
    }
    uint32_t count = 0;
    CHUI_HIP(hipMemcpyAsync(&count, nt.count, sizeof(count), hipMemcpyDeviceToHost, s));
    CHUI_HIP(hipStreamSynchronize(s));

    if (count > nt.cap) {
      if (attempt > 0) throw std::logic_error("force VJP: accepted-node count changed between identical walks");
      cap = static_cast<std::size_t>(count) + n;
      c->vjp_accepted_cap = cap;
      continue;
    }
    chui::vjp_node_corrections(c->tree, wp, c->src64.as<double4>(), uniform ? 1 : 0, um, c->vjp_lam.as<double4>(), nt,
                               count, out, s);
    if (safe_mode) {
      // the list a recorded VJP sorts: twice this count (reserved after the
      // corrections above have read this one's records)
      const std::size_t want = std::min<std::size_t>(2 * static_cast<std::size_t>(count) + 1024, 0xFFFFFFFFull);
      if (want > c->vjp_safe_cap) {
        CHUI_HIP(hipStreamSynchronize(s));
        c->vjp_safe_cap = want;
        c->vjp_rec.reserve(want * sizeof(uint64_t));
        c->vjp_rec_alt.reserve(want * sizeof(uint64_t));
        c->vjp_vals.reserve(want * sizeof(uint32_t));
        c->vjp_vals_alt.reserve(want * sizeof(uint32_t));
        c->vjp_sort.reserve(chui::radix_sort_scratch_bytes(want));
      }
    }
    break;
  }
  if (!rec) post_fault_word(c, s);
}

chui_status chui_treepm_force_vjp_device(chui_treepm_context* c, const double* x, const double* y, const double* z,
                                         const double* m, uint64_t n, const double* lam_x, const double* lam_y,
                                         const double* lam_z, double* out_x, double* out_y, double* out_z,
                                         const chui_treepm_options* o) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    force_vjp_impl(c, x, y, z, m, n, lam_x, lam_y, lam_z, out_x, out_y, out_z, o);
  });
}

chui_status chui_treepm_solve_device(chui_treepm_context* c, const double* x, const double* y, const double* z,
                                     const double* m, uint64_t n, const uint32_t* active, uint64_t n_active,
                                     void* ax, void* ay, void* az, void* pmx, void* pmy, void* pmz, void* srx,
                                     void* sry, void* srz, const chui_treepm_options* o, int32_t refresh,
                                     chui_treepm_diagnostics* diag) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    solve_device_impl(c, x, y, z, m, n, active, n_active, ax, ay, az, pmx, pmy, pmz, srx, sry, srz, o, refresh,
                      diag);
  });
}

chui_status chui_treepm_default_adaptive_options(chui_treepm_adaptive_options* a) {
  return guarded([&] {
    if (!a) throw std::invalid_argument("TreePM adaptive options pointer is null");
    *a = adaptive_defaults();
  });
}

chui_status chui_treepm_adaptive_options_set(chui_treepm_adaptive_options* a, const char* key_c,
                                             const char* value_c) {
  return guarded([&] {
    if (!a || !key_c || !value_c) throw std::invalid_argument("null argument");
    std::string key(key_c), value(value_c);
    if (key.rfind("numerics.", 0) == 0) key = key.substr(9);
    auto num = [&]() {
      char* end = nullptr;
      const double v = std::strtod(value.c_str(), &end);
      if (end == value.c_str()) throw std::invalid_argument("value of " + key + " is not a number: " + value);
      return v;
    };
    auto boolean = [&]() {
      if (value == "true" || value == "1" || value == "yes" || value == "on") return 1;
      if (value == "false" || value == "0" || value == "no" || value == "off") return 0;
      throw std::invalid_argument("value of " + key + " is not a boolean: " + value);
    };
    // config.cpp:1068-1094 (estimator and cutoff policy names), :1741-1754 (keys)
    if (key == "treepm_adaptive_maximum_opening_angle") {
      a->maximum_opening_angle = num();
    } else if (key == "treepm_adaptive_error_estimator") {
      if (value == "residual_derivative") a->error_estimator = CHUI_ESTIMATOR_RESIDUAL_DERIVATIVE;
      else if (value == "relative_monopole") a->error_estimator = CHUI_ESTIMATOR_RELATIVE_MONOPOLE;
      else throw std::invalid_argument("key 'numerics.treepm_adaptive_error_estimator': invalid value '" + value +
                                       "' (supported: residual_derivative, relative_monopole)");
    } else if (key == "treepm_cutoff_node_policy") {
      if (value == "contain") a->cutoff_node_policy = CHUI_CUTOFF_CONTAIN_NODE;
      else if (value == "com_distance") a->cutoff_node_policy = CHUI_CUTOFF_COM_DISTANCE;
      else throw std::invalid_argument("key 'numerics.treepm_cutoff_node_policy': invalid value '" + value +
                                       "' (supported: contain, com_distance)");
    } else if (key == "treepm_group_walk_enabled") {
      a->group_walk = boolean();
    } else if (key == "treepm_group_walk_target_size") {
      const double v = num();
      if (!(v >= 0.0 && v <= 256.0) || v == 1.0 || v != std::floor(v))
        throw std::invalid_argument(
            "numerics.treepm_group_walk_target_size must be 0 (one source leaf) or in [2, treepm_residual_block_size]");
      a->group_target_size = static_cast<uint32_t>(v);
    } else if (key == "treepm_gaussian_pair_lookup_enabled") {
      a->gaussian_pair_lookup = boolean();
    } else if (key == "treepm_radial_coefficients_enabled") {
      a->radial_coefficients = boolean();
    } else {
      throw std::invalid_argument("unknown TreePM adaptive option key: " + key);
    }
  });
}

chui_status chui_treepm_solve_adaptive_device(chui_treepm_context* c, const double* x, const double* y,
                                              const double* z, const double* m, uint64_t n, const uint32_t* active,
                                              uint64_t n_active, void* ax, void* ay, void* az, void* pmx,
                                              void* pmy, void* pmz, void* srx, void* sry, void* srz,
                                              void* nearest_pair_r2, const double* previous_acceleration,
                                              const chui_treepm_options* o,
                                              const chui_treepm_adaptive_options* adaptive, int32_t refresh,
                                              chui_treepm_walk_counters* counters) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (!adaptive) throw std::invalid_argument("TreePM adaptive options pointer is null");
    solve_device_impl(c, x, y, z, m, n, active, n_active, ax, ay, az, pmx, pmy, pmz, srx, sry, srz, o, refresh,
                      nullptr, previous_acceleration, adaptive, counters, nearest_pair_r2);
  });
}

chui_status chui_treepm_solve(chui_treepm_context* c, const double* x, const double* y, const double* z,
                              const double* m, uint64_t n, const uint32_t* active, uint64_t n_active, double* ax,
                              double* ay, double* az, double* pmx, double* pmy, double* pmz, double* srx,
                              double* sry, double* srz, const chui_treepm_options* o, int32_t refresh,
                              chui_treepm_diagnostics* diag) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    validate_options(o, c->mesh);
    const uint64_t na = active ? n_active : n;
    hipStream_t s = c->stream;
    const std::size_t nb = std::max<uint64_t>(n, 1) * sizeof(double);
    c->hx.reserve(nb);
    c->hy.reserve(nb);
    c->hz.reserve(nb);
    c->hm.reserve(nb);
    if (n > 0) {
      if (!x || !y || !z || !m) throw std::invalid_argument("TreePM source lanes are null");
      CHUI_HIP(hipMemcpyAsync(c->hx.ptr, x, n * 8, hipMemcpyHostToDevice, s));
      CHUI_HIP(hipMemcpyAsync(c->hy.ptr, y, n * 8, hipMemcpyHostToDevice, s));
      CHUI_HIP(hipMemcpyAsync(c->hz.ptr, z, n * 8, hipMemcpyHostToDevice, s));
      CHUI_HIP(hipMemcpyAsync(c->hm.ptr, m, n * 8, hipMemcpyHostToDevice, s));
    }
    if (active && na > 0) {
      c->hactive.reserve(na * 4);
      CHUI_HIP(hipMemcpyAsync(c->hactive.ptr, active, na * 4, hipMemcpyHostToDevice, s));
    }
    const bool mixed = o->precision == CHUI_PRECISION_MIXED;
    const std::size_t eb = mixed ? 4 : 8;
    // lanes 0-5 per active slot, 6-8 (short range) per source row
    for (int i = 0; i < 9; ++i) c->hout[i].reserve(std::max<uint64_t>(i < 6 ? na : n, 1) * eb);
    double* hosts[9] = {ax, ay, az, pmx, pmy, pmz, srx, sry, srz};
    void* dev[9];
    for (int i = 0; i < 9; ++i) dev[i] = hosts[i] ? c->hout[i].ptr : nullptr;
    solve_device_impl(c, c->hx.as<double>(), c->hy.as<double>(), c->hz.as<double>(), c->hm.as<double>(), n,
                      active ? c->hactive.as<uint32_t>() : nullptr, na, dev[0], dev[1], dev[2], dev[3], dev[4],
                      dev[5], dev[6], dev[7], dev[8], o, refresh, diag);
    std::vector<double> tmp64;
    for (int i = 0; i < 9; ++i) {
      if (!hosts[i] || !na) continue;
      const uint64_t len = i < 6 ? na : n;
      if (mixed) c->host_float.resize(len);
      else tmp64.resize(len);
      CHUI_HIP(hipMemcpyAsync(mixed ? static_cast<void*>(c->host_float.data()) : static_cast<void*>(tmp64.data()),
                              dev[i], len * eb, hipMemcpyDeviceToHost, s));
      CHUI_HIP(hipStreamSynchronize(s));
      auto value = [&](uint64_t k) { return mixed ? static_cast<double>(c->host_float[k]) : tmp64[k]; };
      if (i < 6 || !active) {
        for (uint64_t k = 0; k < len; ++k) hosts[i][k] = value(k);
      } else {  // only the active rows of a source-row lane are written
        for (uint64_t k = 0; k < na; ++k) hosts[i][active[k]] = value(active[k]);
      }
    }
    check_fault(c, true);
  });
}

chui_status chui_treepm_profiling(chui_treepm_context* c, int32_t enable) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    CHUI_HIP(hipStreamSynchronize(c->stream));
    c->pending = false;
    c->profiling = enable != 0;
    std::fill(c->prof_ms, c->prof_ms + 8, 0.0);
    c->prof_solves = 0;
  });
}

chui_status chui_treepm_profile_read(chui_treepm_context* c, double stage_ms[8], uint64_t* solves) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    CHUI_HIP(hipStreamSynchronize(c->stream));
    harvest(c);
    check_fault(c, true);
    if (stage_ms) std::copy(c->prof_ms, c->prof_ms + 8, stage_ms);
    if (solves) *solves = c->prof_solves;
  });
}

chui_status chui_treepm_kdk_factors(double a0, double a1, double H0, double om, double ol, chui_kdk_factors* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null output");
    if (!(a0 > 0.0) || !(a1 > 0.0)) throw std::invalid_argument("scale factors must be positive");
    if (a1 < a0) throw std::invalid_argument("scale_factor_end must be >= scale_factor_begin");
    if (!(H0 > 0.0)) throw std::invalid_argument("H0 must be positive");
    const double ok = 1.0 - om - ol;
    auto H = [&](double a) { return H0 * std::sqrt(om / (a * a * a) + ok / (a * a) + ol); };
    // cosmic time t(a) from a0 (composite Simpson, 4096 panels)
    auto cosmic = [&](double lo, double hi) {
      const int np = 4096;
      const double hstep = (hi - lo) / np;
      if (hstep == 0.0) return 0.0;
      double acc = 0.0;
      for (int i = 0; i <= np; ++i) {
        const double a = lo + i * hstep;
        const double f = 1.0 / (a * H(a));
        acc += (i == 0 || i == np) ? f : (i % 2 ? 4.0 * f : 2.0 * f);
      }
      return acc * hstep / 3.0;
    };
    const double t_total = cosmic(a0, a1);
    double lo = a0, hi = a1;
    for (int it = 0; it < 200; ++it) {  // scale factor at half the cosmic time
      const double mid = 0.5 * (lo + hi);
      if (cosmic(a0, mid) < 0.5 * t_total) lo = mid;
      else hi = mid;
    }
    const double am = 0.5 * (lo + hi);
    auto driftlike = [&](double b, double e) {  // time_integration.cpp:242-266, 64 midpoint samples
      const int samples = 64;
      const double da = (e - b) / samples;
      if (da == 0.0) return 0.0;
      double acc = 0.0;
      for (int i = 0; i < samples; ++i) {
        const double a = b + (i + 0.5) * da;
        acc += 1.0 / (a * a * H(a));
      }
      return acc * da;
    };
    out->scale_factor_begin = a0;
    out->scale_factor_midpoint = am;
    out->scale_factor_end = a1;
    out->drift_factor = driftlike(a0, a1);
    out->first_kick_factor = driftlike(a0, am) / am;
    out->second_kick_factor = driftlike(am, a1) / a1;
    out->first_hubble_drag = a0 / am;
    out->second_hubble_drag = am / a1;
  });
}

chui_status chui_treepm_kick_device(chui_treepm_context* c, double* vx, double* vy, double* vz, const void* ax,
                                    const void* ay, const void* az, int32_t is_float, const uint32_t* active,
                                    uint64_t n, double kick, double drag) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    chui::kick_device(vx, vy, vz, ax, ay, az, is_float, active, n, kick, drag, c->stream);
  });
}

chui_status chui_treepm_drift_device(chui_treepm_context* c, double* x, double* y, double* z, const double* vx,
                                     const double* vy, const double* vz, uint64_t n, double d, const double L[3]) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (!L) throw std::invalid_argument("box size pointer is null");
    chui::drift_device(x, y, z, vx, vy, vz, n, d, L, c->stream);
  });
}

chui_status chui_treepm_cic_density_device(chui_treepm_context* c, const double* x, const double* y, const double* z,
                                          uint64_t n, double box, int32_t mesh, double* delta) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (!(box > 0.0) || mesh <= 0) throw std::invalid_argument("CIC density needs box > 0 and mesh > 0");
    if (!delta || (n > 0 && (!x || !y || !z))) throw std::invalid_argument("CIC density lanes are null");
    chui::cic_density_device(x, y, z, n, box, mesh, delta, c->stream);
  });
}

chui_status chui_treepm_cic_density_vjp_device(chui_treepm_context* c, const double* x, const double* y,
                                              const double* z, uint64_t n, double box, int32_t mesh,
                                              const double* grad_delta, double* out_x, double* out_y,
                                              double* out_z) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (!(box > 0.0) || mesh <= 0) throw std::invalid_argument("CIC density needs box > 0 and mesh > 0");
    if (n > 0 && (!x || !y || !z || !grad_delta || !out_x || !out_y || !out_z))
      throw std::invalid_argument("CIC density VJP lanes are null");
    chui::cic_density_vjp_device(x, y, z, n, box, mesh, grad_delta, out_x, out_y, out_z, c->stream);
  });
}

chui_status chui_treepm_mass_momentum_device(chui_treepm_context* c, const double* m, const double* vx,
                                             const double* vy, const double* vz, uint64_t n, double out[4]) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    c->red_scratch.reserve((4 * 1024 + 8) * sizeof(double));
    chui::mass_momentum_device(m, vx, vy, vz, n, c->red_scratch.as<double>(), out, c->stream);
  });
}

chui_status chui_treepm_mass_momentum_noalloc_device(chui_treepm_context* c, const double* m, const double* vx,
                                                     const double* vy, const double* vz, uint64_t n,
                                                     double* out_device) {
  return guarded([&] {
    require_ctx(c);
    if (!out_device) throw std::invalid_argument("reduction output pointer is null");
    if (n > 0 && (!m || !vx || !vy || !vz)) throw std::invalid_argument("reduction input lanes are null");
    set_device(c);
    chui::mass_momentum_noalloc_device(c->device, m, vx, vy, vz, n, out_device, c->stream);
  });
}

chui_status chui_treepm_unwrap_device(chui_treepm_context* c, const double* in, uint64_t n, double L, double* out,
                                      double* anchor) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    const double a = chui::unwrap_axis_device(c->tree, in, n, L, out, c->stream);
    CHUI_HIP(hipStreamSynchronize(c->stream));
    if (anchor) *anchor = a;
  });
}

chui_status chui_treepm_morton_device(chui_treepm_context* c, const double* x, const double* y, const double* z,
                                      uint64_t n, uint64_t* keys, uint32_t* order) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (n > 0xFFFFFFFEull) throw std::overflow_error("Morton ordering exceeds the configured local tree-index contract");
    chui::morton_order_device(c->tree, x, y, z, n, keys, order, c->stream);
    CHUI_HIP(hipStreamSynchronize(c->stream));
  });
}

chui_status chui_treepm_tree_size(chui_treepm_context* c, uint64_t* nodes, uint64_t* particles) {
  return guarded([&] {
    require_ctx(c);
    if (c->tree.counts_on_device)
      throw std::logic_error("the last tree was built by a capture-safe solve; its node count stays on the device");
    if (nodes) *nodes = c->tree.node_count;
    if (particles) *particles = c->tree.n;
  });
}

chui_status chui_treepm_tree_copy(chui_treepm_context* c, double* geometry, double* moments, uint32_t* pbegin,
                                  uint32_t* pcount, uint32_t* child_index, uint32_t* order) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (c->tree.counts_on_device)
      throw std::logic_error("the last tree was built by a capture-safe solve; its node count stays on the device");
    const std::size_t nn = c->tree.node_count;
    hipStream_t s = c->stream;
    if (nn > 0) {
      chui::ensure_records64(c->tree, s);
      std::vector<chui::NodeLink> link(nn);
      std::vector<chui::NodeGeo64> geo(nn);
      std::vector<chui::NodeMom64> mom(nn);
      std::vector<chui::NodeQuad64> quad(nn);
      CHUI_HIP(hipMemcpyAsync(link.data(), c->tree.link.ptr, nn * sizeof(chui::NodeLink), hipMemcpyDeviceToHost, s));
      CHUI_HIP(hipMemcpyAsync(geo.data(), c->tree.geo64.ptr, nn * sizeof(chui::NodeGeo64), hipMemcpyDeviceToHost, s));
      CHUI_HIP(hipMemcpyAsync(mom.data(), c->tree.mom64.ptr, nn * sizeof(chui::NodeMom64), hipMemcpyDeviceToHost, s));
      CHUI_HIP(hipMemcpyAsync(quad.data(), c->tree.quad64.ptr, nn * sizeof(chui::NodeQuad64), hipMemcpyDeviceToHost, s));
      if (child_index)
        CHUI_HIP(hipMemcpyAsync(child_index, c->tree.child_index.ptr, nn * 8 * 4, hipMemcpyDeviceToHost, s));
      CHUI_HIP(hipStreamSynchronize(s));
      for (std::size_t i = 0; i < nn; ++i) {
        if (geometry) {
          geometry[4 * i] = geo[i].cx;
          geometry[4 * i + 1] = geo[i].cy;
          geometry[4 * i + 2] = geo[i].cz;
          geometry[4 * i + 3] = geo[i].half;
        }
        if (moments) {
          double* q = moments + 11 * i;
          q[0] = mom[i].mass; q[1] = mom[i].comx; q[2] = mom[i].comy; q[3] = mom[i].comz;
          q[4] = quad[i].qxx; q[5] = quad[i].qxy; q[6] = quad[i].qxz; q[7] = quad[i].qyy; q[8] = quad[i].qyz;
          q[9] = quad[i].qzz; q[10] = quad[i].trace;
        }
        if (pbegin) pbegin[i] = link[i].begin;
        if (pcount) pcount[i] = link[i].count;
      }
    }
    if (order && c->tree.n > 0) {
      CHUI_HIP(hipMemcpyAsync(order, c->tree.order.ptr, c->tree.n * 4, hipMemcpyDeviceToHost, s));
      CHUI_HIP(hipStreamSynchronize(s));
    }
  });
}

chui_status chui_treepm_ewald_device(chui_treepm_context* c, const double* x, const double* y, const double* z,
                                     const double* m, uint64_t n, const uint32_t* targets, uint64_t nt,
                                     const double L[3], double G, double alpha, int32_t real_images, int32_t modes,
                                     double eps, double* ax, double* ay, double* az) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    if (!L) throw std::invalid_argument("box size pointer is null");
    if (real_images < 0 || modes < 0 || !(alpha > 0.0)) throw std::invalid_argument("invalid Ewald parameters");
    c->ewald_scratch.reserve(chui::ewald_scratch_bytes(modes));
    chui::ewald_device(x, y, z, m, n, targets, nt, L, G, alpha, real_images, modes, eps, ax, ay, az,
                       c->ewald_scratch.ptr, c->stream);
    CHUI_HIP(hipStreamSynchronize(c->stream));
  });
}

chui_status chui_treepm_split_diagnostics(double asmth, double rcut_cells, double h, chui_split_diagnostics* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null output");
    if (!(asmth > 0.0 && rcut_cells > 0.0 && h > 0.0)) throw std::invalid_argument("TreePM split policy must be > 0");
    const double a = asmth * h, rc = rcut_cells * h;
    auto S = [&](double r) {
      const double q = (0.5 * r) / a;
      return std::erfc(q) + 2.0 * 0.564189583547756286948079451560772586 * q * std::exp(-q * q);
    };
    auto Lf = [&](double r) {
      const double q = (0.5 * r) / a;
      return std::erf(q) - 2.0 * 0.564189583547756286948079451560772586 * q * std::exp(-q * q);
    };
    out->split_scale_comoving = a;
    out->cutoff_radius_comoving = rc;
    out->short_range_factor_at_split = S(a);
    out->long_range_factor_at_split = Lf(a);
    out->short_range_factor_at_cutoff = S(rc);
    out->long_range_factor_at_cutoff = Lf(rc);
    out->composition_error_at_split = std::abs(S(a) + Lf(a) - 1.0);
    double worst = 0.0;
    for (int i = 1; i <= 256; ++i) {
      const double r = rc * i / 256.0;
      worst = std::max(worst, std::abs(S(r) + Lf(r) - 1.0));
    }
    out->max_relative_composition_error = worst;
  });
}

chui_status chui_treepm_debug_inject_walk_fault(chui_treepm_context* c) {
  return guarded([&] {
    require_ctx(c);
    set_device(c);
    walk_counters(c, c->stream);
    chui::fill_u64(c->counters.as<uint64_t>() + chui::kFaultSlot, 1ull, 1, c->stream);
  });
}

const char* chui_treepm_last_error(void) { return g_last_error.c_str(); }

int32_t chui_treepm_abi_version(void) { return CHUI_TREEPM_ABI_VERSION; }

}  // extern "C"

// tree.hip — periodic unwrap, octant-path keys, level-synchronous octree build,
// preorder numbering, in-leaf Morton order and FP64 multipoles on gfx950.
//
// The reference builds its octree top-down by recursive stable octant
// partitions of the Morton-presorted particle list against FP64 cube centres
// (tree_gravity.cpp:1081-1206). The same tree is produced here without
// recursion:
//   1. every particle computes its 21-level octant path against the exact FP64
//      centre sequence c +- h/2 (the centres the recursion would visit);
//   2. one stable radix sort by that path gives the preorder leaf sequence;
//   3. nodes are discovered level by level with binary searches on the sorted
//      paths (a node is split while count > leaf_size and half > 1e-12);
//   4. inside each leaf the particles are put back into Morton order, because
//      stable partitions preserve the Morton order of a leaf's members.
// All FP64 arithmetic that decides topology or feeds multipoles is written in
// the reference's operation order and compiled with -ffp-contract=off, so keys,
// order, node membership and moments are bit-identical to the CPU path.
#include "tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "primitives.hpp"

namespace chui {
namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr uint32_t kInvalidIndex = 0xFFFFFFFFu;
constexpr int kPathLevels = 21;

struct Axes3 {
  const double* p[3];
};
struct UnwrapArgs {
  double L[3];
  int canonical[3];
  double anchor[3];
};

__device__ __forceinline__ double wrap_coord(double x, double L, int canonical) {
  // tree_pm_coupling.cpp:164-175
  double v = canonical ? x : x - L * floor(x / L);
  if (v >= L) v = 0.0;
  else if (v < 0.0) v += L;
  if (v == 0.0) v = 0.0;
  return v;
}

// ---------------------------------------------------------------------------
// Source statistics: per-axis min/max (order keys), finiteness and mass checks.
struct SourceStats {
  uint64_t lo[3];
  uint64_t hi[3];
  uint32_t bad_coord;
  uint32_t bad_mass;
  uint32_t mass_nonuniform;
  uint32_t pad;
  double mass0;  // m[0] (the uniform mass when every mass equals it)
};

// Each block writes its partial statistics to part[blockIdx.x]; one block then
// folds the partials (same-address atomics from every block would serialise
// at the L2, about 75 us for 1024 blocks whatever n is).
__global__ __launch_bounds__(256) void source_stats_kernel(Axes3 p, const double* __restrict__ m, std::size_t n,
                                                           SourceStats* __restrict__ part) {
  uint64_t lo[3] = {kEmpty, kEmpty, kEmpty}, hi[3] = {0, 0, 0};
  uint32_t bc = 0, bm = 0, nu = 0;
  const double m0 = (m != nullptr && n > 0) ? m[0] : 0.0;
  auto take = [&](const double* v, double mm) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!isfinite(v[a])) {
        bc = 1;
      } else {
        const uint64_t k = order_key(v[a]);
        lo[a] = k < lo[a] ? k : lo[a];
        hi[a] = k > hi[a] ? k : hi[a];
      }
    }
    if (m != nullptr) {
      if (!(isfinite(mm) && mm >= 0.0)) bm = 1;
      if (mm != m0) nu = 1;
    }
  };
  const std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
  std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  constexpr int kU = 4;  // loads of kU elements in flight per thread
  for (; i + (kU - 1) * stride < n; i += kU * stride) {
    double v[kU][3], mm[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
#pragma unroll
      for (int a = 0; a < 3; ++a) v[u][a] = p.p[a][i + u * stride];
      mm[u] = m != nullptr ? m[i + u * stride] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) take(v[u], mm[u]);
  }
  for (; i < n; i += stride) {
    const double v[3] = {p.p[0][i], p.p[1][i], p.p[2][i]};
    take(v, m != nullptr ? m[i] : 0.0);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = wave_min(lo[a]);
    hi[a] = wave_max(hi[a]);
  }
  bc = wave_max(bc);
  bm = wave_max(bm);
  nu = wave_max(nu);
  // block reduction, then one set of atomics per block (few blocks)
  __shared__ uint64_t slo[4][3], shi[4][3];
  __shared__ uint32_t sflag[4][3];
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      slo[w][a] = lo[a];
      shi[w][a] = hi[a];
    }
    sflag[w][0] = bc;
    sflag[w][1] = bm;
    sflag[w][2] = nu;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 1; v < static_cast<int>(blockDim.x) / kWave; ++v) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = slo[v][a] < lo[a] ? slo[v][a] : lo[a];
        hi[a] = shi[v][a] > hi[a] ? shi[v][a] : hi[a];
      }
      bc |= sflag[v][0];
      bm |= sflag[v][1];
      nu |= sflag[v][2];
    }
    SourceStats r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r.lo[a] = lo[a];
      r.hi[a] = hi[a];
    }
    r.bad_coord = bc;
    r.bad_mass = bm;
    r.mass_nonuniform = nu;
    r.pad = 0u;
    part[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void stats_final_kernel(const SourceStats* __restrict__ part, uint32_t nb,
                                                          const double* __restrict__ m, SourceStats* __restrict__ st) {
  uint64_t lo[3] = {kEmpty, kEmpty, kEmpty}, hi[3] = {0, 0, 0};
  uint32_t bc = 0, bm = 0, nu = 0;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const SourceStats& q = part[b];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = q.lo[a] < lo[a] ? q.lo[a] : lo[a];
      hi[a] = q.hi[a] > hi[a] ? q.hi[a] : hi[a];
    }
    bc |= q.bad_coord;
    bm |= q.bad_mass;
    nu |= q.mass_nonuniform;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = wave_min(lo[a]);
    hi[a] = wave_max(hi[a]);
  }
  bc = wave_max(bc);
  bm = wave_max(bm);
  nu = wave_max(nu);
  __shared__ SourceStats sw[256 / kWave];
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      sw[w].lo[a] = lo[a];
      sw[w].hi[a] = hi[a];
    }
    sw[w].bad_coord = bc;
    sw[w].bad_mass = bm;
    sw[w].mass_nonuniform = nu;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    SourceStats r = sw[0];
    for (int v = 1; v < static_cast<int>(blockDim.x) / kWave; ++v) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = sw[v].lo[a] < r.lo[a] ? sw[v].lo[a] : r.lo[a];
        r.hi[a] = sw[v].hi[a] > r.hi[a] ? sw[v].hi[a] : r.hi[a];
      }
      r.bad_coord |= sw[v].bad_coord;
      r.bad_mass |= sw[v].bad_mass;
      r.mass_nonuniform |= sw[v].mass_nonuniform;
    }
    r.pad = 0u;
    r.mass0 = m ? m[0] : 0.0;
    *st = r;
  }
}

// ---------------------------------------------------------------------------
// Periodic unwrap (tree_pm_coupling.cpp:141-218) without a sort.
// Values are binned into B >= 2n buckets of width w = L/B. The largest
// circular gap is >= L/n = 2w, while two values in one bucket differ by less
// than w (up to rounding), so the largest gap, and every gap tying with it,
// runs from the maximum of an occupied bucket p to the minimum of the next
// occupied bucket i. If e = i - p - 1 empty buckets lie between them, the gap
// is within (e w, (e + 2) w) up to rounding of the bucket computation, so only
// runs with e >= R - 2 (R the longest empty run) can hold the largest gap.
// Steps per axis: an occupancy bitmap (one byte store per value, packed); R from a scan
// of the bitmap; the end buckets of candidate runs and their predecessors are
// marked; exact FP64 min/max (IEEE bits; non-negative doubles order like their
// bits) are gathered only for marked buckets, ranked by a popcount scan; the
// candidate gaps are evaluated with the reference's tie rule (smaller anchor).
struct BucketArgs {
  double L[3];
  double scale[3];  // B / L
  int canonical[3];
  int naxes;
  uint64_t B;       // multiple of 32
  uint64_t W;       // B / 32 bitmap words per axis
};

// The per-build values that a capture-safe build keeps on the device instead
// of reading them back: the canonical flags (from the source statistics), the
// unwrap anchors, the root cube and the in-leaf Morton frame (from the unwrap
// result). Kernels given a Frame take these fields from it.
struct Frame {
  int canonical[3];
  int pad;
  double anchor[3];
  double centre[3];
  double half;
  double mlo[3];
  double mext;
};

__device__ __forceinline__ uint64_t bucket_of(double v, const BucketArgs& a, int ax) {
  uint64_t b = static_cast<uint64_t>(v * a.scale[ax]);
  return b >= a.B ? a.B - 1 : b;
}

// Occupancy as one byte per bucket: every value stores 1 into its bucket's
// byte. The stores are idempotent, so they need no atomics (a device-scope
// atomic per value and axis ran at about 1e10 per second, scattered); a
// coalesced pass then packs 32 bytes into each bitmap word.
__global__ __launch_bounds__(256) void occupancy_bytes_kernel(Axes3 p, std::size_t n, BucketArgs a,
                                                              uint8_t* __restrict__ occb,
                                                              const Frame* __restrict__ fr) {
  if (fr)
    for (int ax = 0; ax < 3; ++ax) a.canonical[ax] = fr->canonical[ax];
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    for (int ax = 0; ax < a.naxes; ++ax)
      occb[ax * a.B + bucket_of(wrap_coord(p.p[ax][i], a.L[ax], a.canonical[ax]), a, ax)] = 1u;
  }
}

__global__ __launch_bounds__(256) void occupancy_pack_kernel(const uint8_t* __restrict__ occb, std::size_t words,
                                                             uint32_t* __restrict__ occ) {
  const std::size_t wi = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (wi >= words) return;
  const uint4* src = reinterpret_cast<const uint4*>(occb + 32 * wi);
  const uint4 b0 = src[0], b1 = src[1];
  const uint32_t q[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  uint32_t bits = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
#pragma unroll
    for (int c = 0; c < 4; ++c) bits |= ((q[k] >> (8 * c)) & 1u) << (4 * k + c);
  }
  occ[wi] = bits;
}

// Exact min/max (IEEE bits, max stored +1 so 0 = empty) of the marked buckets.
__global__ __launch_bounds__(256) void marked_minmax_kernel(Axes3 p, std::size_t n, BucketArgs a,
                                                            const uint32_t* __restrict__ mark,
                                                            const uint32_t* __restrict__ rank_base,
                                                            uint64_t* __restrict__ cmin, uint64_t* __restrict__ cmax,
                                                            const Frame* __restrict__ fr) {
  if (fr)
    for (int ax = 0; ax < 3; ++ax) a.canonical[ax] = fr->canonical[ax];
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    for (int ax = 0; ax < a.naxes; ++ax) {
      const double v = wrap_coord(p.p[ax][i], a.L[ax], a.canonical[ax]);
      const uint64_t b = bucket_of(v, a, ax);
      const uint64_t wi = ax * a.W + (b >> 5);
      const uint32_t mw = mark[wi];
      const uint32_t bit = 1u << (b & 31);
      if (mw & bit) {
        const uint32_t r = rank_base[wi] + __popc(mw & (bit - 1u));
        const uint64_t bits = dbits(v);
        atomicMin(reinterpret_cast<unsigned long long*>(&cmin[r]), bits);
        atomicMax(reinterpret_cast<unsigned long long*>(&cmax[r]), bits + 1);
      }
    }
  }
}

__global__ __launch_bounds__(256) void popc_kernel(const uint32_t* __restrict__ w, std::size_t nw,
                                                   uint32_t* __restrict__ out) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < nw) out[i] = __popc(w[i]);
}

constexpr int kGapThreads = 256;
constexpr int kGapItems = 8;
constexpr int kGapTile = kGapThreads * kGapItems;

__device__ __forceinline__ uint64_t block_max_u64(uint64_t v) {
  __shared__ uint64_t wmax[kGapThreads / kWave];
  v = wave_max(v);
  if (lane_id() == 0) wmax[threadIdx.x / kWave] = v;
  __syncthreads();
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < kGapThreads / kWave; ++i) r = wmax[i] > r ? wmax[i] : r;
  __syncthreads();
  return r;
}

// Exclusive max-scan over the block (blocked thread order); returns the
// exclusive prefix for this thread and the block total.
__device__ __forceinline__ uint64_t block_excl_max_u64(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wtot[kGapThreads / kWave];
  const int lane = lane_id(), w = threadIdx.x / kWave;
  uint64_t inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint64_t o = __shfl_up(inc, off, kWave);
    if (lane >= off) inc = o > inc ? o : inc;
  }
  uint64_t excl = __shfl_up(inc, 1, kWave);
  if (lane == 0) excl = 0;
  if (lane == kWave - 1) wtot[w] = inc;
  __syncthreads();
  uint64_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kGapThreads / kWave; ++i) {
    if (i < w) woff = wtot[i] > woff ? wtot[i] : woff;
    tot = wtot[i] > tot ? wtot[i] : tot;
  }
  __syncthreads();
  *total = tot;
  return excl > woff ? excl : woff;
}

// Per block of kGapTile bitmap words: 1 + the last occupied bucket (0 if none).
__global__ __launch_bounds__(kGapThreads) void occ_blockmax_kernel(const uint32_t* __restrict__ occ, uint64_t W,
                                                                   uint64_t* __restrict__ blockmax, uint64_t nblk) {
  const int ax = blockIdx.y;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kGapTile;
  uint64_t mx = 0;
#pragma unroll
  for (int k = 0; k < kGapItems; ++k) {
    const uint64_t wi = base + static_cast<uint64_t>(k) * kGapThreads + threadIdx.x;
    if (wi < W) {
      const uint32_t v = occ[ax * W + wi];
      if (v) {
        const uint64_t last = wi * 32 + 31 - __clz(v);
        mx = last + 1 > mx ? last + 1 : mx;
      }
    }
  }
  mx = block_max_u64(mx);
  if (threadIdx.x == 0) blockmax[ax * nblk + blockIdx.x] = mx;
}

// One block per axis: exclusive max-scan of the block maxima in place.
__global__ __launch_bounds__(kGapThreads) void blockmax_scan_kernel(uint64_t* blockmax, uint64_t nblk,
                                                                    uint64_t* total) {
  const int ax = blockIdx.x;
  uint64_t carry = 0;
  for (uint64_t c = 0; c < nblk; c += kGapThreads) {
    const uint64_t idx = c + threadIdx.x;
    const uint64_t v = idx < nblk ? blockmax[ax * nblk + idx] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_max_u64(v, &tot);
    ex = ex > carry ? ex : carry;
    if (idx < nblk) blockmax[ax * nblk + idx] = ex;
    carry = tot > carry ? tot : carry;
  }
  if (threadIdx.x == 0) total[ax] = carry;
}

struct GapBest {
  double gap;
  double anchor;
  double umax;  // largest unwrapped coordinate if this gap is chosen
  int valid;
  int pad;
};

__device__ __forceinline__ bool better(const GapBest& a, const GapBest& b) {
  if (!a.valid) return false;
  if (!b.valid) return true;
  if (a.gap > b.gap) return true;
  if (a.gap == b.gap && a.anchor < b.anchor) return true;
  return false;
}

__device__ __forceinline__ GapBest shfl_best(const GapBest& v, int src) {
  GapBest r;
  r.gap = __shfl(v.gap, src, kWave);
  r.anchor = __shfl(v.anchor, src, kWave);
  r.umax = __shfl(v.umax, src, kWave);
  r.valid = __shfl(v.valid, src, kWave);
  r.pad = 0;
  return r;
}

__device__ GapBest block_best(GapBest b) {
  __shared__ GapBest wb[kGapThreads / kWave];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const GapBest o = shfl_best(b, lane_id() ^ off);
    if (better(o, b)) b = o;
  }
  if (lane_id() == 0) wb[threadIdx.x / kWave] = b;
  __syncthreads();
  GapBest r = wb[0];
  for (int i = 1; i < kGapThreads / kWave; ++i)
    if (better(wb[i], r)) r = wb[i];
  __syncthreads();
  return r;
}

// Walks the occupied buckets of a block's words in order. For each occupied
// bucket i it knows its predecessor p (the previous occupied bucket, or for the
// first one the last occupied bucket minus B: the wrap) and the empty run
// e = i - p - 1. MODE 0: the longest run per axis (atomicMax). MODE 1: marks
// the end buckets of runs with e >= R - 2 and their predecessors. MODE 2:
// evaluates the marked gaps from the exact min/max and keeps the best per block.
struct RunArgs {
  const uint32_t* occ;
  const uint64_t* block_prefix;  // exclusive max-scan of the block maxima
  const uint64_t* total;         // 1 + last occupied bucket, per axis
  unsigned long long* longest;   // per axis (MODE 0 writes, MODE 1 reads)
  uint32_t* mark;                // MODE 1 writes
  uint32_t* ends;                // MODE 1 writes
  const uint32_t* rank_base;     // MODE 2
  const uint64_t* cmin;          // MODE 2
  const uint64_t* cmax;          // MODE 2
  GapBest* block_out;            // MODE 2
  uint64_t nblk;
};

template <int MODE>
__global__ __launch_bounds__(kGapThreads) void runs_kernel(BucketArgs a, RunArgs r) {
  const int ax = blockIdx.y;
  const uint64_t W = a.W, B = a.B;
  const uint64_t wbase = static_cast<uint64_t>(blockIdx.x) * kGapTile + static_cast<uint64_t>(threadIdx.x) * kGapItems;
  // 1 + last occupied bucket among this thread's words
  uint64_t tmax = 0;
  uint32_t words[kGapItems];
#pragma unroll
  for (int k = 0; k < kGapItems; ++k) {
    const uint64_t wi = wbase + k;
    words[k] = wi < W ? r.occ[ax * W + wi] : 0u;
    if (words[k]) {
      const uint64_t last = wi * 32 + 31 - __clz(words[k]);
      tmax = last + 1 > tmax ? last + 1 : tmax;
    }
  }
  uint64_t tot;
  uint64_t run = block_excl_max_u64(tmax, &tot);
  const uint64_t bp = r.block_prefix[ax * r.nblk + blockIdx.x];
  run = run > bp ? run : bp;
  const uint64_t last_all = r.total[ax] - 1;  // last occupied bucket of the axis
  const unsigned long long R = MODE == 1 ? r.longest[ax] : 0ull;
  const double L = a.L[ax];
  uint64_t longest = 0;
  GapBest best{0.0, 0.0, 0.0, 0, 0};
  const auto rank_of = [&](uint64_t b) -> uint32_t {
    const uint64_t wi = ax * W + (b >> 5);
    const uint32_t bit = 1u << (b & 31);
    return r.rank_base[wi] + __popc(r.mark[wi] & (bit - 1u));
  };
  for (int k = 0; k < kGapItems; ++k) {
    uint32_t v = words[k];
    const uint64_t wi = wbase + k;
    uint32_t endbits = 0;
    while (v) {
      const int bitpos = __ffs(v) - 1;
      v &= v - 1;
      const uint64_t i = wi * 32 + bitpos;
      // run == 0: first occupied bucket; its predecessor is the last one, across the wrap
      const uint64_t e = run == 0 ? (B - 1 - last_all) + i : i - run;  // run = 1 + predecessor
      if (MODE == 0) {
        longest = e > longest ? e : longest;
      } else if (MODE == 1) {
        if (e + 2 >= R) {
          endbits |= 1u << bitpos;
          const uint64_t pb = run == 0 ? last_all : run - 1;
          atomicOr(&r.mark[ax * W + (pb >> 5)], 1u << (pb & 31));
          atomicOr(&r.mark[ax * W + (i >> 5)], 1u << (i & 31));
        }
      } else {
        if ((r.ends[ax * W + wi] >> bitpos) & 1u) {
          const double vmin = bitsd(r.cmin[rank_of(i)]);
          GapBest c;
          c.valid = 1;
          c.pad = 0;
          c.anchor = vmin;
          if (run == 0) {
            const double gmax = bitsd(r.cmax[rank_of(last_all)] - 1);
            c.gap = (vmin + L) - gmax;
            c.umax = gmax;
          } else {
            const double pr = bitsd(r.cmax[rank_of(run - 1)] - 1);
            c.gap = vmin - pr;
            c.umax = pr + L;
          }
          if (better(c, best)) best = c;
        }
      }
      run = i + 1;
    }
    if (MODE == 1 && endbits) r.ends[ax * W + wi] = endbits;
  }
  if (MODE == 0) {
    longest = wave_max(longest);
    if (lane_id() == 0 && longest) atomicMax(&r.longest[ax], static_cast<unsigned long long>(longest));
  } else if (MODE == 2) {
    best = block_best(best);
    if (threadIdx.x == 0) r.block_out[ax * r.nblk + blockIdx.x] = best;
  }
}

__global__ __launch_bounds__(kGapThreads) void gap_final_kernel(const GapBest* __restrict__ block_out, uint64_t nblk,
                                                                UnwrapResult* res) {
  const int ax = blockIdx.x;
  GapBest best{0.0, 0.0, 0.0, 0, 0};
  for (uint64_t i = threadIdx.x; i < nblk; i += kGapThreads) {
    const GapBest c = block_out[ax * nblk + i];
    if (better(c, best)) best = c;
  }
  best = block_best(best);
  if (threadIdx.x == 0) {
    res->anchor[ax] = best.anchor;
    res->umin[ax] = best.anchor;
    res->umax[ax] = best.umax;
  }
}

// ---------------------------------------------------------------------------
// Keys.
__device__ __forceinline__ uint64_t expand21(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

struct MortonArgs {
  double lo[3];
  double extent;  // max(max extent, 1e-12)  (tree_ordering.cpp:151)
};

// tree_ordering.cpp:159-167: clamp((x-min)/extent,0,1) * (2^21-1), llround.
__device__ __forceinline__ uint64_t morton_key(double x, double y, double z, const MortonArgs& a) {
  const double grid = 2097151.0;
  const double nx = fmin(fmax((x - a.lo[0]) / a.extent, 0.0), 1.0);
  const double ny = fmin(fmax((y - a.lo[1]) / a.extent, 0.0), 1.0);
  const double nz = fmin(fmax((z - a.lo[2]) / a.extent, 0.0), 1.0);
  const uint32_t qx = static_cast<uint32_t>(llround(nx * grid));
  const uint32_t qy = static_cast<uint32_t>(llround(ny * grid));
  const uint32_t qz = static_cast<uint32_t>(llround(nz * grid));
  return (expand21(qx) << 2) | (expand21(qy) << 1) | expand21(qz);
}

__device__ __forceinline__ double unwrap_coord(double x, double L, int canonical, double anchor) {
  const double v = wrap_coord(x, L, canonical);
  return v < anchor ? v + L : v;
}

struct PathArgs {
  UnwrapArgs u;
  double c0[3];
  double h0;
};

// 21-level octant path of every particle against the reference's FP64 centres.
// Also writes the sources in input order as canonical coordinates and mass
// (double4), so that the tree-order gather reads one 32-byte record per source.
__global__ __launch_bounds__(256) void path_key_kernel(Axes3 p, const double* __restrict__ m, std::size_t n,
                                                       PathArgs a, uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals, double4* __restrict__ packed,
                                                       const Frame* __restrict__ fr) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fr) {
    for (int ax = 0; ax < 3; ++ax) {
      a.u.canonical[ax] = fr->canonical[ax];
      a.u.anchor[ax] = fr->anchor[ax];
      a.c0[ax] = fr->centre[ax];
    }
    a.h0 = fr->half;
  }
  double x = p.p[0][i], y = p.p[1][i], z = p.p[2][i];
  if (fr) {
    // a capture-safe build reports non-finite coordinates through the fault
    // word and carries on; zeros keep every later cell index in range
    x = isfinite(x) ? x : 0.0;
    y = isfinite(y) ? y : 0.0;
    z = isfinite(z) ? z : 0.0;
  }
  packed[i] = make_double4(wrap_coord(x, a.u.L[0], a.u.canonical[0]), wrap_coord(y, a.u.L[1], a.u.canonical[1]),
                           wrap_coord(z, a.u.L[2], a.u.canonical[2]), m[i]);
  const double ux = unwrap_coord(x, a.u.L[0], a.u.canonical[0], a.u.anchor[0]);
  const double uy = unwrap_coord(y, a.u.L[1], a.u.canonical[1], a.u.anchor[1]);
  const double uz = unwrap_coord(z, a.u.L[2], a.u.canonical[2], a.u.anchor[2]);
  double cx = a.c0[0], cy = a.c0[1], cz = a.c0[2], h = a.h0;
  uint64_t key = 0;
  for (int lvl = 0; lvl < kPathLevels; ++lvl) {
    const uint32_t ox = ux >= cx ? 1u : 0u, oy = uy >= cy ? 1u : 0u, oz = uz >= cz ? 1u : 0u;
    key = (key << 3) | (ox << 2) | (oy << 1) | oz;
    const double ch = 0.5 * h;
    cx = cx + (ox ? ch : -ch);
    cy = cy + (oy ? ch : -ch);
    cz = cz + (oz ? ch : -ch);
    h = ch;
  }
  keys[i] = key;
  vals[i] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(256) void morton_keys_kernel(Axes3 p, std::size_t n, MortonArgs a,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = morton_key(p.p[0][i], p.p[1][i], p.p[2][i], a);
  vals[i] = static_cast<uint32_t>(i);
}

// Tree-order copy of the sources: canonical coordinates (wrap_coord) and mass.
// One gather through the sort order of path_key_kernel's input-order records;
// everything after it reads the copy.
__global__ __launch_bounds__(256) void gather_sources_kernel(const double4* __restrict__ packed, std::size_t n,
                                                             const uint32_t* __restrict__ order,
                                                             double4* __restrict__ pk) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pk[i] = packed[order[i]];
}

// unwrap_coord of a canonical coordinate
__device__ __forceinline__ double unwrap_canonical(double v, double anchor, double L) {
  return v < anchor ? v + L : v;
}

__global__ __launch_bounds__(256) void leaf_morton_kernel(const double4* __restrict__ pk, std::size_t n, UnwrapArgs u,
                                                          MortonArgs a, uint64_t* __restrict__ mk,
                                                          const Frame* __restrict__ fr) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fr) {
    for (int ax = 0; ax < 3; ++ax) {
      u.anchor[ax] = fr->anchor[ax];
      a.lo[ax] = fr->mlo[ax];
    }
    a.extent = fr->mext;
  }
  const double4 q = pk[i];
  mk[i] = morton_key(unwrap_canonical(q.x, u.anchor[0], u.L[0]), unwrap_canonical(q.y, u.anchor[1], u.L[1]),
                     unwrap_canonical(q.z, u.anchor[2], u.L[2]), a);
}

// ---------------------------------------------------------------------------
// Level-synchronous node discovery on the sorted path keys (BFS numbering).
// meta = depth | octant << 8 | leaf << 16
__device__ __forceinline__ uint32_t lower_bound_digit(const uint64_t* __restrict__ k, uint32_t b, uint32_t e,
                                                      int shift, uint32_t digit) {
  while (b < e) {
    const uint32_t mid = b + ((e - b) >> 1);
    if (static_cast<uint32_t>((k[mid] >> shift) & 7u) < digit) b = mid + 1;
    else e = mid;
  }
  return b;
}

// ---- level discovery driven from the device ---------------------------------
// One level of the octree per three launches, with the level's node range read
// from the device (lv[d], lv[d + 1]), so the host issues every level without
// reading anything back and reads the level starts and flags once at the end.
// A level's nodes are split into up to kLevelBlocks contiguous chunks, one per
// block. Count: eight threads per node, thread o of a splitting node
// binary-searches the first particle of octant o (the searches are
// independent, so a level costs one search depth rather than seven); the
// octant boundaries are kept for the emit pass and the group counts the
// non-empty octants. Offsets: one block turns the block sums into offsets and
// the next level's start. Emit: each block numbers its children in node order,
// so children are in BFS order and each node's children are contiguous in
// octant order.
constexpr int kLevelBlocks = 2048;
constexpr int kMaxLevels = kPathLevels + 1;            // depths 0..21
constexpr int kLvTooDeep = kMaxLevels + 2;              // lv[0, kMaxLevels + 2): level starts
constexpr int kLvOverflow = kLvTooDeep + 1;             // children beyond the node capacity
constexpr int kLvDeepSplit = kLvTooDeep + 2;            // a node below the partial sort's depth splits
constexpr int kLvLeaves = kLvTooDeep + 3;               // nodes without children
constexpr int kLvTotal = kLvTooDeep + 4;                // node count (capture-safe builds)
constexpr int kLvWords = kLvTooDeep + 5;

// nblocks: the grid of the count and emit passes of level d (at most
// kLevelBlocks); the offsets pass scans that many block sums.
struct LevelDevArgs {
  uint32_t d, leaf_size, capacity, sorted_levels, nblocks;
};

// Grid of the count and emit passes of level d: level d holds at most
// min(8^d, n) nodes and a block decides 32 nodes per step, so the shallow
// levels launch a few blocks instead of kLevelBlocks.
inline uint32_t level_blocks(int d, std::size_t n) {
  const uint64_t bound = d >= 8 ? static_cast<uint64_t>(n) : std::min<uint64_t>(1ull << (3 * d), n);
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(kLevelBlocks, (bound + 31) / 32)));
}

// exclusive prefix sum over the block (blockDim.x a multiple of the wave); total to *total
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
  const int lane = lane_id(), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += t;
  }
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < nw; ++i) {
      const uint32_t t = wsum[i];
      wsum[i] = run;
      run += t;
    }
    *total = run;
  }
  __syncthreads();
  const uint32_t r = wsum[w] + inc - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void level_count_dev_kernel(LevelDevArgs a, uint32_t* __restrict__ lv,
                                                              const uint32_t* __restrict__ nb,
                                                              const uint32_t* __restrict__ nc,
                                                              uint32_t* __restrict__ meta,
                                                              const double4* __restrict__ geo,
                                                              const uint64_t* __restrict__ keys,
                                                              uint32_t* __restrict__ nchild,
                                                              uint32_t* __restrict__ bounds,
                                                              uint32_t* __restrict__ bsum) {
  __shared__ uint32_t wsum[256 / kWave];
  __shared__ uint32_t tot;
  const uint32_t lb = lv[a.d], cnt = lv[a.d + 1] - lb;
  const uint32_t chunk = (cnt + gridDim.x - 1) / gridDim.x;
  const uint32_t k0 = min(cnt, blockIdx.x * chunk), k1 = min(cnt, k0 + chunk);
  uint32_t sum = 0, leaves = 0;
  bool deep = false, too = false;
  for (uint32_t base = k0; base < k1; base += blockDim.x / 8) {  // the same trip count for the whole block
    const uint32_t k = base + threadIdx.x / 8, o = threadIdx.x & 7u;
    const bool in = k < k1;
    const uint32_t node = lb + (in ? k : k0);
    const uint32_t count = in ? nc[node] : 0u;
    const double half = in ? geo[node].w : 0.0;
    const uint32_t depth = in ? (meta[node] & 0xFFu) : 0u;
    const bool leaf = count <= a.leaf_size || half <= 1.0e-12;  // tree_gravity.cpp:1133
    const bool too_deep = in && !leaf && depth >= static_cast<uint32_t>(kPathLevels);
    const bool split = in && !leaf && !too_deep;
    const uint32_t b = in ? nb[node] : 0u, e = b + count;
    uint32_t lo = b;
    if (split && o > 0) lo = lower_bound_digit(keys, b, e, 3 * (kPathLevels - 1 - static_cast<int>(depth)), o);
    const uint32_t nxt = __shfl_down(lo, 1, 8);
    const uint32_t hi = o == 7u ? e : nxt;
    uint32_t c = split && hi > lo ? 1u : 0u;
    c += __shfl_xor(c, 1, 8);
    c += __shfl_xor(c, 2, 8);
    c += __shfl_xor(c, 4, 8);
    if (split) bounds[8ull * k + o] = lo;
    if (in && o == 0u) {
      if (split) {
        nchild[node] = c;
        sum += c;
        deep = deep || depth + 1u > a.sorted_levels;
      } else {
        nchild[node] = 0;
        meta[node] |= 1u << 16;
        too = too || too_deep;
        ++leaves;
      }
    }
  }
  if (deep) atomicOr(&lv[kLvDeepSplit], 1u);
  if (too) atomicOr(&lv[kLvTooDeep], 1u);
  (void)block_exclusive_scan(leaves, wsum, &tot);
  if (threadIdx.x == 0 && tot) atomicAdd(&lv[kLvLeaves], tot);
  (void)block_exclusive_scan(sum, wsum, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// One block of kLevelBlocks / 2 threads: block sums -> exclusive offsets and
// the next level's start (or the overflow flag).
__global__ __launch_bounds__(kLevelBlocks / 2) void level_offsets_dev_kernel(LevelDevArgs a, uint32_t* __restrict__ lv,
                                                                            uint32_t* __restrict__ bsum) {
  __shared__ uint32_t wsum[kLevelBlocks / 2 / kWave];
  __shared__ uint32_t tot;
  const uint32_t i = 2 * threadIdx.x;
  const uint32_t v0 = i < a.nblocks ? bsum[i] : 0u, v1 = i + 1 < a.nblocks ? bsum[i + 1] : 0u;
  const uint32_t off = block_exclusive_scan(v0 + v1, wsum, &tot);
  if (i < a.nblocks) bsum[i] = off;
  if (i + 1 < a.nblocks) bsum[i + 1] = off + v0;
  if (threadIdx.x == 0) {
    const uint32_t le = lv[a.d + 1];
    const uint64_t next = static_cast<uint64_t>(le) + tot;
    const bool over = lv[kLvOverflow] != 0u || next > a.capacity;
    if (over) lv[kLvOverflow] = 1u;
    lv[a.d + 2] = over ? le : static_cast<uint32_t>(next);
  }
}

__global__ __launch_bounds__(256) void level_emit_dev_kernel(LevelDevArgs a, const uint32_t* __restrict__ lv,
                                                             uint32_t* __restrict__ nb, uint32_t* __restrict__ nc,
                                                             uint32_t* __restrict__ meta, double4* __restrict__ geo,
                                                             const uint32_t* __restrict__ bounds,
                                                             const uint32_t* __restrict__ nchild,
                                                             const uint32_t* __restrict__ bsum,
                                                             uint32_t* __restrict__ first_child) {
  __shared__ uint32_t wsum[256 / kWave];
  __shared__ uint32_t tot;
  if (lv[kLvOverflow]) return;  // block-uniform
  const uint32_t lb = lv[a.d], le = lv[a.d + 1], cnt = le - lb;
  const uint32_t chunk = (cnt + gridDim.x - 1) / gridDim.x;
  const uint32_t k0 = min(cnt, blockIdx.x * chunk), k1 = min(cnt, k0 + chunk);
  uint32_t run = le + bsum[blockIdx.x];
  for (uint32_t base = k0; base < k1; base += blockDim.x) {
    const uint32_t k = base + threadIdx.x;
    const bool in = k < k1;
    const uint32_t node = lb + (in ? k : k0);
    const uint32_t c = in ? nchild[node] : 0u;
    const uint32_t pre = block_exclusive_scan(c, wsum, &tot);
    if (in) {
      if (c == 0) {
        first_child[node] = kInvalidIndex;
      } else {
        const uint32_t depth = meta[node] & 0xFFu;
        const uint32_t e = nb[node] + nc[node];
        const double4 g = geo[node];
        const uint32_t fc = run + pre;
        first_child[node] = fc;
        uint32_t j = 0;
        for (uint32_t o = 0; o < 8; ++o) {
          const uint32_t lo = bounds[8ull * k + o];
          const uint32_t hi = o == 7 ? e : bounds[8ull * k + o + 1];
          if (hi > lo) {
            const uint32_t ch = fc + j++;
            nb[ch] = lo;
            nc[ch] = hi - lo;
            meta[ch] = (depth + 1u) | (o << 8);
            const double chh = 0.5 * g.w;  // tree_gravity.cpp:1180-1184
            double4 cg;
            cg.x = g.x + ((o & 4u) ? chh : -chh);
            cg.y = g.y + ((o & 2u) ? chh : -chh);
            cg.z = g.z + ((o & 1u) ? chh : -chh);
            cg.w = chh;
            geo[ch] = cg;
          }
        }
      }
    }
    run += tot;
  }
}

__global__ void level_dev_init_kernel(uint32_t* lv) {
  const int i = threadIdx.x;
  if (i < kLvWords) lv[i] = i == 1 ? 1u : 0u;
}

__global__ void root_init_kernel(uint32_t* nb, uint32_t* nc, uint32_t* meta, double4* geo, uint32_t n, double cx,
                                 double cy, double cz, double h, const Frame* __restrict__ fr) {
  nb[0] = 0;
  nc[0] = n;
  meta[0] = 0;
  geo[0] = fr ? make_double4(fr->centre[0], fr->centre[1], fr->centre[2], fr->half) : make_double4(cx, cy, cz, h);
}

// A level's node range: [lb, lb + cnt) from the host, or [lv[d], lv[d + 1])
// from the device when lv is given.
struct LevelRange {
  uint32_t lb, cnt;
  const uint32_t* lv;
  int d;
};
__device__ __forceinline__ void level_range(const LevelRange& r, uint32_t& lb, uint32_t& cnt) {
  if (r.lv) {
    lb = r.lv[r.d];
    cnt = r.lv[r.d + 1] - lb;
  } else {
    lb = r.lb;
    cnt = r.cnt;
  }
}

// Bottom-up subtree sizes (in nodes), one level per launch.
__global__ __launch_bounds__(256) void subtree_kernel(LevelRange r, const uint32_t* __restrict__ nchild,
                                                      const uint32_t* __restrict__ first_child,
                                                      uint32_t* __restrict__ subtree) {
  uint32_t lb, cnt;
  level_range(r, lb, cnt);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
    const uint32_t node = lb + k;
    uint32_t s = 1;
    const uint32_t c = nchild[node];
    for (uint32_t j = 0; j < c; ++j) s += subtree[first_child[node] + j];
    subtree[node] = s;
  }
}

// Top-down preorder indices, one level per launch.
__global__ __launch_bounds__(256) void preorder_kernel(LevelRange r, const uint32_t* __restrict__ nchild,
                                                       const uint32_t* __restrict__ first_child,
                                                       const uint32_t* __restrict__ subtree,
                                                       uint32_t* __restrict__ pre) {
  uint32_t lb, cnt;
  level_range(r, lb, cnt);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
    const uint32_t node = lb + k;
    uint32_t run = pre[node] + 1;
    const uint32_t c = nchild[node];
    for (uint32_t j = 0; j < c; ++j) {
      const uint32_t ch = first_child[node] + j;
      pre[ch] = run;
      run += subtree[ch];
    }
  }
}

// In-leaf Morton order: insertion sort of (morton key, source row) pairs.
__device__ __forceinline__ void leaf_sort_node(uint32_t node, const uint32_t* __restrict__ nb,
                                               const uint32_t* __restrict__ nc, uint64_t* __restrict__ mk,
                                               uint32_t* __restrict__ order, uint32_t* __restrict__ from,
                                               double4* __restrict__ pk) {
  const uint32_t b = nb[node], e = b + nc[node];
  for (uint32_t i = b; i < e; ++i) from[i] = i;
  for (uint32_t i = b + 1; i < e; ++i) {
    const uint64_t kk = mk[i];
    const uint32_t kv = order[i];
    const uint32_t kf = from[i];
    uint32_t j = i;
    while (j > b && (mk[j - 1] > kk || (mk[j - 1] == kk && order[j - 1] > kv))) {
      mk[j] = mk[j - 1];
      order[j] = order[j - 1];
      from[j] = from[j - 1];
      --j;
    }
    mk[j] = kk;
    order[j] = kv;
    from[j] = kf;
  }
  // Apply the leaf's permutation to the source copy in place, one cycle at a
  // time: position j takes the record that was at from[j]; a written position
  // is marked by from[j] = j, like a fixed point.
  for (uint32_t c = b; c < e; ++c) {
    if (from[c] == c) continue;
    const double4 first = pk[c];
    uint32_t j = c;
    for (;;) {
      const uint32_t k = from[j];
      from[j] = j;
      if (k == c) {
        pk[j] = first;
        break;
      }
      pk[j] = pk[k];
      j = k;
    }
  }
}

// total_dev (when given) holds the node count of a capture-safe build.
__global__ __launch_bounds__(256) void leaf_sort_kernel(uint32_t total_nodes, const uint32_t* __restrict__ total_dev,
                                                        const uint32_t* __restrict__ nchild,
                                                        const uint32_t* __restrict__ nb, const uint32_t* __restrict__ nc,
                                                        uint64_t* __restrict__ mk, uint32_t* __restrict__ order,
                                                        uint32_t* __restrict__ from, double4* __restrict__ pk) {
  const uint32_t total = total_dev ? *total_dev : total_nodes;
  for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < total; node += gridDim.x * blockDim.x)
    if (nchild[node] == 0) leaf_sort_node(node, nb, nc, mk, order, from, pk);
}

// Multipoles, one BFS level per launch from the deepest level up
// (tree_gravity.cpp:1208-1380, same operation order).
struct MomArgs {
  UnwrapArgs u;
  LevelRange r;
  int quad;
};

__device__ __forceinline__ void multipole_node(uint32_t node, const MomArgs& a, const double4* __restrict__ pk,
                                               const uint32_t* __restrict__ nb, const uint32_t* __restrict__ nc,
                                               const uint32_t* __restrict__ nchild,
                                               const uint32_t* __restrict__ first_child,
                                               const double4* __restrict__ geo, double4* __restrict__ mom,
                                               double* __restrict__ quad, double* __restrict__ box) {
  const double4 g = geo[node];
  double* q = quad + 8ull * node;
  const uint32_t c = nchild[node];
  // tight bounding box of the node's particles (unwrapped): lo[3], hi[3]
  double lo[3] = {g.x, g.y, g.z}, hi[3] = {g.x, g.y, g.z};
  if (c == 0) {
    const uint32_t b = nb[node], e = b + nc[node];
    double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    if (e > b) {
      for (int ax = 0; ax < 3; ++ax) {
        lo[ax] = INFINITY;
        hi[ax] = -INFINITY;
      }
    }
    for (uint32_t i = b; i < e; ++i) {
      const double4 q = pk[i];
      const double mm = q.w;
      const double ux = unwrap_canonical(q.x, a.u.anchor[0], a.u.L[0]);
      const double uy = unwrap_canonical(q.y, a.u.anchor[1], a.u.L[1]);
      const double uz = unwrap_canonical(q.z, a.u.anchor[2], a.u.L[2]);
      tm += mm;
      wx += mm * ux;
      wy += mm * uy;
      wz += mm * uz;
      lo[0] = fmin(lo[0], ux);
      lo[1] = fmin(lo[1], uy);
      lo[2] = fmin(lo[2], uz);
      hi[0] = fmax(hi[0], ux);
      hi[1] = fmax(hi[1], uy);
      hi[2] = fmax(hi[2], uz);
    }
    if (box) {
      for (int ax = 0; ax < 3; ++ax) {
        box[6ull * node + ax] = lo[ax];
        box[6ull * node + 3 + ax] = hi[ax];
      }
    }
    double4 M = make_double4(g.x, g.y, g.z, tm);
    if (tm > 0.0) {
      M.x = wx / tm;
      M.y = wy / tm;
      M.z = wz / tm;
    }
    mom[node] = M;
    double ixx = 0, ixy = 0, ixz = 0, iyy = 0, iyz = 0, izz = 0;
    if (a.quad && tm > 0.0) {
      for (uint32_t i = b; i < e; ++i) {
        const double4 q = pk[i];
        const double mm = q.w;
        const double rx = unwrap_canonical(q.x, a.u.anchor[0], a.u.L[0]) - M.x;
        const double ry = unwrap_canonical(q.y, a.u.anchor[1], a.u.L[1]) - M.y;
        const double rz = unwrap_canonical(q.z, a.u.anchor[2], a.u.L[2]) - M.z;
        ixx += mm * rx * rx;
        ixy += mm * rx * ry;
        ixz += mm * rx * rz;
        iyy += mm * ry * ry;
        iyz += mm * ry * rz;
        izz += mm * rz * rz;
      }
      const double tr = ixx + iyy + izz;
      q[0] = 3.0 * ixx - tr;
      q[1] = 3.0 * ixy;
      q[2] = 3.0 * ixz;
      q[3] = 3.0 * iyy - tr;
      q[4] = 3.0 * iyz;
      q[5] = 3.0 * izz - tr;
      q[6] = tr;
    } else {
      for (int j = 0; j < 7; ++j) q[j] = 0.0;
    }
    const double ox = g.x - M.x, oy = g.y - M.y, oz = g.z - M.z;
    q[7] = sqrt(ox * ox + oy * oy + oz * oz);
    return;
  }
  const uint32_t fc = first_child[node];
  double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
  for (uint32_t j = 0; j < c; ++j) {
    const double4 cm = mom[fc + j];
    tm += cm.w;
    wx += cm.w * cm.x;
    wy += cm.w * cm.y;
    wz += cm.w * cm.z;
  }
  if (box) {
    for (int ax = 0; ax < 3; ++ax) {
      lo[ax] = INFINITY;
      hi[ax] = -INFINITY;
    }
    for (uint32_t j = 0; j < c; ++j) {
      const double* cb = box + 6ull * (fc + j);
      for (int ax = 0; ax < 3; ++ax) {
        lo[ax] = fmin(lo[ax], cb[ax]);
        hi[ax] = fmax(hi[ax], cb[3 + ax]);
      }
    }
    for (int ax = 0; ax < 3; ++ax) {
      box[6ull * node + ax] = lo[ax];
      box[6ull * node + 3 + ax] = hi[ax];
    }
  }
  double4 M = make_double4(g.x, g.y, g.z, tm);
  if (tm > 0.0) {
    M.x = wx / tm;
    M.y = wy / tm;
    M.z = wz / tm;
  }
  mom[node] = M;
  if (a.quad && tm > 0.0) {
    double qxx = 0, qxy = 0, qxz = 0, qyy = 0, qyz = 0, qzz = 0, tr = 0;
    for (uint32_t j = 0; j < c; ++j) {
      const double4 cm = mom[fc + j];
      const double* cq = quad + 8ull * (fc + j);
      const double dx = cm.x - M.x, dy = cm.y - M.y, dz = cm.z - M.z;
      const double mm = cm.w;
      const double d2 = dx * dx + dy * dy + dz * dz;
      qxx += cq[0] + mm * (3.0 * dx * dx - d2);
      qxy += cq[1] + mm * (3.0 * dx * dy);
      qxz += cq[2] + mm * (3.0 * dx * dz);
      qyy += cq[3] + mm * (3.0 * dy * dy - d2);
      qyz += cq[4] + mm * (3.0 * dy * dz);
      qzz += cq[5] + mm * (3.0 * dz * dz - d2);
      tr += cq[6] + mm * d2;
    }
    q[0] = qxx; q[1] = qxy; q[2] = qxz; q[3] = qyy; q[4] = qyz; q[5] = qzz; q[6] = tr;
  } else {
    for (int j = 0; j < 7; ++j) q[j] = 0.0;
  }
  const double ox = g.x - M.x, oy = g.y - M.y, oz = g.z - M.z;
  q[7] = sqrt(ox * ox + oy * oy + oz * oz);
}

__global__ __launch_bounds__(128) void multipole_kernel(MomArgs a, const double4* __restrict__ pk,
                                                        const uint32_t* __restrict__ nb,
                                                        const uint32_t* __restrict__ nc,
                                                        const uint32_t* __restrict__ nchild,
                                                        const uint32_t* __restrict__ first_child,
                                                        const double4* __restrict__ geo, double4* __restrict__ mom,
                                                        double* __restrict__ quad /* 8 per node */,
                                                        double* __restrict__ box /* 6 per node or null */,
                                                        const Frame* __restrict__ fr) {
  if (fr)
    for (int ax = 0; ax < 3; ++ax) a.u.anchor[ax] = fr->anchor[ax];
  uint32_t lb, cnt;
  level_range(a.r, lb, cnt);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x)
    multipole_node(lb + k, a, pk, nb, nc, nchild, first_child, geo, mom, quad, box);
}

__device__ __forceinline__ uint32_t to_fixed(double c, double scale) {
  // scale = 2^32 / L; wraps modulo 2^32 (periodic)
  const double f = floor(c * scale);
  return static_cast<uint32_t>(static_cast<uint64_t>(static_cast<int64_t>(f)));
}

struct RecordArgs {
  uint32_t total;
  int fp64;   // FP64 records and child indices
  int mixed;  // fixed-point/FP32 records and tight boxes
  int cubic;  // separation units of the mixed walk: fixed point (cubic box) or lengths
  double fscale[3];
};

// x rounded up to a float (x >= 0)
__device__ __forceinline__ float float_up(double x) {
  const float f = static_cast<float>(x);
  return static_cast<double>(f) < x ? nextafterf(f, INFINITY) : f;
}

// BFS -> preorder records: links always, FP64 records and/or the mixed walk's
// fixed-point/FP32 records.
__global__ __launch_bounds__(256) void records_kernel(RecordArgs a, const uint32_t* __restrict__ pre,
                                                      const uint32_t* __restrict__ nb, const uint32_t* __restrict__ nc,
                                                      const uint32_t* __restrict__ nchild,
                                                      const uint32_t* __restrict__ first_child,
                                                      const uint32_t* __restrict__ meta,
                                                      const uint32_t* __restrict__ subtree,
                                                      const double4* __restrict__ geo,
                                                      const double4* __restrict__ mom,
                                                      const double* __restrict__ quad, NodeLink* __restrict__ link,
                                                      NodeGeo64* __restrict__ geo64, NodeMom64* __restrict__ mom64,
                                                      NodeQuad64* __restrict__ quad64,
                                                      uint32_t* __restrict__ child_index, NodeGeo32* __restrict__ geo32,
                                                      NodeMom32* __restrict__ mom32,
                                                      NodeQuad32* __restrict__ quad32,
                                                      const double* __restrict__ box,
                                                      NodeBox32* __restrict__ box32,
                                                      const uint32_t* __restrict__ total_dev) {
  const uint32_t total = total_dev ? *total_dev : a.total;
  for (uint32_t node = blockIdx.x * blockDim.x + threadIdx.x; node < total; node += gridDim.x * blockDim.x) {
  const uint32_t p = pre[node];
  const uint32_t c = nchild[node];
  link[p] = NodeLink{nb[node], nc[node], subtree[node], c};
  const double4 g = geo[node];
  const double4 M = mom[node];
  const double* q = quad + 8ull * node;
  if (a.fp64) {
    geo64[p] = NodeGeo64{g.x, g.y, g.z, g.w};
    mom64[p] = NodeMom64{M.x, M.y, M.z, M.w};
    quad64[p] = NodeQuad64{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]};
    uint32_t ci[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) ci[o] = kInvalidIndex;
    for (uint32_t j = 0; j < c; ++j) {
      const uint32_t ch = first_child[node] + j;
      ci[(meta[ch] >> 8) & 7u] = pre[ch];
    }
#pragma unroll
    for (int o = 0; o < 8; ++o) child_index[8ull * p + o] = ci[o];
  }
  if (a.mixed) {
    geo32[p] = NodeGeo32{to_fixed(g.x, a.fscale[0]), to_fixed(g.y, a.fscale[1]), to_fixed(g.z, a.fscale[2]),
                         static_cast<float>(g.w)};
    mom32[p] = NodeMom32{to_fixed(M.x, a.fscale[0]), to_fixed(M.y, a.fscale[1]), to_fixed(M.z, a.fscale[2]),
                         static_cast<float>(M.w)};
    quad32[p] = NodeQuad32{static_cast<float>(q[0]), static_cast<float>(q[1]), static_cast<float>(q[2]),
                           static_cast<float>(q[3]), static_cast<float>(q[4]), static_cast<float>(q[5]),
                           static_cast<float>(q[6]), static_cast<float>(q[7])};
    // Fixed point floors: the centre and every particle move by less than one
    // unit each, so two units on top of the half-extent keep every particle's
    // fixed-point position inside the box.
    const double* bx = box + 6ull * node;
    NodeBox32 r;
    uint32_t cf[3];
    float hf[3];
    for (int ax = 0; ax < 3; ++ax) {
      const double half = 0.5 * (bx[3 + ax] - bx[ax]);
      cf[ax] = to_fixed(0.5 * (bx[ax] + bx[3 + ax]), a.fscale[ax]);
      hf[ax] = a.cubic ? float_up(half * a.fscale[ax] + 2.0) : float_up(half + 2.0 / a.fscale[ax]);
    }
    r.cx = cf[0];
    r.cy = cf[1];
    r.cz = cf[2];
    r.hx = hf[0];
    r.hy = hf[1];
    r.hz = hf[2];
    r.cell_half = static_cast<float>(g.w);
    r.pad = 0u;
    box32[p] = r;
  }
  }
}

__global__ void unwrap_apply_kernel(const double* in, std::size_t n, double L, int canonical, double anchor,
                                    double* out) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = unwrap_coord(in[i], L, canonical, anchor);
}

__global__ void pos_of_row_kernel(const uint32_t* __restrict__ order, std::size_t n, uint32_t* __restrict__ pos) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) pos[order[i]] = static_cast<uint32_t>(i);
}

// ---- capture-safe build: per-build values kept on the device ------------------
struct FrameCheck {
  double L[3];
  int expect_uniform;  // mass uniformity of the prepared build
  double expect_mass;  // its uniform mass
};

// Canonical flags from the statistics (tree_pm_coupling.cpp:3185-3195, as the
// ordinary build decides them on the host); invalid inputs and a mass
// uniformity other than the prepared build's set fault bits.
__global__ void frame_stats_kernel(const SourceStats* __restrict__ st, const double* __restrict__ m, FrameCheck c,
                                   Frame* __restrict__ fr, unsigned long long* __restrict__ fault) {
  if (threadIdx.x != 0) return;
  unsigned long long f = 0;
  if (st->bad_coord) f |= kFaultCoordinates;
  if (st->bad_mass) f |= kFaultMasses;
  const int uniform = st->mass_nonuniform == 0 ? 1 : 0;
  if (uniform != c.expect_uniform || (uniform && m[0] != c.expect_mass)) f |= kFaultPrepared;
  for (int a = 0; a < 3; ++a) {
    const double lo = order_key_inv(st->lo[a]), hi = order_key_inv(st->hi[a]);
    fr->canonical[a] = (lo >= 0.0 && hi < c.L[a]) ? 1 : 0;
  }
  fr->pad = 0;
  if (f) atomicOr(fault, f);
}

// Root cube (tree_gravity.cpp:443-452) and the in-leaf Morton frame
// (tree_ordering.cpp:151) from the unwrap result, with the ordinary build's
// host arithmetic (max as (a < b) ? b : a, no contraction).
__global__ void frame_root_kernel(const UnwrapResult* __restrict__ res, Frame* __restrict__ fr) {
  if (threadIdx.x != 0) return;
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double e = res->umax[a] - res->umin[a];
    ext = ext < e ? e : ext;
  }
  const double max_extent = ext < 1.0e-10 ? 1.0e-10 : ext;
  for (int a = 0; a < 3; ++a) {
    fr->anchor[a] = res->anchor[a];
    fr->centre[a] = 0.5 * (res->umin[a] + res->umax[a]);
    fr->mlo[a] = res->umin[a];
  }
  fr->half = 0.5 * max_extent * (1.0 + 1.0e-8);
  fr->mext = ext < 1.0e-12 ? 1.0e-12 : ext;
}

// After a capture-safe level discovery of `issued` levels: the node count,
// and fault bits for a node capacity or depth beyond the prepared build or a
// tree deeper than 21 levels.
__global__ void level_check_kernel(uint32_t* __restrict__ lv, int issued, unsigned long long* __restrict__ fault) {
  if (threadIdx.x != 0) return;
  unsigned long long f = 0;
  if (lv[kLvOverflow]) f |= kFaultPrepared;
  if (lv[kLvTooDeep]) f |= kFaultTooDeep;
  if (issued < kMaxLevels && lv[issued + 1] > lv[issued]) f |= kFaultPrepared;
  if (f) {
    // The tree is incomplete (children counted but not emitted, or levels left
    // out): every level and the node count become empty, so the passes after
    // this one, the deposit and the walk do nothing; the fault word reports it.
    for (int d = 0; d < kLvTooDeep; ++d) lv[d] = 0;
    lv[kLvTotal] = 0;
    atomicOr(fault, f);
  } else {
    lv[kLvTotal] = lv[issued];
  }
}

// p[0, *count) = v, the count read on the device.
template <class T>
__global__ __launch_bounds__(256) void fill_count_kernel(T* __restrict__ p, T v, const uint32_t* __restrict__ count) {
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <class T>
void grow_copy(DeviceBuffer& b, std::size_t old_count, std::size_t new_count, hipStream_t s) {
  if (new_count * sizeof(T) <= b.bytes) return;
  DeviceBuffer nb;
  nb.reserve(new_count * sizeof(T));
  if (old_count) CHUI_HIP(hipMemcpyAsync(nb.ptr, b.ptr, old_count * sizeof(T), hipMemcpyDeviceToDevice, s));
  std::swap(b.ptr, nb.ptr);
  std::swap(b.bytes, nb.bytes);
  CHUI_HIP(hipStreamSynchronize(s));  // old buffer freed by nb's destructor after the copy
}

// Source statistics -> host.
SourceStats compute_stats(Tree& t, Axes3 p, const double* m, std::size_t n, hipStream_t s) {
  const unsigned nb = n > 0 ? std::min<unsigned>(grid_for(n, 256), 1024u) : 0u;
  t.bounds.reserve((1 + std::max(nb, 1u)) * sizeof(SourceStats));
  SourceStats* st = t.bounds.as<SourceStats>();
  if (n > 0) {
    source_stats_kernel<<<nb, 256, 0, s>>>(p, m, n, st + 1);
    CHUI_LAUNCH_CHECK("source_stats_kernel");
  }
  stats_final_kernel<<<1, 256, 0, s>>>(st + 1, nb, n > 0 ? m : nullptr, st);
  CHUI_LAUNCH_CHECK("stats_final_kernel");
  SourceStats h{};
  t.readback.read(&h, st, sizeof(SourceStats), s);
  return h;
}

double host_order_key_inv(uint64_t k) {
  const uint64_t b = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

// Unwraps naxes axes at once; fills res (host).
// With fr (a capture-safe build) the canonical flags come from the frame, the
// marked-bucket buffers are sized by their bound (marked buckets are occupied
// buckets, at most n per axis) and nothing is read back: the result stays in
// the device UnwrapResult (returned).
UnwrapResult* unwrap_axes(Tree& t, Axes3 p, int naxes, std::size_t n, const double* L, const int* canonical,
                          UnwrapResult& res, hipStream_t s, const Frame* fr = nullptr, bool size_bound = false) {
  for (int ax = 0; ax < naxes; ++ax) {
    res.anchor[ax] = 0.0;
    res.umin[ax] = 0.0;
    res.umax[ax] = 0.0;
    res.canonical[ax] = canonical[ax];
  }
  if (n == 0) return nullptr;
  BucketArgs a{};
  const uint64_t B = (std::max<uint64_t>(2ull * n, 64ull) + 31) / 32 * 32;
  a.B = B;
  a.W = B / 32;
  a.naxes = naxes;
  for (int ax = 0; ax < naxes; ++ax) {
    a.L[ax] = L[ax];
    a.scale[ax] = static_cast<double>(B) / L[ax];
    a.canonical[ax] = canonical[ax];
  }
  const uint64_t W = a.W, NW = naxes * W;
  const uint64_t nblk = (W + kGapTile - 1) / kGapTile;
  // bitmaps: occupancy, marks, run ends; ranks; block scan data
  t.bucket_min.reserve(3 * NW * sizeof(uint32_t) + 2 * NW * sizeof(uint32_t));
  uint32_t* occ = t.bucket_min.as<uint32_t>();
  uint32_t* mark = occ + NW;
  uint32_t* ends = mark + NW;
  uint32_t* popc = ends + NW;
  uint32_t* rank_base = popc + NW;
  t.block_max.reserve((naxes * nblk + 8) * sizeof(uint64_t));
  uint64_t* blockmax = t.block_max.as<uint64_t>();
  uint64_t* total = blockmax + naxes * nblk;               // 3 entries
  unsigned long long* longest = reinterpret_cast<unsigned long long*>(total + 3);  // 3 entries
  uint32_t* marked = reinterpret_cast<uint32_t*>(total + 6);  // total marked buckets
  t.block_best.reserve((naxes * nblk + 1) * sizeof(GapBest) + sizeof(UnwrapResult));
  GapBest* best = t.block_best.as<GapBest>();
  UnwrapResult* dres = reinterpret_cast<UnwrapResult*>(best + naxes * nblk + 1);
  fill_u32(occ, 0u, 3 * NW, s);
  fill_u64(total, 0ull, 8, s);
  const unsigned pgrid = std::min<unsigned>(grid_for(n, 256), 8192u);
  // bytes per bucket (B per axis, a multiple of 32), then packed into occ
  t.occ_bytes.reserve(naxes * B);
  uint8_t* occb = t.occ_bytes.as<uint8_t>();
  fill_u32(reinterpret_cast<uint32_t*>(occb), 0u, naxes * B / 4, s);
  occupancy_bytes_kernel<<<pgrid, 256, 0, s>>>(p, n, a, occb, fr);
  CHUI_LAUNCH_CHECK("occupancy_bytes_kernel");
  occupancy_pack_kernel<<<grid_for(NW, 256), 256, 0, s>>>(occb, NW, occ);
  CHUI_LAUNCH_CHECK("occupancy_pack_kernel");
  occ_blockmax_kernel<<<dim3(static_cast<unsigned>(nblk), naxes), kGapThreads, 0, s>>>(occ, W, blockmax, nblk);
  CHUI_LAUNCH_CHECK("occ_blockmax_kernel");
  blockmax_scan_kernel<<<naxes, kGapThreads, 0, s>>>(blockmax, nblk, total);
  CHUI_LAUNCH_CHECK("blockmax_scan_kernel");
  RunArgs ra{};
  ra.occ = occ;
  ra.block_prefix = blockmax;
  ra.total = total;
  ra.longest = longest;
  ra.mark = mark;
  ra.ends = ends;
  ra.rank_base = rank_base;
  ra.block_out = best;
  ra.nblk = nblk;
  const dim3 rgrid(static_cast<unsigned>(nblk), naxes);
  runs_kernel<0><<<rgrid, kGapThreads, 0, s>>>(a, ra);
  CHUI_LAUNCH_CHECK("runs_kernel<0>");
  runs_kernel<1><<<rgrid, kGapThreads, 0, s>>>(a, ra);
  CHUI_LAUNCH_CHECK("runs_kernel<1>");
  popc_kernel<<<grid_for(NW, 256), 256, 0, s>>>(mark, NW, popc);
  CHUI_LAUNCH_CHECK("popc_kernel");
  t.bfs_scan.reserve(scan_scratch_bytes(NW) + 64);
  exclusive_scan_u32(popc, rank_base, NW, marked, t.bfs_scan.ptr, s);
  uint64_t* cmin = nullptr;
  uint64_t* cmax = nullptr;
  if (fr) {
    const uint64_t cap = static_cast<uint64_t>(naxes) * n + 64;
    t.bucket_max.reserve(2 * cap * sizeof(uint64_t));
    cmin = t.bucket_max.as<uint64_t>();
    cmax = cmin + cap;
    fill_count_kernel<uint64_t><<<256, 256, 0, s>>>(cmin, kEmpty, marked);
    CHUI_LAUNCH_CHECK("fill_count_kernel");
    fill_count_kernel<uint64_t><<<256, 256, 0, s>>>(cmax, 0ull, marked);
    CHUI_LAUNCH_CHECK("fill_count_kernel");
  } else {
    uint32_t hm = 0;
    t.readback.read(&hm, marked, sizeof(hm), s);
    if (size_bound) t.bucket_max.reserve(2 * (static_cast<uint64_t>(naxes) * n + 64) * sizeof(uint64_t));
    t.bucket_max.reserve(2 * std::max<uint64_t>(hm, 1) * sizeof(uint64_t));
    cmin = t.bucket_max.as<uint64_t>();
    cmax = cmin + std::max<uint64_t>(hm, 1);
    fill_u64(cmin, kEmpty, hm, s);
    fill_u64(cmax, 0ull, hm, s);
  }
  marked_minmax_kernel<<<pgrid, 256, 0, s>>>(p, n, a, mark, rank_base, cmin, cmax, fr);
  CHUI_LAUNCH_CHECK("marked_minmax_kernel");
  ra.cmin = cmin;
  ra.cmax = cmax;
  runs_kernel<2><<<rgrid, kGapThreads, 0, s>>>(a, ra);
  CHUI_LAUNCH_CHECK("runs_kernel<2>");
  gap_final_kernel<<<naxes, kGapThreads, 0, s>>>(best, nblk, dres);
  CHUI_LAUNCH_CHECK("gap_final_kernel");
  if (fr) return dres;
  UnwrapResult h{};
  t.readback.read(&h, dres, sizeof(UnwrapResult), s);
  for (int ax = 0; ax < naxes; ++ax) {
    res.anchor[ax] = h.anchor[ax];
    res.umin[ax] = h.umin[ax];
    res.umax[ax] = h.umax[ax];
  }
  return dres;
}

void launch_records(Tree& t, bool fp64, bool mixed, hipStream_t s) {
  const std::size_t total = t.node_count;
  const std::size_t nsz = t.sized_for_capture ? std::max(total, t.bfs_capacity) : total;
  if (fp64) {
    t.geo64.reserve(nsz * sizeof(NodeGeo64));
    t.mom64.reserve(nsz * sizeof(NodeMom64));
    t.quad64.reserve(nsz * sizeof(NodeQuad64));
    t.child_index.reserve(nsz * 8 * sizeof(uint32_t));
  }
  RecordArgs ra{};
  ra.total = static_cast<uint32_t>(total);
  ra.fp64 = fp64 ? 1 : 0;
  ra.mixed = mixed ? 1 : 0;
  ra.cubic = t.cubic ? 1 : 0;
  for (int a = 0; a < 3; ++a) ra.fscale[a] = t.fscale[a];
  records_kernel<<<grid_for(total, 256), 256, 0, s>>>(
      ra, t.bfs_pre.as<uint32_t>(), t.bfs_begin.as<uint32_t>(), t.bfs_count.as<uint32_t>(),
      t.bfs_nchild.as<uint32_t>(), t.bfs_first_child.as<uint32_t>(), t.bfs_meta.as<uint32_t>(),
      t.bfs_subtree.as<uint32_t>(), t.bfs_geo.as<double4>(), t.bfs_mom.as<double4>(), t.bfs_quad.as<double>(),
      t.link.as<NodeLink>(), fp64 ? t.geo64.as<NodeGeo64>() : nullptr, fp64 ? t.mom64.as<NodeMom64>() : nullptr,
      fp64 ? t.quad64.as<NodeQuad64>() : nullptr, fp64 ? t.child_index.as<uint32_t>() : nullptr,
      mixed ? t.geo32.as<NodeGeo32>() : nullptr, mixed ? t.mom32.as<NodeMom32>() : nullptr,
      mixed ? t.quad32.as<NodeQuad32>() : nullptr, mixed ? t.bfs_box.as<double>() : nullptr,
      mixed ? t.box32.as<NodeBox32>() : nullptr, device_node_count(t));
  CHUI_LAUNCH_CHECK("records_kernel");
  if (fp64) t.records64 = true;
}

}  // namespace

void ensure_records64(Tree& t, hipStream_t s) {
  if (t.records64 || t.node_count == 0) return;
  launch_records(t, true, false, s);
}

// ---------------------------------------------------------------------------
void build_tree(Tree& t, const double* x, const double* y, const double* z, const double* m, const TreeParams& P,
                double softening, hipStream_t s) {
  (void)softening;
  const std::size_t n = P.n;
  if (n > 0xFFFFFFFEull) throw std::overflow_error("Tree gravity source count exceeds the 32-bit tree-index contract");
  const bool safe = P.capture_safe;
  if (safe) {
    if (!P.fault) throw std::logic_error("capture-safe tree builds report through a fault word");
    if (n == 0 || n != t.prepared_n || !t.sized_for_capture || t.safe_levels <= 0 || t.bfs_capacity == 0)
      throw std::runtime_error(
          "capture-safe solve on a context not prepared for these sources (run an ordinary solve of the same "
          "source count first)");
  }
  t.n = n;
  t.node_count = 0;
  t.leaf_count = 0;
  t.counts_on_device = false;
  if (!safe) {
    t.max_depth = 0;
    t.level_start.clear();
  }
  Axes3 p{{x, y, z}};
  Frame* fr = nullptr;
  UnwrapArgs ua{};
  for (int a = 0; a < 3; ++a) ua.L[a] = P.L[a];
  if (safe) {
    // statistics -> canonical flags and checks on the device
    t.frame.reserve(sizeof(Frame));
    fr = t.frame.as<Frame>();
    const unsigned nb = std::min<unsigned>(grid_for(n, 256), 1024u);
    t.bounds.reserve((1 + nb) * sizeof(SourceStats));
    SourceStats* st = t.bounds.as<SourceStats>();
    source_stats_kernel<<<nb, 256, 0, s>>>(p, m, n, st + 1);
    CHUI_LAUNCH_CHECK("source_stats_kernel");
    stats_final_kernel<<<1, 256, 0, s>>>(st + 1, nb, m, st);
    CHUI_LAUNCH_CHECK("stats_final_kernel");
    FrameCheck fc{};
    for (int a = 0; a < 3; ++a) fc.L[a] = P.L[a];
    fc.expect_uniform = t.mass_uniform ? 1 : 0;
    fc.expect_mass = t.uniform_mass;
    frame_stats_kernel<<<1, 64, 0, s>>>(st, m, fc, fr, P.fault);
    CHUI_LAUNCH_CHECK("frame_stats_kernel");
    int unused[3] = {0, 0, 0};
    UnwrapResult* dres = unwrap_axes(t, p, 3, n, P.L, unused, t.unwrap, s, fr);
    frame_root_kernel<<<1, 64, 0, s>>>(dres, fr);
    CHUI_LAUNCH_CHECK("frame_root_kernel");
  } else {
    const SourceStats st = compute_stats(t, p, m, n, s);
    if (st.bad_coord) throw std::invalid_argument("Periodic TreePM tree geometry requires finite source coordinates");
    if (st.bad_mass)
      throw std::invalid_argument("Tree gravity requires finite coordinates and finite non-negative masses");
    if (n == 0) return;
    t.mass_uniform = st.mass_nonuniform == 0;
    t.uniform_mass = st.mass0;
    int canonical[3];
    for (int a = 0; a < 3; ++a) {
      const double lo = host_order_key_inv(st.lo[a]);
      const double hi = host_order_key_inv(st.hi[a]);
      canonical[a] = (lo >= 0.0 && hi < P.L[a]) ? 1 : 0;  // tree_pm_coupling.cpp:3185-3195
    }
    unwrap_axes(t, p, 3, n, P.L, canonical, t.unwrap, s, nullptr, P.size_for_capture);
    if (P.size_for_capture) t.frame.reserve(sizeof(Frame));
    const UnwrapResult& U = t.unwrap;
    // Root cube (tree_gravity.cpp:443-452) from the unwrapped bounds.
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, U.umax[a] - U.umin[a]);
    const double max_extent = std::max(ext, 1.0e-10);
    for (int a = 0; a < 3; ++a) t.root_centre[a] = 0.5 * (U.umin[a] + U.umax[a]);
    t.root_half = 0.5 * max_extent * (1.0 + 1.0e-8);
    for (int a = 0; a < 3; ++a) {
      ua.canonical[a] = U.canonical[a];
      ua.anchor[a] = U.anchor[a];
    }
  }

  // Path keys + stable sort.
  t.keys.reserve(n * sizeof(uint64_t));
  t.keys_alt.reserve(n * sizeof(uint64_t));
  t.order.reserve(n * sizeof(uint32_t));
  t.vals_alt.reserve(n * sizeof(uint32_t));
  t.sort_scratch.reserve(radix_sort_scratch_bytes(n));
  PathArgs pa{};
  pa.u = ua;
  for (int a = 0; a < 3; ++a) pa.c0[a] = t.root_centre[a];
  pa.h0 = t.root_half;
  uint64_t* keys = t.keys.as<uint64_t>();
  uint32_t* order = t.order.as<uint32_t>();
  t.pk.reserve(n * sizeof(double4));
  t.pk_in.reserve(n * sizeof(double4));
  path_key_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, m, n, pa, keys, order, t.pk_in.as<double4>(), fr);
  CHUI_LAUNCH_CHECK("path_key_kernel");
  // Sort by the bits of levels 1..16 only (6 digit passes instead of 8). This
  // is the reference's partition order for every node up to depth 16, and the
  // order inside a leaf is rebuilt below by (Morton key, row), so it suffices
  // unless some depth-16 node still splits; the discovery then restarts after
  // sorting by all 21 levels. A capture-safe build cannot restart: it sorts by
  // all 21 levels at once (the same partition, and the same order after the
  // in-leaf sort).
  constexpr int kPartialSortLevels = 16;
  constexpr int kPartialSortLow = 3 * (kPathLevels - kPartialSortLevels);
  bool partial = !safe;
  radix_sort_pairs(keys, order, t.keys_alt.as<uint64_t>(), t.vals_alt.as<uint32_t>(), n,
                   partial ? kPartialSortLow : 0, 3 * kPathLevels, t.sort_scratch.ptr, s);

  // Level-synchronous discovery. Every level is issued from the host without a
  // read-back (the kernels read each level's range from lv); the level starts
  // and flags come back once. A node capacity too small for the tree, or a
  // depth-16 node that splits while only levels 1..16 are sorted, restarts it.
  // A capture-safe build issues the prepared depth + 4 levels into the
  // prepared capacity and reads nothing back; a tree that does not fit sets a
  // fault bit.
  const std::size_t initial_cap = std::max<std::size_t>(1024, n / 4 + 64);
  auto ensure = [&](std::size_t need, std::size_t used) {
    if (need <= t.bfs_capacity && t.bfs_begin.ptr) return;
    std::size_t nc = std::max({need, 2 * t.bfs_capacity, initial_cap});
    grow_copy<uint32_t>(t.bfs_begin, used, nc, s);
    grow_copy<uint32_t>(t.bfs_count, used, nc, s);
    grow_copy<uint32_t>(t.bfs_first_child, used, nc, s);
    grow_copy<uint32_t>(t.bfs_nchild, used, nc, s);
    grow_copy<uint32_t>(t.bfs_meta, used, nc, s);
    grow_copy<double4>(t.bfs_geo, used, nc, s);
    t.bfs_capacity = nc;
  };
  if (!safe) ensure(initial_cap, 0);
  t.level_dev.reserve(kLvWords * sizeof(uint32_t));
  t.level_scratch.reserve(kLevelBlocks * sizeof(uint32_t));
  uint32_t* lv = t.level_dev.as<uint32_t>();
  uint32_t* bsum = t.level_scratch.as<uint32_t>();
  int nlevels = 0;       // BFS levels to run the per-level passes over
  std::size_t total = 0; // node count, or the capacity for a capture-safe build
  for (;;) {
    const std::size_t cap = std::min<std::size_t>(t.bfs_capacity, 0xFFFFFFFFull);
    t.level_bounds.reserve(8ull * cap * sizeof(uint32_t));
    root_init_kernel<<<1, 1, 0, s>>>(t.bfs_begin.as<uint32_t>(), t.bfs_count.as<uint32_t>(),
                                     t.bfs_meta.as<uint32_t>(), t.bfs_geo.as<double4>(), static_cast<uint32_t>(n),
                                     t.root_centre[0], t.root_centre[1], t.root_centre[2], t.root_half, fr);
    CHUI_LAUNCH_CHECK("root_init_kernel");
    level_dev_init_kernel<<<1, 64, 0, s>>>(lv);
    CHUI_LAUNCH_CHECK("level_dev_init_kernel");
    auto issue_levels = [&](int d0, int d1) {
      for (int d = d0; d < d1; ++d) {
        const LevelDevArgs la{static_cast<uint32_t>(d), P.leaf_size, static_cast<uint32_t>(cap),
                              static_cast<uint32_t>(partial ? kPartialSortLevels : kMaxLevels), level_blocks(d, n)};
        level_count_dev_kernel<<<la.nblocks, 256, 0, s>>>(
            la, lv, t.bfs_begin.as<uint32_t>(), t.bfs_count.as<uint32_t>(), t.bfs_meta.as<uint32_t>(),
            t.bfs_geo.as<double4>(), keys, t.bfs_nchild.as<uint32_t>(), t.level_bounds.as<uint32_t>(), bsum);
        CHUI_LAUNCH_CHECK("level_count_dev_kernel");
        level_offsets_dev_kernel<<<1, kLevelBlocks / 2, 0, s>>>(la, lv, bsum);
        CHUI_LAUNCH_CHECK("level_offsets_dev_kernel");
        level_emit_dev_kernel<<<la.nblocks, 256, 0, s>>>(
            la, lv, t.bfs_begin.as<uint32_t>(), t.bfs_count.as<uint32_t>(), t.bfs_meta.as<uint32_t>(),
            t.bfs_geo.as<double4>(), t.level_bounds.as<uint32_t>(), t.bfs_nchild.as<uint32_t>(), bsum,
            t.bfs_first_child.as<uint32_t>());
        CHUI_LAUNCH_CHECK("level_emit_dev_kernel");
      }
    };
    if (safe) {
      nlevels = std::min(kMaxLevels, t.safe_levels);
      issue_levels(0, nlevels);
      level_check_kernel<<<1, 64, 0, s>>>(lv, nlevels, P.fault);
      CHUI_LAUNCH_CHECK("level_check_kernel");
      total = cap;
      break;
    }
    // Levels up to the last build's depth + 2 first (an empty level costs
    // three launches of about 4 us each); the read-back that follows anyway
    // tells whether the tree went deeper, and only then are the rest issued.
    int issued = t.level_hint > 0 ? std::min(kMaxLevels, t.level_hint) : kMaxLevels;
    issue_levels(0, issued);
    uint32_t h[kLvWords];
    for (;;) {
      t.readback.read(h, lv, sizeof(h), s);
      // level `issued` is [h[issued], h[issued + 1]), written by the last pass
      if (issued < kMaxLevels && !h[kLvOverflow] && !(partial && h[kLvDeepSplit]) && h[issued + 1] > h[issued]) {
        issue_levels(issued, kMaxLevels);
        issued = kMaxLevels;
        continue;
      }
      break;
    }
    if (partial && h[kLvDeepSplit]) {
      // A depth-16 node splits (whether it does depends on its count, not on
      // the unsorted low digits). Sort by every level and start over; the sort
      // is stable, so equal keys keep their row order.
      radix_sort_pairs(keys, order, t.keys_alt.as<uint64_t>(), t.vals_alt.as<uint32_t>(), n, 0, 3 * kPathLevels,
                       t.sort_scratch.ptr, s);
      partial = false;
      continue;
    }
    if (h[kLvOverflow]) {
      ensure(2 * t.bfs_capacity, 0);
      continue;
    }
    if (h[kLvTooDeep]) {
      throw UnsupportedError(
          "octree deeper than 21 levels (more than max_leaf_size coincident or near-coincident particles) "
          "is not built by the GPU path yet");
    }
    t.level_start = {h[0], h[1]};
    for (int d = 0; d + 1 < kMaxLevels && h[d + 2] > h[d + 1]; ++d) t.level_start.push_back(h[d + 2]);
    t.max_depth = static_cast<uint32_t>(t.level_start.size() - 2);
    t.leaf_count = h[kLvLeaves];
    t.level_hint = static_cast<int>(t.max_depth) + 3;
    t.safe_levels = std::min(kMaxLevels, static_cast<int>(t.max_depth) + 5);
    t.prepared_n = n;
    t.sized_for_capture = P.size_for_capture;
    nlevels = static_cast<int>(t.level_start.size()) - 1;
    total = t.level_start.back();
    break;
  }
  t.node_count = total;
  t.counts_on_device = safe;
  const uint32_t* total_dev = device_node_count(t);
  // node arrays: the node count, or the capacity when capture-safe builds follow
  const std::size_t nsz = P.size_for_capture ? std::max(total, t.bfs_capacity) : total;
  // per-level passes: one launch per level; a capture-safe build reads each
  // level's range on the device and strides a fixed grid over it
  auto range_of = [&](int d) {
    LevelRange r{};
    if (safe) {
      r.lv = lv;
      r.d = d;
    } else {
      r.lb = static_cast<uint32_t>(t.level_start[d]);
      r.cnt = static_cast<uint32_t>(t.level_start[d + 1] - t.level_start[d]);
    }
    return r;
  };
  auto level_grid = [&](int d, unsigned block) {
    if (safe) return std::min<unsigned>(grid_for(total, block), 1024u);
    return grid_for(t.level_start[d + 1] - t.level_start[d], block);
  };
  t.bfs_subtree.reserve(nsz * sizeof(uint32_t));
  t.bfs_pre.reserve(nsz * sizeof(uint32_t));
  t.bfs_mom.reserve(nsz * sizeof(double4));
  t.bfs_quad.reserve(nsz * 8 * sizeof(double));
  uint32_t* nchild = t.bfs_nchild.as<uint32_t>();
  uint32_t* fchild = t.bfs_first_child.as<uint32_t>();
  for (int d = nlevels - 1; d >= 0; --d) {
    subtree_kernel<<<level_grid(d, 256), 256, 0, s>>>(range_of(d), nchild, fchild, t.bfs_subtree.as<uint32_t>());
    CHUI_LAUNCH_CHECK("subtree_kernel");
  }
  fill_u32(t.bfs_pre.as<uint32_t>(), 0u, 1, s);
  for (int d = 0; d < nlevels; ++d) {
    preorder_kernel<<<level_grid(d, 256), 256, 0, s>>>(range_of(d), nchild, fchild, t.bfs_subtree.as<uint32_t>(),
                                                       t.bfs_pre.as<uint32_t>());
    CHUI_LAUNCH_CHECK("preorder_kernel");
  }
  // In-leaf Morton order (tree_ordering.cpp:151 extent with the 1e-12 floor).
  MortonArgs ma{};
  if (!safe) {
    const UnwrapResult& U = t.unwrap;
    double mext = 0.0;
    for (int a = 0; a < 3; ++a) {
      ma.lo[a] = U.umin[a];
      mext = std::max(mext, U.umax[a] - U.umin[a]);
    }
    ma.extent = std::max(mext, 1.0e-12);
  }
  uint64_t* mk = t.keys_alt.as<uint64_t>();
  double4* pk = t.pk.as<double4>();
  gather_sources_kernel<<<grid_for(n, 256), 256, 0, s>>>(t.pk_in.as<double4>(), n, order, pk);
  CHUI_LAUNCH_CHECK("gather_sources_kernel");
  leaf_morton_kernel<<<grid_for(n, 256), 256, 0, s>>>(pk, n, ua, ma, mk, fr);
  CHUI_LAUNCH_CHECK("leaf_morton_kernel");
  leaf_sort_kernel<<<grid_for(total, 256), 256, 0, s>>>(static_cast<uint32_t>(total), total_dev, nchild,
                                                        t.bfs_begin.as<uint32_t>(), t.bfs_count.as<uint32_t>(), mk,
                                                        order, t.vals_alt.as<uint32_t>(), pk);
  CHUI_LAUNCH_CHECK("leaf_sort_kernel");
  // Multipoles and tight particle boxes bottom-up.
  t.bfs_box.reserve(nsz * 6 * sizeof(double));
  for (int d = nlevels - 1; d >= 0; --d) {
    MomArgs ma2{};
    ma2.u = ua;
    ma2.r = range_of(d);
    ma2.quad = P.quadrupole ? 1 : 0;
    multipole_kernel<<<level_grid(d, 128), 128, 0, s>>>(
        ma2, pk, t.bfs_begin.as<uint32_t>(), t.bfs_count.as<uint32_t>(), nchild, fchild,
        t.bfs_geo.as<double4>(), t.bfs_mom.as<double4>(), t.bfs_quad.as<double>(), t.bfs_box.as<double>(), fr);
    CHUI_LAUNCH_CHECK("multipole_kernel");
  }
  // Preorder records: links and, for mixed builds, the fixed-point records;
  // the FP64 records for FP64 builds (or later, on demand).
  t.link.reserve(nsz * sizeof(NodeLink));
  if (P.mixed) {
    t.geo32.reserve(nsz * sizeof(NodeGeo32));
    t.mom32.reserve(nsz * sizeof(NodeMom32));
    t.quad32.reserve(nsz * sizeof(NodeQuad32));
    t.box32.reserve(nsz * sizeof(NodeBox32));
  }
  t.records64 = false;
  t.cubic = (P.L[0] == P.L[1] && P.L[0] == P.L[2]);
  for (int a = 0; a < 3; ++a) t.fscale[a] = 4294967296.0 / P.L[a];
  launch_records(t, !P.mixed, P.mixed, s);
}

// ---- mesh-deposit units ---------------------------------------------------------
// The first cell a particle at unwrapped coordinate u writes (CIC: its cell;
// TSC: the cell below the nearest one); the deposit kernels wrap it into the
// mesh, so the tile of a box is [first(lo), first(hi) + span - 1] per axis.
__device__ __forceinline__ long deposit_first_cell(double u, double inv_h, int scheme) {
  const double g = u * inv_h;
  return scheme == 0 ? static_cast<long>(floor(g)) : static_cast<long>(floor(g + 0.5)) - 1;
}

__device__ bool deposit_tile(const double* box, const DepositTiling& d, int32_t lo[3], int32_t dims[3]) {
  const long span = d.scheme == 0 ? 2 : 3;
  long vol = 1;
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    const long l = deposit_first_cell(box[a], d.inv_h[a], d.scheme);
    const long ext = deposit_first_cell(box[3 + a], d.inv_h[a], d.scheme) + span - l;
    long w = l % d.n[a];
    if (w < 0) w += d.n[a];
    lo[a] = static_cast<int32_t>(w);
    ok = ok && ext <= d.n[a];
    dims[a] = ok ? static_cast<int32_t>(ext) : 0;
    vol *= ok ? ext : 1;
  }
  return ok && vol <= static_cast<long>(d.max_cells);
}

// Units form a cut of the tree: a node is one when it fits (or is a leaf) and
// its parent does not, so every particle lies in exactly one unit. The parent
// decides for its children; flags start at zero.
__global__ __launch_bounds__(256) void deposit_flag_kernel(uint32_t total_nodes, const uint32_t* __restrict__ total_dev,
                                                           const uint32_t* __restrict__ nchild,
                                                           const uint32_t* __restrict__ first_child,
                                                           const double* __restrict__ box, DepositTiling d,
                                                           uint32_t* __restrict__ flag) {
  const uint32_t total = total_dev ? *total_dev : total_nodes;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
    int32_t lo[3], dims[3];
    const uint32_t c = nchild[k];
    const bool fits = deposit_tile(box + 6ull * k, d, lo, dims);
    if (k == 0) flag[0] = (fits || c == 0) ? 1u : 0u;
    if (fits || c == 0) continue;
    const uint32_t fc = first_child[k];
    for (uint32_t j = 0; j < c; ++j) {
      const uint32_t ch = fc + j;
      flag[ch] = (nchild[ch] == 0 || deposit_tile(box + 6ull * ch, d, lo, dims)) ? 1u : 0u;
    }
  }
}

__global__ __launch_bounds__(256) void deposit_emit_kernel(uint32_t total_nodes, const uint32_t* __restrict__ total_dev,
                                                           const uint32_t* __restrict__ flag,
                                                           const uint32_t* __restrict__ index,
                                                           const uint32_t* __restrict__ nb,
                                                           const uint32_t* __restrict__ nc,
                                                           const double* __restrict__ box, DepositTiling d,
                                                           DepositUnit* __restrict__ units,
                                                           uint32_t* __restrict__ blocks) {
  const uint32_t total = total_dev ? *total_dev : total_nodes;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
    if (!flag[k]) continue;
    DepositUnit u;
    u.begin = nb[k];
    u.count = nc[k];
    if (!deposit_tile(box + 6ull * k, d, u.lo, u.dims)) u.dims[0] = 0;  // leaf spanning too many cells
    const uint32_t i = index[k];
    units[i] = u;
    blocks[i] = (u.count + d.chunk - 1) / d.chunk;
  }
}

DepositUnits build_deposit_units(Tree& t, const DepositTiling& d, hipStream_t s) {
  DepositUnits r;
  const std::size_t total = t.node_count;
  if (total == 0 || t.n == 0) return r;
  const std::size_t nsz = t.sized_for_capture ? std::max(total, t.bfs_capacity) : total;
  t.dep_flag.reserve(nsz * sizeof(uint32_t));
  t.dep_index.reserve(nsz * sizeof(uint32_t));
  t.dep_units.reserve(nsz * sizeof(DepositUnit));
  t.dep_blocks.reserve((nsz + 1) * sizeof(uint32_t));
  t.dep_scan.reserve(scan_scratch_bytes(nsz) + 64);
  t.dep_total.reserve(2 * sizeof(uint32_t));
  uint32_t* flag = t.dep_flag.as<uint32_t>();
  uint32_t* totals = t.dep_total.as<uint32_t>();  // [0] units, [1] blocks
  const uint32_t* total_dev = device_node_count(t);
  fill_u32(flag, 0u, total, s);
  deposit_flag_kernel<<<grid_for(total, 256), 256, 0, s>>>(static_cast<uint32_t>(total), total_dev,
                                                           t.bfs_nchild.as<uint32_t>(),
                                                           t.bfs_first_child.as<uint32_t>(), t.bfs_box.as<double>(),
                                                           d, flag);
  CHUI_LAUNCH_CHECK("deposit_flag_kernel");
  exclusive_scan_u32(flag, t.dep_index.as<uint32_t>(), total, totals, t.dep_scan.ptr, s);
  if (total_dev) fill_u32(t.dep_blocks.as<uint32_t>(), 0u, total, s);  // blocks past the last unit stay 0
  deposit_emit_kernel<<<grid_for(total, 256), 256, 0, s>>>(
      static_cast<uint32_t>(total), total_dev, flag, t.dep_index.as<uint32_t>(), t.bfs_begin.as<uint32_t>(),
      t.bfs_count.as<uint32_t>(), t.bfs_box.as<double>(), d, t.dep_units.as<DepositUnit>(),
      t.dep_blocks.as<uint32_t>());
  CHUI_LAUNCH_CHECK("deposit_emit_kernel");
  r.units = t.dep_units.as<DepositUnit>();
  r.first_block = t.dep_blocks.as<uint32_t>();
  if (total_dev) {
    // capture-safe: scan the blocks of every capacity slot (zero past the last
    // unit) and leave both counts on the device; the deposit's grid is a fixed
    // size that strides over the real block count
    exclusive_scan_u32(t.dep_blocks.as<uint32_t>(), t.dep_blocks.as<uint32_t>(), total, totals + 1, t.dep_scan.ptr,
                       s);
    r.dev_counts = totals;
    r.count = 0;
    r.blocks = static_cast<uint32_t>(std::min<std::size_t>(total + t.n / d.chunk + 1, kDepositSafeGrid));
    return r;
  }
  uint32_t host[2] = {0, 0};
  t.readback.read(host, totals, sizeof(uint32_t), s);
  exclusive_scan_u32(t.dep_blocks.as<uint32_t>(), t.dep_blocks.as<uint32_t>(), host[0], totals + 1, t.dep_scan.ptr, s);
  t.readback.read(host + 1, totals + 1, sizeof(uint32_t), s);
  r.count = host[0];
  r.blocks = host[1];
  return r;
}

const uint32_t* device_node_count(const Tree& t) {
  return t.counts_on_device ? t.level_dev.as<uint32_t>() + kLvTotal : nullptr;
}

void build_pos_of_row(Tree& t, hipStream_t s) {
  t.pos_of_row.reserve(std::max<std::size_t>(t.n, 1) * sizeof(uint32_t));
  if (t.n == 0) return;
  pos_of_row_kernel<<<grid_for(t.n, 256), 256, 0, s>>>(t.order.as<uint32_t>(), t.n, t.pos_of_row.as<uint32_t>());
  CHUI_LAUNCH_CHECK("pos_of_row_kernel");
}

void morton_order_device(Tree& t, const double* x, const double* y, const double* z, std::size_t n,
                         uint64_t* keys_sorted, uint32_t* order, hipStream_t s) {
  Axes3 p{{x, y, z}};
  const SourceStats st = compute_stats(t, p, nullptr, n, s);
  if (st.bad_coord) throw std::invalid_argument("Tree bounds requires finite particle coordinates");
  if (n == 0) return;
  MortonArgs ma{};
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    ma.lo[a] = host_order_key_inv(st.lo[a]);
    ext = std::max(ext, host_order_key_inv(st.hi[a]) - ma.lo[a]);
  }
  ma.extent = std::max(ext, 1.0e-12);
  t.keys_alt.reserve(n * sizeof(uint64_t));
  t.vals_alt.reserve(n * sizeof(uint32_t));
  t.sort_scratch.reserve(radix_sort_scratch_bytes(n));
  morton_keys_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, n, ma, keys_sorted, order);
  CHUI_LAUNCH_CHECK("morton_keys_kernel");
  radix_sort_pairs(keys_sorted, order, t.keys_alt.as<uint64_t>(), t.vals_alt.as<uint32_t>(), n, 0, 63,
                   t.sort_scratch.ptr, s);
}

double unwrap_axis_device(Tree& t, const double* in, std::size_t n, double L, double* out, hipStream_t s) {
  if (!(std::isfinite(L) && L > 0.0))
    throw std::invalid_argument("Periodic TreePM tree geometry requires finite positive axis lengths");
  Axes3 p{{in, in, in}};
  const SourceStats st = compute_stats(t, p, nullptr, n, s);
  if (st.bad_coord) throw std::invalid_argument("Periodic TreePM tree geometry requires finite source coordinates");
  if (n == 0) return 0.0;
  const double lo = host_order_key_inv(st.lo[0]), hi = host_order_key_inv(st.hi[0]);
  int canonical[3] = {(lo >= 0.0 && hi < L) ? 1 : 0, 0, 0};
  double Ls[3] = {L, L, L};
  UnwrapResult r{};
  unwrap_axes(t, p, 1, n, Ls, canonical, r, s);
  unwrap_apply_kernel<<<grid_for(n, 256), 256, 0, s>>>(in, n, L, canonical[0], r.anchor[0], out);
  CHUI_LAUNCH_CHECK("unwrap_apply_kernel");
  return r.anchor[0];
}

}  // namespace chui

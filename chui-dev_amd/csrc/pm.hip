// pm.hip — PM long-range solve: CIC/TSC deposit, hand-written batched
// Stockham FFT passes (radix 8/4/2) through LDS, Green's function with the Gaussian split filter and
// spectral gradient (pm_solver.cpp:925-951, :2398-2705, :4016-4061, :4581-4940).
//
// Layout: complex spectrum C[ix][iy][kz], kz in [0, nz/2] (row stride nzc);
// real meshes use the same storage with padded rows of 2*nzc reals, so the
// r2c and the three c2r transforms run in place. Forward passes: z (r2c, two
// real rows packed into one complex FFT), y, x (+ Green's function). Inverse:
// x twice (with/without -i kx), y three times, z (c2r) three times; the
// constant 1/(cell volume * M) of the reference's density normalisation and
// inverse FFT is folded into the Green's function.
#include "pm.hpp"
#include "primitives.hpp"

#include <mutex>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace chui {
namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kFftThreads = 256;
// Butterflies a thread keeps in registers per pass (float2: 4 x 8 points,
// double2: 2 x 8 points); blocks are sized so that lines * n / R fits.
template <class T> constexpr int fft_max_butterflies() { return sizeof(T) == 4 ? 4 : 2; }
// Global loads a thread issues before its LDS stores. A block's whole tile is
// 16 lines of 512 points (FP32) or 8 lines (FP64) over 256 threads, so all of
// a thread's loads are in flight at once: about 64 KB per CU with two blocks,
// which HBM3E needs to approach its rate (8 loads per thread reached 2.3 TB/s).
template <class T> constexpr int fft_load_batch() { return sizeof(T) == 4 ? 32 : 16; }
// The z passes keep two values per load (real row pairs), so half of that.
template <class T> constexpr int fft_z_load_batch() { return sizeof(T) == 4 ? 16 : 8; }

// Smallest radix of the pass plan of lds_fft for a line of 2^lg points.
inline int fft_min_radix(int lg) {
  const int n8 = (lg % 3 == 1 && lg >= 4) ? (lg - 4) / 3 : lg / 3;
  const int tail = lg - 3 * n8;
  return tail == 0 ? 8 : (tail == 1 ? 2 : 4);
}

template <class T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };

template <class V> __device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <class V> __device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class V> __device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <class V, class T> __device__ __forceinline__ V mk(T x, T y) { V r; r.x = x; r.y = y; return r; }

// ---- LDS line FFTs: Stockham autosort passes of radix 8, 4 and 2 ---------------
// Pass with radix R after sub-transforms of length Ns: butterfly j reads
// x[j + r n/R], multiplies by exp(-+2 pi i (j mod Ns) r / (Ns R)), applies the
// length-R DFT and writes x[(j / Ns) Ns R + (j mod Ns) + r Ns]. Natural order
// in and out (no bit reversal); log2(n) = 9 takes three passes of radix 8.
// Each thread keeps its butterflies in registers across the barrier between
// the read and the write phase, so one LDS buffer serves every pass.
template <class V, bool INV> __device__ __forceinline__ V rot_mi(V a) {  // * (-i) forward, * (+i) inverse
  V r;
  if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

template <class V, bool INV> __device__ __forceinline__ void dft4(V& a0, V& a1, V& a2, V& a3) {
  const V t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = rot_mi<V, INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <class V, int R, bool INV> __device__ __forceinline__ void dft(V* v) {
  if constexpr (R == 2) {
    const V t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
  } else if constexpr (R == 4) {
    dft4<V, INV>(v[0], v[1], v[2], v[3]);
  } else {
    V e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    V o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<V, INV>(e0, e1, e2, e3);
    dft4<V, INV>(o0, o1, o2, o3);
    using T = decltype(o1.x);
    const T h = T(0.70710678118654752440084436210484904);
    // o_k *= W8^k, W8 = exp(-+2 pi i / 8)
    V w1, w3;
    if (INV) {
      w1.x = (o1.x - o1.y) * h; w1.y = (o1.x + o1.y) * h;
      w3.x = -(o3.x + o3.y) * h; w3.y = (o3.x - o3.y) * h;
    } else {
      w1.x = (o1.x + o1.y) * h; w1.y = (o1.y - o1.x) * h;
      w3.x = (o3.y - o3.x) * h; w3.y = -(o3.x + o3.y) * h;
    }
    const V w2 = rot_mi<V, INV>(o2);
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1); v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2); v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3); v[7] = csub(e3, w3);
  }
}

// exp(-+2 pi i m / n) for 0 <= m < n from the half table tw[k] = exp(-2 pi i k / n).
template <class V, bool INV> __device__ __forceinline__ V twiddle(const V* tw, int m, int halfn) {
  V w = tw[m < halfn ? m : m - halfn];
  if (m >= halfn) { w.x = -w.x; w.y = -w.y; }
  if (INV) w.y = -w.y;
  return w;
}

// LDS slot of point i of a power-of-two line: one padding slot after every 8
// points. The radix-8 passes store each butterfly's outputs 8 points (first
// pass) or 64 points (second pass) apart; unpadded, the 64 lanes of such a
// store fall into 4 or 8 of the 64 LDS banks (16- and 8-way conflicts), padded
// into 32 (2-way). Lines of other lengths are not padded.
__host__ __device__ __forceinline__ int lds_slot(int i, bool pad) { return pad ? i + (i >> 3) : i; }
// Slots between consecutive LDS lines: odd, so that the lines of a block
// start in different banks.
__host__ __device__ __forceinline__ int lds_line_stride(int n, bool pad) { return pad ? n + (n >> 3) + 1 : n + 1; }

template <class T, int R, bool INV>
__device__ __attribute__((noinline)) void stockham_pass(typename C2<T>::type* X, int nlines, int lgn, int stride,
                                              const typename C2<T>::type* tw, int lgNs) {
  using V = typename C2<T>::type;
  constexpr int LR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  constexpr int MAXB = fft_max_butterflies<T>();  // the host sizes blocks so lines * n / R <= MAXB * kFftThreads
  const int lgq = lgn - LR;
  const int q = 1 << lgq;
  const int total = nlines << lgq;
  const int Ns = 1 << lgNs;
  const int halfn = 1 << (lgn - 1);
  const int sh = lgn - lgNs - LR;
  V v[MAXB][R];
  int base[MAXB], dst[MAXB];
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    const int b = threadIdx.x + u * kFftThreads;
    dst[u] = -1;
    base[u] = 0;
    if (b < total) {
      const int line = b >> lgq, j = b & (q - 1);
      const V* L = X + line * stride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u][r] = L[lds_slot(j + r * q, true)];
      const int jm = j & (Ns - 1);
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[u][r] = cmul(v[u][r], twiddle<V, INV>(tw, (jm * r) << sh, halfn));
      }
      dft<V, R, INV>(v[u]);
      base[u] = line * stride;
      dst[u] = ((j >> lgNs) << (lgNs + LR)) + jm;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    if (dst[u] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) X[base[u] + lds_slot(dst[u] + r * Ns, true)] = v[u][r];
    }
  }
  __syncthreads();
}

template <class T, bool INV>
__device__ void lds_fft_dir(typename C2<T>::type* X, int nlines, int lg, int stride, const typename C2<T>::type* tw) {
  // plan: radix-8 passes, then a tail of 4+4, 4 or 2 (log2 tail 4, 2 or 1)
  const int n8 = (lg % 3 == 1 && lg >= 4) ? (lg - 4) / 3 : lg / 3;
  const int tail = lg - 3 * n8;
#pragma nounroll
  for (int i = 0; i < n8; ++i) stockham_pass<T, 8, INV>(X, nlines, lg, stride, tw, 3 * i);
  const int done = 3 * n8;
  if (tail >= 2) stockham_pass<T, 4, INV>(X, nlines, lg, stride, tw, done);
  if (tail == 4) stockham_pass<T, 4, INV>(X, nlines, lg, stride, tw, done + 2);
  if (tail == 1) stockham_pass<T, 2, INV>(X, nlines, lg, stride, tw, done);
}

// ---- lines of n = 3^b 2^a points (meshes that are not powers of two) -----------
// The same Stockham pass with integer index arithmetic and a full-circle
// twiddle table tw[m] = exp(-2 pi i m / n), m < n. Radix-3 butterflies:
// y0 = a0 + a1 + a2, y1,2 = a0 - (a1 + a2)/2 -+ i (sqrt 3 / 2)(a1 - a2) (forward).
template <class V, bool INV> __device__ __forceinline__ void dft3(V* v) {
  using T = decltype(v[0].x);
  const T s3 = T(0.86602540378443864676372317075293618);
  const V sum = cadd(v[1], v[2]);
  const V t1 = mk<V, T>(v[0].x - T(0.5) * sum.x, v[0].y - T(0.5) * sum.y);
  const V d = rot_mi<V, INV>(csub(v[1], v[2]));
  const V t2 = mk<V, T>(s3 * d.x, s3 * d.y);
  v[0] = cadd(v[0], sum);
  v[1] = cadd(t1, t2);
  v[2] = csub(t1, t2);
}

template <class T, int R, bool INV>
__device__ __attribute__((noinline)) void stockham_pass_gen(typename C2<T>::type* X, int nlines, int n, int stride,
                                                  const typename C2<T>::type* tw, int Ns) {
  using V = typename C2<T>::type;
  constexpr int MAXB = fft_max_butterflies<T>();  // the host sizes blocks so lines * n / R <= MAXB * kFftThreads
  const int q = n / R;
  const int total = nlines * q;
  const int step = n / (Ns * R);  // twiddle index of exp(-2 pi i / (Ns R))
  V v[MAXB][R];
  int dst[MAXB];
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    const int b = threadIdx.x + u * kFftThreads;
    dst[u] = -1;
    if (b < total) {
      const int line = b / q, j = b - line * q;
      const V* L = X + line * stride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u][r] = L[j + r * q];
      const int jg = j / Ns;
      const int jm = j - jg * Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          V w = tw[jm * r * step];
          if (INV) w.y = -w.y;
          v[u][r] = cmul(v[u][r], w);
        }
      }
      if constexpr (R == 3) dft3<V, INV>(v[u]);
      else dft<V, R, INV>(v[u]);
      dst[u] = line * stride + jg * Ns * R + jm;
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < MAXB; ++u) {
    if (dst[u] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) X[dst[u] + r * Ns] = v[u][r];
    }
  }
  __syncthreads();
}

template <class T, bool INV>
__device__ void lds_fft_gen_dir(typename C2<T>::type* X, int nlines, int n, int stride,
                                const typename C2<T>::type* tw) {
  // plan: radix-3 passes for 3^b, then radix 8 / 4 / 2 for 2^a as in lds_fft_dir
  int Ns = 1, m = n;
#pragma nounroll
  while (m % 3 == 0) {
    stockham_pass_gen<T, 3, INV>(X, nlines, n, stride, tw, Ns);
    Ns *= 3;
    m /= 3;
  }
  int lg = 0;
  while ((1 << lg) < m) ++lg;
  const int n8 = (lg % 3 == 1 && lg >= 4) ? (lg - 4) / 3 : lg / 3;
  const int tail = lg - 3 * n8;
#pragma nounroll
  for (int i = 0; i < n8; ++i) {
    stockham_pass_gen<T, 8, INV>(X, nlines, n, stride, tw, Ns);
    Ns *= 8;
  }
  if (tail >= 2) {
    stockham_pass_gen<T, 4, INV>(X, nlines, n, stride, tw, Ns);
    Ns *= 4;
  }
  if (tail == 4) stockham_pass_gen<T, 4, INV>(X, nlines, n, stride, tw, Ns);
  if (tail == 1) stockham_pass_gen<T, 2, INV>(X, nlines, n, stride, tw, Ns);
}

// Smallest radix of the pass plan for a line of n points (lg < 0: not a power of two).
inline int fft_plan_min_radix(int n) {
  int m = n, r = 1 << 30;
  while (m % 3 == 0) {
    m /= 3;
    r = 3;
  }
  int lg = 0;
  while ((1 << lg) < m) ++lg;
  if (lg > 0) r = std::min(r, fft_min_radix(lg));
  return r;
}

// Twiddle entries a kernel keeps in LDS: half a circle for power-of-two lines,
// the full circle otherwise.
__host__ __device__ __forceinline__ int fft_tw_len(int n, int lg) { return lg >= 0 ? n / 2 : n; }

// In-place FFT of nlines LDS lines (natural order in and out); blockDim must
// be kFftThreads. lg = log2(n), or -1 when n = 3^b 2^a is not a power of two.
template <class T>
__device__ void lds_fft(typename C2<T>::type* X, int nlines, int n, int lg, int stride,
                        const typename C2<T>::type* tw, bool inverse) {
  if (lg < 0) {
    if (inverse) lds_fft_gen_dir<T, true>(X, nlines, n, stride, tw);
    else lds_fft_gen_dir<T, false>(X, nlines, n, stride, tw);
    return;
  }
  if (inverse) lds_fft_dir<T, true>(X, nlines, lg, stride, tw);
  else lds_fft_dir<T, false>(X, nlines, lg, stride, tw);
}

// idx -> (idx / n, idx % n) for n = 2^lg (lg >= 0) or any n.
__device__ __forceinline__ void split_index(int idx, int n, int lg, int& hi, int& lo) {
  if (lg >= 0) {
    hi = idx >> lg;
    lo = idx & (n - 1);
  } else {
    hi = idx / n;
    lo = idx - hi * n;
  }
}

// ---- deposit ------------------------------------------------------------------
__device__ __forceinline__ double wrap_position(double x, double L) {
  const double w = fmod(x, L);
  return w < 0.0 ? w + L : w;
}
__device__ __forceinline__ int wrap_index(long i, int n) {
  long w = i % n;
  if (w < 0) w += n;
  return static_cast<int>(w);
}

struct DepositArgs {
  int n[3];
  int lg[3];
  double inv_h[3];  // n / L
  double L[3];
  std::size_t row;
  int scheme;
};

// Stencil of one source: wrapped cell indices and weights per axis (2 cells
// for CIC, 3 for TSC), first the FP64 form (pm_solver.cpp assignment with
// periodic wrap), then the fixed-point form of the mixed path, whose top lg
// bits are the cell and whose remaining bits the position inside it.
__device__ __forceinline__ void stencil(const DepositArgs& a, const double4& sp, int idx[3][3], double w[3][3]) {
  const double g[3] = {wrap_position(sp.x, a.L[0]) * a.inv_h[0], wrap_position(sp.y, a.L[1]) * a.inv_h[1],
                       wrap_position(sp.z, a.L[2]) * a.inv_h[2]};
  for (int ax = 0; ax < 3; ++ax) {
    if (a.scheme == 0) {
      const double base = floor(g[ax]);
      const long i0 = static_cast<long>(base);
      const double t = g[ax] - base;
      idx[ax][0] = wrap_index(i0, a.n[ax]);
      idx[ax][1] = wrap_index(i0 + 1, a.n[ax]);
      w[ax][0] = 1.0 - t;
      w[ax][1] = t;
    } else {
      const double c = floor(g[ax] + 0.5);
      const long ic = static_cast<long>(c);
      const double d = g[ax] - c;
      idx[ax][0] = wrap_index(ic - 1, a.n[ax]);
      idx[ax][1] = wrap_index(ic, a.n[ax]);
      idx[ax][2] = wrap_index(ic + 1, a.n[ax]);
      w[ax][0] = 0.5 * (0.5 - d) * (0.5 - d);
      w[ax][1] = 0.75 - d * d;
      w[ax][2] = 0.5 * (0.5 + d) * (0.5 + d);
    }
  }
}
__device__ __forceinline__ double source_mass(const double4& sp) { return sp.w; }

// Index i in [-n, 2n) wrapped into [0, n).
__device__ __forceinline__ int wrap_once(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Fixed-point coordinate q (x = q L / 2^32) on a mesh of n cells: q n / 2^32 =
// cell + t exactly in 64-bit integers (for n = 2^lg: the top lg bits of q).
__device__ __forceinline__ void fixed_cell(uint32_t q, int n, int& cell, float& t) {
  const uint64_t p = static_cast<uint64_t>(q) * static_cast<uint32_t>(n);
  cell = static_cast<int>(p >> 32);
  t = static_cast<float>(static_cast<uint32_t>(p)) * 0x1p-32f;
}

__device__ __forceinline__ void stencil(const DepositArgs& a, const uint4& sp, int idx[3][3], float w[3][3]) {
  const uint32_t q[3] = {sp.x, sp.y, sp.z};
  for (int ax = 0; ax < 3; ++ax) {
    const int nn = a.n[ax];
    int cell;
    float t;
    fixed_cell(q[ax], nn, cell, t);
    if (a.scheme == 0) {
      idx[ax][0] = cell;
      idx[ax][1] = wrap_once(cell + 1, nn);
      w[ax][0] = 1.0f - t;
      w[ax][1] = t;
    } else {
      const int c = t >= 0.5f ? cell + 1 : cell;
      const float d = t >= 0.5f ? t - 1.0f : t;
      idx[ax][0] = wrap_once(c - 1, nn);
      idx[ax][1] = wrap_once(c, nn);
      idx[ax][2] = wrap_once(c + 1, nn);
      w[ax][0] = 0.5f * (0.5f - d) * (0.5f - d);
      w[ax][1] = 0.75f - d * d;
      w[ax][2] = 0.5f * (0.5f + d) * (0.5f + d);
    }
  }
}
__device__ __forceinline__ float source_mass(const uint4& sp) { return __uint_as_float(sp.w); }

// One block per chunk of a deposit unit. A tiled unit accumulates its chunk in
// an LDS block of cells and then adds the block to the mesh one z row at a
// time, so neighbouring lanes hit neighbouring cells and each atomic
// instruction touches one or two cache lines instead of up to 64. A source
// whose stencil leaves the block (only through rounding at the block's edge)
// and every source of an untiled unit go straight to the mesh.
template <class T, class S>
__global__ __launch_bounds__(256) void deposit_kernel(DepositArgs a, const S* __restrict__ src, int uniform, T um,
                                                      DepositUnits du, T* __restrict__ rho) {
  __shared__ T tile[kDepositTileBytes / sizeof(T)];
  __shared__ uint32_t s_unit;
  // Block vb of the unit list; the grid covers every block (one iteration per
  // workgroup) or, for a capture-safe solve, strides over the block count
  // read on the device.
  const uint32_t nunits = du.dev_counts ? du.dev_counts[0] : du.count;
  const uint32_t nblocks = du.dev_counts ? du.dev_counts[1] : du.blocks;
  for (uint32_t vb = blockIdx.x; vb < nblocks; vb += gridDim.x) {
    __syncthreads();  // the previous block's flush is done with tile and s_unit
    if (threadIdx.x == 0) {
      uint32_t lo = 0, hi = nunits - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (du.first_block[mid] <= vb) lo = mid;
        else hi = mid - 1;
      }
      s_unit = lo;
    }
    __syncthreads();
    const uint32_t ui = s_unit;
    const DepositUnit u = du.units[ui];
    const uint32_t p0 = u.begin + (vb - du.first_block[ui]) * kDepositChunk;
    const uint32_t p1 = min(p0 + kDepositChunk, u.begin + u.count);
    const int cnt = a.scheme == 0 ? 2 : 3;
    const bool tiled = u.dims[0] > 0;
    const int dy = u.dims[1], dz = u.dims[2];
    const int vol = tiled ? u.dims[0] * dy * dz : 0;
    for (int i = threadIdx.x; i < vol; i += blockDim.x) tile[i] = T(0);
    __syncthreads();
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
      const S sp = src[p];
      int idx[3][3];
      T w[3][3];
      stencil(a, sp, idx, w);
      const T mm = uniform ? um : source_mass(sp);
      int o[3];
      bool in = tiled;
      for (int ax = 0; ax < 3; ++ax) {
        o[ax] = wrap_once(idx[ax][0] - u.lo[ax], a.n[ax]);
        in = in && o[ax] + cnt <= u.dims[ax];
      }
      for (int i = 0; i < cnt; ++i)
        for (int j = 0; j < cnt; ++j) {
          const T mij = mm * w[0][i] * w[1][j];
          if (in) {
            T* t = &tile[((o[0] + i) * dy + (o[1] + j)) * dz + o[2]];
            for (int k = 0; k < cnt; ++k) atomicAdd(t + k, mij * w[2][k]);
          } else {
            const std::size_t rb = (static_cast<std::size_t>(idx[0][i]) * a.n[1] + idx[1][j]) * a.row;
            for (int k = 0; k < cnt; ++k) atomicAdd(&rho[rb + idx[2][k]], mij * w[2][k]);
          }
        }
    }
    if (!tiled) continue;  // block-uniform
    __syncthreads();
    const int rows = u.dims[0] * dy;
    const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int r = wv; r < rows; r += blockDim.x / kWave) {
      const int ix = r / dy, iy = r - ix * dy;
      const std::size_t base =
          (static_cast<std::size_t>(wrap_once(u.lo[0] + ix, a.n[0])) * a.n[1] + wrap_once(u.lo[1] + iy, a.n[1])) *
          a.row;
      for (int iz = lane; iz < dz; iz += kWave) {
        const T v = tile[r * dz + iz];
        if (v != T(0)) atomicAdd(&rho[base + wrap_once(u.lo[2] + iz, a.n[2])], v);
      }
    }
  }
}

// ---- z passes (real <-> half complex, two rows per complex FFT) -----------------
template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_z_r2c_kernel(T* data, std::size_t rows, int nz, int lg,
                                                                std::size_t row, const typename C2<T>::type* twz,
                                                                int pairs_per_block) {
  using V = typename C2<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  const int twn = fft_tw_len(nz, lg);
  V* X = tw + twn;
  const bool pad = lg >= 0;
  const int stride = lds_line_stride(nz, pad);
  for (int i = threadIdx.x; i < twn; i += blockDim.x) tw[i] = twz[i];
  const std::size_t pair0 = static_cast<std::size_t>(blockIdx.x) * pairs_per_block;
  const int npts = pairs_per_block * nz;
  for (int base = threadIdx.x; base < npts; base += fft_z_load_batch<T>() * kFftThreads) {
    V v[fft_z_load_batch<T>()];
#pragma unroll
    for (int u = 0; u < fft_z_load_batch<T>(); ++u) {  // all loads in flight before the LDS stores
      const int idx = base + u * kFftThreads;
      v[u] = mk<V, T>(T(0), T(0));
      int pp, i;
      split_index(idx, nz, lg, pp, i);
      const std::size_t pr = pair0 + pp;
      if (idx < npts && 2 * pr < rows) {
        const T* ra = data + (2 * pr) * row;
        v[u].x = ra[i];
        if (2 * pr + 1 < rows) v[u].y = ra[row + i];
      }
    }
#pragma unroll
    for (int u = 0; u < fft_z_load_batch<T>(); ++u) {
      const int idx = base + u * kFftThreads;
      if (idx < npts) {
        int pp, i;
        split_index(idx, nz, lg, pp, i);
        X[pp * stride + lds_slot(i, pad)] = v[u];
      }
    }
  }
  __syncthreads();
  lds_fft<T>(X, pairs_per_block, nz, lg, stride, tw, false);
  const int nzc = (nz >> 1) + 1;
  for (int idx = threadIdx.x; idx < pairs_per_block * nzc; idx += blockDim.x) {
    const int pp = idx / nzc, k = idx - pp * nzc;
    const std::size_t pr = pair0 + pp;
    if (2 * pr >= rows) continue;
    const V zk = X[pp * stride + lds_slot(k == 0 ? 0 : k, pad)];
    const V zm = X[pp * stride + lds_slot(k == 0 ? 0 : nz - k, pad)];
    const V zmc = mk<V, T>(zm.x, -zm.y);
    const V s = cadd(zk, zmc), d = csub(zk, zmc);
    V* ca = reinterpret_cast<V*>(data + (2 * pr) * row);
    ca[k] = mk<V, T>(T(0.5) * s.x, T(0.5) * s.y);
    if (2 * pr + 1 < rows) {
      V* cb = reinterpret_cast<V*>(data + (2 * pr + 1) * row);
      cb[k] = mk<V, T>(T(0.5) * d.y, T(-0.5) * d.x);  // (d) * (-i/2)
    }
  }
}

template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_z_c2r_kernel(T* data, std::size_t rows, int nz, int lg,
                                                                std::size_t row, const typename C2<T>::type* twz,
                                                                const T* __restrict__ dkz, int premul,
                                                                int pairs_per_block) {
  using V = typename C2<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  const int twn = fft_tw_len(nz, lg);
  V* X = tw + twn;
  const bool pad = lg >= 0;
  const int stride = lds_line_stride(nz, pad);
  const int nzc = (nz >> 1) + 1;
  for (int i = threadIdx.x; i < twn; i += blockDim.x) tw[i] = twz[i];
  const std::size_t pair0 = static_cast<std::size_t>(blockIdx.x) * pairs_per_block;
  const int npts = pairs_per_block * nzc;
  for (int base = threadIdx.x; base < npts; base += fft_z_load_batch<T>() * kFftThreads) {
    V va[fft_z_load_batch<T>()], vb[fft_z_load_batch<T>()];
#pragma unroll
    for (int u = 0; u < fft_z_load_batch<T>(); ++u) {  // all loads in flight before the LDS stores
      const int idx = base + u * kFftThreads;
      va[u] = mk<V, T>(T(0), T(0));
      vb[u] = mk<V, T>(T(0), T(0));
      const int pp = idx / nzc, k = idx - pp * nzc;
      const std::size_t pr = pair0 + pp;
      if (idx < npts && 2 * pr < rows) {
        va[u] = reinterpret_cast<const V*>(data + (2 * pr) * row)[k];
        if (2 * pr + 1 < rows) vb[u] = reinterpret_cast<const V*>(data + (2 * pr + 1) * row)[k];
      }
    }
#pragma unroll
    for (int u = 0; u < fft_z_load_batch<T>(); ++u) {
      const int idx = base + u * kFftThreads;
      if (idx >= npts) continue;
      const int pp = idx / nzc, k = idx - pp * nzc;
      V xa = va[u], xb = vb[u];
      if (premul) {  // multiply by -i k_z
        const T kk = dkz[k];
        xa = mk<V, T>(xa.y * kk, -xa.x * kk);
        xb = mk<V, T>(xb.y * kk, -xb.x * kk);
      }
      if (k == 0 || k == nz / 2) {
        X[pp * stride + lds_slot(k, pad)] = mk<V, T>(xa.x, xb.x);
      } else {
        X[pp * stride + lds_slot(k, pad)] = mk<V, T>(xa.x - xb.y, xa.y + xb.x);          // xa + i xb
        X[pp * stride + lds_slot(nz - k, pad)] = mk<V, T>(xa.x + xb.y, -xa.y + xb.x);  // conj(xa) + i conj(xb)
      }
    }
  }
  __syncthreads();
  lds_fft<T>(X, pairs_per_block, nz, lg, stride, tw, true);
  for (int idx = threadIdx.x; idx < pairs_per_block * nz; idx += blockDim.x) {
    const int pp = idx / nz, i = idx - pp * nz;
    const std::size_t pr = pair0 + pp;
    if (2 * pr >= rows) continue;
    const V v = X[pp * stride + lds_slot(i, pad)];
    data[(2 * pr) * row + i] = v.x;
    if (2 * pr + 1 < rows) data[(2 * pr + 1) * row + i] = v.y;
  }
}

// ---- complex line passes (y and x) ---------------------------------------------
struct LineArgs {
  int n, lg;
  std::size_t line_stride;   // complex elements between points of a line
  std::size_t inner_count;   // contiguous line starts per outer index
  std::size_t outer_stride;
  int tj;                    // lines per block (power of two)
  int inverse;
  int premul;                // multiply by -i dk[i] before the transform
  int green;                 // multiply by the Green's function after (x pass only)
  // Green's function support (x pass): inner index j = iy * nzc + kz
  std::size_t nzc;
  double pref, split, fold;
  int deconv;
};

template <class T>
__global__ __launch_bounds__(kFftThreads) void fft_lines_kernel(const typename C2<T>::type* in,
                                                                typename C2<T>::type* out, LineArgs a,
                                                                const typename C2<T>::type* twl,
                                                                const T* __restrict__ dk,
                                                                const T* __restrict__ kx, const T* __restrict__ ky,
                                                                const T* __restrict__ kz, const T* __restrict__ wx,
                                                                const T* __restrict__ wy, const T* __restrict__ wz) {
  using V = typename C2<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V* tw = reinterpret_cast<V*>(smem_raw);
  const int twn = fft_tw_len(a.n, a.lg);
  V* X = tw + twn;
  const bool pad = a.lg >= 0;
  const int stride = lds_line_stride(a.n, pad);
  for (int i = threadIdx.x; i < twn; i += blockDim.x) tw[i] = twl[i];
  const std::size_t j0 = static_cast<std::size_t>(blockIdx.x) * a.tj;
  const std::size_t obase = static_cast<std::size_t>(blockIdx.y) * a.outer_stride;
  const int tjm = a.tj - 1;
  int tjl = 0;
  while ((1 << tjl) < a.tj) ++tjl;
  const int npts = a.tj * a.n;
  for (int base = threadIdx.x; base < npts; base += fft_load_batch<T>() * kFftThreads) {
    V v[fft_load_batch<T>()];
#pragma unroll
    for (int u = 0; u < fft_load_batch<T>(); ++u) {  // all loads in flight before the LDS stores
      const int idx = base + u * kFftThreads;
      const int jj = idx & tjm, i = idx >> tjl;
      const std::size_t j = j0 + jj;
      v[u] = mk<V, T>(T(0), T(0));
      if (idx < npts && j < a.inner_count) v[u] = in[obase + j + static_cast<std::size_t>(i) * a.line_stride];
    }
#pragma unroll
    for (int u = 0; u < fft_load_batch<T>(); ++u) {
      const int idx = base + u * kFftThreads;
      if (idx >= npts) continue;
      const int jj = idx & tjm, i = idx >> tjl;
      V vv = v[u];
      if (a.premul) {
        const T kk = dk[i];
        vv = mk<V, T>(vv.y * kk, -vv.x * kk);
      }
      X[jj * stride + lds_slot(i, pad)] = vv;
    }
  }
  __syncthreads();
  lds_fft<T>(X, a.tj, a.n, a.lg, stride, tw, a.inverse != 0);
  for (int idx = threadIdx.x; idx < a.tj * a.n; idx += blockDim.x) {
    const int jj = idx & tjm, i = idx >> tjl;
    const std::size_t j = j0 + jj;
    if (j >= a.inner_count) continue;
    V v = X[jj * stride + lds_slot(i, pad)];
    if (a.green) {
      const std::size_t iy = j / a.nzc, iz = j - iy * a.nzc;
      const T k2 = kx[i] * kx[i] + ky[iy] * ky[iy] + kz[iz] * kz[iz];
      T gval = T(0);
      if (k2 != T(0)) {
        T corr = T(1);
        if (a.deconv) {
          const T tw_ = wx[i] * wy[iy] * wz[iz];
          corr = T(1) / fmax(tw_ * tw_, T(1.0e-12));
        }
        T filt = T(1);
        if (a.split > 0.0) {
          const T ka = sqrt(k2) * T(a.split);
          filt = exp(-(ka * ka));
        }
        gval = T(a.pref) * corr * filt / k2 * T(a.fold);
      }
      v.x *= gval;
      v.y *= gval;
    }
    out[obase + j + static_cast<std::size_t>(i) * a.line_stride] = v;
  }
}

// ---- host -------------------------------------------------------------------
template <class T>
struct Tables {
  T *kx, *ky, *kz, *dkx, *dky, *dkz, *wx, *wy, *wz;
  typename C2<T>::type *twx, *twy, *twz;
  int lines = 0;  // lines per block cap of the line passes (PmOptions::fft_lines)
};

template <class T>
Tables<T> upload_tables(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, hipStream_t s) {
  using V = typename C2<T>::type;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const std::size_t nzc = g.nzc;
  const double key[8] = {static_cast<double>(nx), static_cast<double>(ny), static_cast<double>(nz), g.L[0], g.L[1],
                         g.L[2], static_cast<double>(o.assignment), static_cast<double>(o.deconvolution)};
  auto tables_at = [&](const std::size_t* off) {
    T* base = w.tables.as<T>();
    V* tb = reinterpret_cast<V*>(base + off[9]);
    return Tables<T>{base + off[0], base + off[1], base + off[2], base + off[3], base + off[4], base + off[5],
                     base + off[6], base + off[7], base + off[8], tb + off[10], tb + off[10] + nx,
                     tb + off[10] + nx + ny, o.fft_lines};
  };
  if (w.tables.ptr && std::equal(key, key + 8, w.table_key)) return tables_at(w.table_off);
  std::vector<T> host;
  auto push_k = [&](int n, std::size_t count, double L, bool deriv, bool window, int wexp) {
    const double dk = 2.0 * kPi / L;
    for (std::size_t i = 0; i < count; ++i) {
      const long mode = i <= static_cast<std::size_t>(n) / 2 ? static_cast<long>(i) : static_cast<long>(i) - n;
      const double k = dk * static_cast<double>(mode);
      if (window) {
        double wv = 1.0;
        if (o.deconvolution) {
          const double xx = 0.5 * k * L / static_cast<double>(n);
          const double sn = std::abs(xx) < 1.0e-12 ? 1.0 : std::sin(xx) / xx;
          wv = std::pow(sn, wexp);
        }
        host.push_back(static_cast<T>(wv));
      } else if (deriv) {
        host.push_back(static_cast<T>((n % 2 == 0 && i == static_cast<std::size_t>(n) / 2) ? 0.0 : k));
      } else {
        host.push_back(static_cast<T>(k));
      }
    }
  };
  const int wexp = o.assignment == 0 ? 2 : 3;
  std::size_t off[9];
  off[0] = host.size(); push_k(nx, nx, g.L[0], false, false, wexp);
  off[1] = host.size(); push_k(ny, ny, g.L[1], false, false, wexp);
  off[2] = host.size(); push_k(nz, nzc, g.L[2], false, false, wexp);
  off[3] = host.size(); push_k(nx, nx, g.L[0], true, false, wexp);
  off[4] = host.size(); push_k(ny, ny, g.L[1], true, false, wexp);
  off[5] = host.size(); push_k(nz, nzc, g.L[2], true, false, wexp);
  off[6] = host.size(); push_k(nx, nx, g.L[0], false, true, wexp);
  off[7] = host.size(); push_k(ny, ny, g.L[1], false, true, wexp);
  off[8] = host.size(); push_k(nz, nzc, g.L[2], false, true, wexp);
  while (host.size() % 4) host.push_back(T(0));
  const std::size_t scal = host.size();
  std::vector<V> tws;
  std::size_t toff[3];
  for (int ax = 0; ax < 3; ++ax) {
    toff[ax] = tws.size();
    const int n = g.n[ax];
    for (int k = 0; k < n; ++k) {  // full circle; power-of-two lines use the first half
      const double ang = -2.0 * kPi * static_cast<double>(k) / static_cast<double>(n);
      V v;
      v.x = static_cast<T>(std::cos(ang));
      v.y = static_cast<T>(std::sin(ang));
      tws.push_back(v);
    }
  }
  const std::size_t bytes = scal * sizeof(T) + tws.size() * sizeof(V);
  require_not_capturing(s, "PM tables");
  w.tables.reserve(bytes);
  CHUI_HIP(hipMemcpyAsync(w.tables.ptr, host.data(), scal * sizeof(T), hipMemcpyHostToDevice, s));
  CHUI_HIP(hipMemcpyAsync(static_cast<char*>(w.tables.ptr) + scal * sizeof(T), tws.data(), tws.size() * sizeof(V),
                          hipMemcpyHostToDevice, s));
  CHUI_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  for (int i = 0; i < 9; ++i) w.table_off[i] = off[i];
  w.table_off[9] = scal;
  w.table_off[10] = toff[0];
  w.table_off[11] = 0;
  std::copy(key, key + 8, w.table_key);
  return tables_at(w.table_off);
}

// Dynamic LDS a line-pass block may use: 80 KB (two blocks per CU, 16 lines
// of 512 float2) when the device allows more than 64 KB per block.
std::size_t fft_lds_limit() {
  static std::size_t limit = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return std::size_t(64 * 1024);
    return std::min<std::size_t>(static_cast<std::size_t>(v), 80 * 1024);
  }();
  return limit;
}

int line_log2(int n) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  return (1 << lg) == n ? lg : -1;
}

template <class T>
int lines_per_block(int n) {
  using V = typename C2<T>::type;
  const int lg = line_log2(n);
  const int cap = fft_max_butterflies<T>() * kFftThreads * fft_plan_min_radix(n);  // points per block
  const std::size_t lim = fft_lds_limit();
  const std::size_t twb = static_cast<std::size_t>(fft_tw_len(n, lg)) * sizeof(V);
  int tj = 16;
  const std::size_t ls = static_cast<std::size_t>(lds_line_stride(n, lg >= 0));
  while (tj > 1 && (static_cast<std::size_t>(tj) * ls * sizeof(V) + twb > lim || tj * n > cap)) tj >>= 1;
  if (tj * n > cap || static_cast<std::size_t>(tj) * ls * sizeof(V) + twb > lim)
    throw UnsupportedError("PM mesh axis too long for one LDS line");
  return tj;
}

template <class T>
void launch_lines(const typename C2<T>::type* in, typename C2<T>::type* out, LineArgs a, std::size_t outer_count,
                  const typename C2<T>::type* tw, const T* dk, const Tables<T>& tb, hipStream_t s) {
  using V = typename C2<T>::type;
  a.tj = lines_per_block<T>(a.n);
  if (tb.lines > 0 && tb.lines < a.tj) a.tj = tb.lines;  // a power of two below the LDS-bound maximum
  const std::size_t smem =
      static_cast<std::size_t>(a.tj) * lds_line_stride(a.n, a.lg >= 0) * sizeof(V) + fft_tw_len(a.n, a.lg) * sizeof(V);
  if (smem > 64 * 1024) {
    // contexts of an ensemble launch from several host threads
    static std::mutex mu;
    static std::size_t granted = 0;
    std::lock_guard<std::mutex> lock(mu);
    if (smem > granted) {
      CHUI_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_lines_kernel<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(smem)));
      granted = smem;
    }
  }
  dim3 grid(static_cast<unsigned>((a.inner_count + a.tj - 1) / a.tj), static_cast<unsigned>(outer_count));
  fft_lines_kernel<T><<<grid, kFftThreads, smem, s>>>(in, out, a, tw, dk, tb.kx, tb.ky, tb.kz, tb.wx, tb.wy, tb.wz);
  CHUI_LAUNCH_CHECK("fft_lines_kernel");
}

template <class T>
int pairs_per_block(int nz) {
  using V = typename C2<T>::type;
  const int lg = line_log2(nz);
  const int cap = fft_max_butterflies<T>() * kFftThreads * fft_plan_min_radix(nz);  // points per block
  const std::size_t twb = static_cast<std::size_t>(fft_tw_len(nz, lg)) * sizeof(V);
  int pp = 8;
  const std::size_t ls = static_cast<std::size_t>(lds_line_stride(nz, lg >= 0));
  while (pp > 1 && (static_cast<std::size_t>(pp) * ls * sizeof(V) + twb > 48 * 1024 || pp * nz > cap)) pp >>= 1;
  if (pp * nz > cap || static_cast<std::size_t>(pp) * ls * sizeof(V) + twb > 64 * 1024)
    throw UnsupportedError("PM mesh z axis too long for one LDS line");
  return pp;
}

template <class T>
void z_r2c(T* data, const PmGeometry& g, const Tables<T>& tb, hipStream_t s) {
  using V = typename C2<T>::type;
  const std::size_t rows = static_cast<std::size_t>(g.n[0]) * g.n[1];
  const int pp = pairs_per_block<T>(g.n[2]);
  const std::size_t pairs = (rows + 1) / 2;
  const std::size_t smem =
      static_cast<std::size_t>(pp) * lds_line_stride(g.n[2], g.lg[2] >= 0) * sizeof(V) +
      fft_tw_len(g.n[2], g.lg[2]) * sizeof(V);
  fft_z_r2c_kernel<T><<<static_cast<unsigned>((pairs + pp - 1) / pp), kFftThreads, smem, s>>>(
      data, rows, g.n[2], g.lg[2], g.row, tb.twz, pp);
  CHUI_LAUNCH_CHECK("fft_z_r2c_kernel");
}

template <class T>
void z_c2r(T* data, const PmGeometry& g, const Tables<T>& tb, bool premul, hipStream_t s) {
  using V = typename C2<T>::type;
  const std::size_t rows = static_cast<std::size_t>(g.n[0]) * g.n[1];
  const int pp = pairs_per_block<T>(g.n[2]);
  const std::size_t pairs = (rows + 1) / 2;
  const std::size_t smem =
      static_cast<std::size_t>(pp) * lds_line_stride(g.n[2], g.lg[2] >= 0) * sizeof(V) +
      fft_tw_len(g.n[2], g.lg[2]) * sizeof(V);
  fft_z_c2r_kernel<T><<<static_cast<unsigned>((pairs + pp - 1) / pp), kFftThreads, smem, s>>>(
      data, rows, g.n[2], g.lg[2], g.row, tb.twz, tb.dkz, premul ? 1 : 0, pp);
  CHUI_LAUNCH_CHECK("fft_z_c2r_kernel");
}

// Forward r2c FFT of the real mesh in A followed by the Green's function with
// its normalisation (the operator up to the gradient), in place.
template <class T>
void forward_green(T* A, const PmGeometry& g, const PmOptions& o, const Tables<T>& tb, hipStream_t s) {
  using V = typename C2<T>::type;
  V* Ac = reinterpret_cast<V*>(A);
  const int nx = g.n[0], ny = g.n[1];
  const std::size_t nzc = g.nzc;
  const double M = static_cast<double>(g.n[0]) * g.n[1] * g.n[2];
  const double cell_volume = (g.L[0] * g.L[1] * g.L[2]) / M;
  z_r2c<T>(A, g, tb, s);
  LineArgs ya{};
  ya.n = ny; ya.lg = g.lg[1]; ya.line_stride = nzc; ya.inner_count = nzc; ya.outer_stride = static_cast<std::size_t>(ny) * nzc;
  ya.inverse = 0; ya.premul = 0; ya.green = 0; ya.nzc = nzc;
  launch_lines<T>(Ac, Ac, ya, nx, tb.twy, nullptr, tb, s);
  LineArgs xa{};
  xa.n = nx; xa.lg = g.lg[0]; xa.line_stride = static_cast<std::size_t>(ny) * nzc; xa.inner_count = static_cast<std::size_t>(ny) * nzc;
  xa.outer_stride = 0; xa.inverse = 0; xa.premul = 0; xa.green = 1; xa.nzc = nzc;
  xa.pref = -4.0 * kPi * o.G;
  xa.split = o.split_scale;
  xa.fold = 1.0 / (cell_volume * M);
  xa.deconv = o.deconvolution;
  launch_lines<T>(Ac, Ac, xa, 1, tb.twx, nullptr, tb, s);
}

// nu_hat = i (kx mu0_hat + ky mu1_hat + kz mu2_hat), written over mu0_hat: the
// transpose of the three gradient operators -i k_a (same Nyquist rule), summed.
template <class T>
__global__ __launch_bounds__(256) void divergence_kernel(typename C2<T>::type* __restrict__ a0,
                                                         const typename C2<T>::type* __restrict__ a1,
                                                         const typename C2<T>::type* __restrict__ a2, int ny,
                                                         std::size_t nzc, std::size_t cells, const T* __restrict__ dkx,
                                                         const T* __restrict__ dky, const T* __restrict__ dkz) {
  using V = typename C2<T>::type;
  const std::size_t c = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const std::size_t kz = c % nzc, r = c / nzc;
  const std::size_t iy = r % static_cast<std::size_t>(ny), ix = r / static_cast<std::size_t>(ny);
  const V u0 = a0[c], u1 = a1[c], u2 = a2[c];
  const T kx = dkx[ix], ky = dky[iy], kz_ = dkz[kz];
  V v;
  v.x = -(kx * u0.y + ky * u1.y + kz_ * u2.y);
  v.y = kx * u0.x + ky * u1.x + kz_ * u2.x;
  a0[c] = v;
}

template <class T>
void solve_spectral(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const Tables<T>& tb, hipStream_t s) {
  using V = typename C2<T>::type;
  T* A = w.buf[0].as<T>();
  T* B = w.buf[1].as<T>();
  T* Cb = w.buf[2].as<T>();
  V* Ac = reinterpret_cast<V*>(A);
  V* Bc = reinterpret_cast<V*>(B);
  V* Cc = reinterpret_cast<V*>(Cb);
  const int nx = g.n[0], ny = g.n[1];
  const std::size_t nzc = g.nzc;
  // forward z (r2c), y and x lines, Green's function (Phi in A)
  forward_green<T>(A, g, o, tb, s);
  // lines along y: inner = kz (contiguous), outer = ix; along x: inner = (iy, kz)
  LineArgs ya{};
  ya.n = ny; ya.lg = g.lg[1]; ya.line_stride = nzc; ya.inner_count = nzc; ya.outer_stride = static_cast<std::size_t>(ny) * nzc;
  ya.inverse = 0; ya.premul = 0; ya.green = 0; ya.nzc = nzc;
  LineArgs xi{};
  xi.n = nx; xi.lg = g.lg[0]; xi.line_stride = static_cast<std::size_t>(ny) * nzc; xi.inner_count = static_cast<std::size_t>(ny) * nzc;
  xi.outer_stride = 0; xi.nzc = nzc;
  // inverse x: Qx = IFFT_x(-i kx Phi) -> B ; P = IFFT_x(Phi) -> C
  xi.inverse = 1; xi.green = 0; xi.premul = 1;
  launch_lines<T>(Ac, Bc, xi, 1, tb.twx, tb.dkx, tb, s);
  xi.premul = 0;
  launch_lines<T>(Ac, Cc, xi, 1, tb.twx, nullptr, tb, s);
  // inverse y: B -> B ; C (-i ky) -> A ; C -> C
  LineArgs yi = ya;
  yi.inverse = 1;
  launch_lines<T>(Bc, Bc, yi, nx, tb.twy, nullptr, tb, s);
  yi.premul = 1;
  launch_lines<T>(Cc, Ac, yi, nx, tb.twy, tb.dky, tb, s);
  yi.premul = 0;
  launch_lines<T>(Cc, Cc, yi, nx, tb.twy, nullptr, tb, s);
  // inverse z (c2r) in place: B -> Ax, A -> Ay, C (-i kz) -> Az
  z_c2r<T>(B, g, tb, false, s);
  z_c2r<T>(A, g, tb, false, s);
  z_c2r<T>(Cb, g, tb, true, s);
  w.force[0] = B;
  w.force[1] = A;
  w.force[2] = Cb;
}

template <class T>
void reserve(PmWorkspace& w, const PmGeometry& g) {
  const std::size_t bytes = g.complex_cells() * 2 * sizeof(T);
  for (auto& b : w.buf) b.reserve(bytes);
  w.cells = g.complex_cells();
  w.elem_bytes = sizeof(T);
}

DepositArgs deposit_args(const PmGeometry& g, const PmOptions& o) {
  DepositArgs a{};
  for (int ax = 0; ax < 3; ++ax) {
    a.n[ax] = g.n[ax];
    a.lg[ax] = g.lg[ax];
    a.L[ax] = g.L[ax];
    a.inv_h[ax] = static_cast<double>(g.n[ax]) / g.L[ax];
  }
  a.row = g.row;
  a.scheme = o.assignment;
  return a;
}

}  // namespace

void pm_solve_fp64(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const double4* src, int uniform_mass,
                   double um, std::size_t n, const DepositUnits& units, hipStream_t s) {
  w.valid = false;  // the force meshes are overwritten from here on
  reserve<double>(w, g);
  const Tables<double> tb = upload_tables<double>(w, g, o, s);
  double* rho = w.buf[0].as<double>();
  fill_f64(rho, 0.0, g.complex_cells() * 2, s);
  if (n > 0) {
    if (units.count == 0 && !units.dev_counts) throw std::logic_error("PM deposit without deposit units");
    deposit_kernel<double, double4><<<units.blocks, 256, 0, s>>>(deposit_args(g, o), src, uniform_mass, um, units,
                                                                  rho);
    CHUI_LAUNCH_CHECK("deposit_kernel");
  }
  solve_spectral<double>(w, g, o, tb, s);
  w.valid = true;
}

void pm_adjoint_fp64(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const double4* const comp[3],
                     std::size_t n, const DepositUnits& units, hipStream_t s) {
  using V = double2;
  reserve<double>(w, g);
  const Tables<double> tb = upload_tables<double>(w, g, o, s);
  for (int a = 0; a < 3; ++a) {
    double* mu = w.buf[a].as<double>();
    fill_f64(mu, 0.0, g.complex_cells() * 2, s);
    if (n > 0) {
      if (units.count == 0 && !units.dev_counts) throw std::logic_error("PM adjoint deposit without deposit units");
      deposit_kernel<double, double4><<<units.blocks, 256, 0, s>>>(deposit_args(g, o), comp[a], 0, 0.0, units, mu);
      CHUI_LAUNCH_CHECK("deposit_kernel");
    }
    forward_green<double>(mu, g, o, tb, s);
  }
  V* A = w.buf[0].as<V>();
  const std::size_t cells = g.complex_cells();
  divergence_kernel<double><<<grid_for(cells, 256), 256, 0, s>>>(A, w.buf[1].as<V>(), w.buf[2].as<V>(), g.n[1],
                                                                g.nzc, cells, tb.dkx, tb.dky, tb.dkz);
  CHUI_LAUNCH_CHECK("divergence_kernel");
  LineArgs xi{};
  xi.n = g.n[0]; xi.lg = g.lg[0]; xi.line_stride = static_cast<std::size_t>(g.n[1]) * g.nzc;
  xi.inner_count = static_cast<std::size_t>(g.n[1]) * g.nzc; xi.outer_stride = 0; xi.inverse = 1; xi.premul = 0;
  xi.green = 0; xi.nzc = g.nzc;
  launch_lines<double>(A, A, xi, 1, tb.twx, nullptr, tb, s);
  LineArgs yi{};
  yi.n = g.n[1]; yi.lg = g.lg[1]; yi.line_stride = g.nzc; yi.inner_count = g.nzc;
  yi.outer_stride = static_cast<std::size_t>(g.n[1]) * g.nzc; yi.inverse = 1; yi.premul = 0; yi.green = 0;
  yi.nzc = g.nzc;
  launch_lines<double>(A, A, yi, g.n[0], tb.twy, nullptr, tb, s);
  z_c2r<double>(w.buf[0].as<double>(), g, tb, false, s);
  w.valid = false;  // holds nu, not a force field
}

void pm_solve_mixed(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const uint4* src, int uniform_mass,
                    float um, std::size_t n, const DepositUnits& units, hipStream_t s) {
  w.valid = false;  // the force meshes are overwritten from here on
  reserve<float>(w, g);
  const Tables<float> tb = upload_tables<float>(w, g, o, s);
  float* rho = w.buf[0].as<float>();
  fill_f32(rho, 0.0f, g.complex_cells() * 2, s);
  if (n > 0) {
    if (units.count == 0 && !units.dev_counts) throw std::logic_error("PM deposit without deposit units");
    deposit_kernel<float, uint4><<<units.blocks, 256, 0, s>>>(deposit_args(g, o), src, uniform_mass, um, units, rho);
    CHUI_LAUNCH_CHECK("deposit_kernel");
  }
  solve_spectral<float>(w, g, o, tb, s);
  w.valid = true;
}

}  // namespace chui

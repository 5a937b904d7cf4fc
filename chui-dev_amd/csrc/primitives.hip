// primitives.hip — stable LSD radix sort, exclusive scan and fills for gfx950.
//
// The sort is a classic three-kernel LSD pass (tile histogram, digit-major scan,
// stable scatter). Stability inside a tile comes from wave64 ballot ranking:
// each wave ranks 64 consecutive items per round with 8 ballots (one per digit
// bit), keeps running per-digit counts in LDS across rounds, and the block
// combines the four waves' counts in wave order. Item order inside a tile is
// therefore (wave, round, lane) = input order. The scatter first sorts the
// tile in LDS and then stores it in order, one digit run after another.
#include "primitives.hpp"

namespace chui {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;                 // 4
constexpr int kSortRounds = 8;                                   // rounds of 64 per wave
constexpr int kSortTile = kSortWaves * kSortRounds * kWave;      // 2048 items

// ---- exclusive scan ---------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* block_total) {
  __shared__ uint32_t wave_tot[kScanThreads / kWave];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t o = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += o;
  }
  if (lane == kWave - 1) wave_tot[w] = inc;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / kWave; ++i) {
    if (i < w) wave_off += wave_tot[i];
    tot += wave_tot[i];
  }
  __syncthreads();
  *block_total = tot;
  return wave_off + inc - v;
}

template <class K>
__global__ __launch_bounds__(kSortThreads) void sort_hist_kernel(const K* __restrict__ keys, std::size_t n,
                                                                 int shift, uint32_t dmask,
                                                                 uint32_t* __restrict__ hist, uint32_t tiles) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kSortTile;
#pragma unroll
  for (int k = 0; k < kSortTile / kSortThreads; ++k) {
    const std::size_t idx = base + static_cast<std::size_t>(k) * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[static_cast<uint32_t>(keys[idx] >> shift) & dmask], 1u);
  }
  __syncthreads();
  hist[static_cast<std::size_t>(threadIdx.x) * tiles + blockIdx.x] = cnt[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(kSortThreads) void sort_scatter_kernel(
    const K* __restrict__ keys, const uint32_t* __restrict__ vals, K* __restrict__ okeys,
    uint32_t* __restrict__ ovals, std::size_t n, int shift, uint32_t dmask,
    const uint32_t* __restrict__ offsets, uint32_t tiles) {
  __shared__ uint32_t wcount[kSortWaves][256];
  __shared__ uint32_t gbase[256];
  __shared__ uint32_t lbase[256];
  __shared__ K skeys[kSortTile];
  __shared__ uint32_t svals[kSortTile];
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  for (int i = threadIdx.x; i < kSortWaves * 256; i += kSortThreads) (&wcount[0][0])[i] = 0;
  gbase[threadIdx.x] = offsets[static_cast<std::size_t>(threadIdx.x) * tiles + blockIdx.x];
  __syncthreads();

  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kSortTile +
                           static_cast<std::size_t>(w) * (kSortRounds * kWave);
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  K k[kSortRounds];
  uint32_t v[kSortRounds];
  uint32_t rank[kSortRounds];
  int dig[kSortRounds];
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    const std::size_t idx = base + static_cast<std::size_t>(r) * kWave + lane;
    const bool valid = idx < n;
    uint32_t d = 0;
    k[r] = 0;
    v[r] = 0;
    if (valid) {
      k[r] = keys[idx];
      v[r] = vals[idx];
      d = static_cast<uint32_t>(k[r] >> shift) & dmask;
    }
    uint64_t mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bm = __ballot(valid && bit);
      mask &= bit ? bm : ~bm;
    }
    uint32_t before = 0;
    if (valid) before = wcount[w][d];
    const uint32_t rk = static_cast<uint32_t>(__popcll(mask & lt));
    if (valid && rk == 0) wcount[w][d] = before + static_cast<uint32_t>(__popcll(mask));
    rank[r] = before + rk;
    dig[r] = valid ? static_cast<int>(d) : -1;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;
    uint32_t acc = 0;
#pragma unroll
    for (int ww = 0; ww < kSortWaves; ++ww) {
      const uint32_t c = wcount[ww][d];
      wcount[ww][d] = acc;
      acc += c;
    }
    uint32_t tile_total;
    lbase[d] = block_exclusive_sum(acc, &tile_total);  // digit d's first slot in the sorted tile
  }
  __syncthreads();
  // Sort the tile in LDS, then write it out in order: consecutive lanes store
  // consecutive positions of one digit's run, so the stores coalesce.
#pragma unroll
  for (int r = 0; r < kSortRounds; ++r) {
    if (dig[r] >= 0) {
      const uint32_t lp = lbase[dig[r]] + wcount[w][dig[r]] + rank[r];
      skeys[lp] = k[r];
      svals[lp] = v[r];
    }
  }
  __syncthreads();
  const std::size_t t0 = static_cast<std::size_t>(blockIdx.x) * kSortTile;
  const uint32_t in_tile = static_cast<uint32_t>(n - t0 < static_cast<std::size_t>(kSortTile) ? n - t0 : kSortTile);
  for (uint32_t i = threadIdx.x; i < in_tile; i += kSortThreads) {
    const K key = skeys[i];
    const uint32_t d = static_cast<uint32_t>(key >> shift) & dmask;
    const uint32_t pos = gbase[d] + (i - lbase[d]);
    if (pos >= n) continue;  // unreachable when the histogram is consistent
    okeys[pos] = key;
    ovals[pos] = svals[i];
  }
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, std::size_t n,
                                                                   uint32_t* __restrict__ partial) {
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const std::size_t idx = base + static_cast<std::size_t>(k) * kScanThreads + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  uint32_t tot;
  (void)block_exclusive_sum(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* in, uint32_t* out, std::size_t n,
                                                                  const uint32_t* __restrict__ partial_scanned,
                                                                  uint32_t* total_dev) {
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile +
                           static_cast<std::size_t>(threadIdx.x) * kScanItems;
  uint32_t x[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const std::size_t idx = base + k;
    x[k] = idx < n ? in[idx] : 0u;
    s += x[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_sum(s, &tot) + (partial_scanned ? partial_scanned[blockIdx.x] : 0u);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const std::size_t idx = base + k;
    if (idx < n) out[idx] = run;
    run += x[k];
  }
  if (total_dev && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total_dev = run;
}

template <class T>
__global__ void fill_kernel(T* p, T v, std::size_t n) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void iota_kernel(uint32_t* p, std::size_t n) {
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
    p[i] = static_cast<uint32_t>(i);
}

template <class T>
void fill_any(T* p, T v, std::size_t n, hipStream_t s) {
  if (n == 0) return;
  const unsigned g = std::min<unsigned>(grid_for(n, 256), 65536u);
  fill_kernel<T><<<g, 256, 0, s>>>(p, v, n);
  CHUI_LAUNCH_CHECK("fill_kernel");
}

std::size_t tiles_for(std::size_t n) { return (n + kSortTile - 1) / kSortTile; }

template <class K>
void radix_sort_impl(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, std::size_t n, int begin_bit,
                     int end_bit, void* scratch, hipStream_t s) {
  if (n < 2 || end_bit <= begin_bit) return;
  const std::size_t tiles = tiles_for(n);
  uint32_t* hist = static_cast<uint32_t*>(scratch);
  void* scan_scratch = hist + tiles * 256;
  K* src_k = keys;
  uint32_t* src_v = vals;
  K* dst_k = keys_alt;
  uint32_t* dst_v = vals_alt;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    const int width = std::min(8, end_bit - shift);
    const uint32_t dmask = (1u << width) - 1u;
    sort_hist_kernel<K><<<static_cast<unsigned>(tiles), kSortThreads, 0, s>>>(src_k, n, shift, dmask, hist,
                                                                             static_cast<uint32_t>(tiles));
    CHUI_LAUNCH_CHECK("sort_hist_kernel");
    exclusive_scan_u32(hist, hist, tiles * 256, nullptr, scan_scratch, s);
    sort_scatter_kernel<K><<<static_cast<unsigned>(tiles), kSortThreads, 0, s>>>(
        src_k, src_v, dst_k, dst_v, n, shift, dmask, hist, static_cast<uint32_t>(tiles));
    CHUI_LAUNCH_CHECK("sort_scatter_kernel");
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  if (src_k != keys) {
    CHUI_HIP(hipMemcpyAsync(keys, src_k, n * sizeof(K), hipMemcpyDeviceToDevice, s));
    CHUI_HIP(hipMemcpyAsync(vals, src_v, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace

std::size_t scan_scratch_bytes(std::size_t n) {
  std::size_t total = 0;
  std::size_t len = n;
  while (len > static_cast<std::size_t>(kScanTile)) {
    len = (len + kScanTile - 1) / kScanTile;
    total += len;
  }
  return (total + 64) * sizeof(uint32_t);
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* total_dev, void* scratch,
                        hipStream_t s) {
  if (n == 0) {
    if (total_dev) fill_u32(total_dev, 0u, 1, s);
    return;
  }
  const std::size_t blocks = (n + kScanTile - 1) / kScanTile;
  if (blocks == 1) {
    scan_apply_kernel<<<1, kScanThreads, 0, s>>>(in, out, n, nullptr, total_dev);
    CHUI_LAUNCH_CHECK("scan_apply_kernel");
    return;
  }
  uint32_t* partial = static_cast<uint32_t*>(scratch);
  scan_reduce_kernel<<<static_cast<unsigned>(blocks), kScanThreads, 0, s>>>(in, n, partial);
  CHUI_LAUNCH_CHECK("scan_reduce_kernel");
  exclusive_scan_u32(partial, partial, blocks, nullptr, partial + blocks, s);
  scan_apply_kernel<<<static_cast<unsigned>(blocks), kScanThreads, 0, s>>>(in, out, n, partial, total_dev);
  CHUI_LAUNCH_CHECK("scan_apply_kernel");
}

std::size_t radix_sort_scratch_bytes(std::size_t n) {
  const std::size_t tiles = tiles_for(std::max<std::size_t>(n, 1));
  return tiles * 256 * sizeof(uint32_t) + scan_scratch_bytes(tiles * 256) + 256;
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, std::size_t n,
                      int begin_bit, int end_bit, void* scratch, hipStream_t s) {
  radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, scratch, s);
}
void radix_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, std::size_t n,
                        int begin_bit, int end_bit, void* scratch, hipStream_t s) {
  radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, begin_bit, end_bit, scratch, s);
}

void fill_u32(uint32_t* p, uint32_t v, std::size_t n, hipStream_t s) { fill_any(p, v, n, s); }
void fill_u64(uint64_t* p, uint64_t v, std::size_t n, hipStream_t s) { fill_any(p, v, n, s); }
void fill_f64(double* p, double v, std::size_t n, hipStream_t s) { fill_any(p, v, n, s); }
void fill_f32(float* p, float v, std::size_t n, hipStream_t s) { fill_any(p, v, n, s); }
void iota_u32(uint32_t* p, std::size_t n, hipStream_t s) {
  if (n == 0) return;
  iota_kernel<<<std::min<unsigned>(grid_for(n, 256), 65536u), 256, 0, s>>>(p, n);
  CHUI_LAUNCH_CHECK("iota_kernel");
}

}  // namespace chui

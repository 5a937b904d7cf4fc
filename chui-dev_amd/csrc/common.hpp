// common.hpp — shared host/device utilities of the MI355X TreePM library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

namespace chui {

// Exceptions used inside the library; the C ABI maps each class to a
// chui_status code (include/chui_treepm.h) so nothing crosses the boundary.
struct DeviceError : std::runtime_error { using std::runtime_error::runtime_error; };
struct UnsupportedError : std::runtime_error { using std::runtime_error::runtime_error; };
struct OutOfMemoryError : std::runtime_error { using std::runtime_error::runtime_error; };

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory)
    throw OutOfMemoryError(std::string(what) + ": " + hipGetErrorString(e));
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}
#define CHUI_HIP(call) ::chui::hip_check((call), #call)
#define CHUI_LAUNCH_CHECK(name) ::chui::hip_check(hipGetLastError(), name)

constexpr int kWave = 64;

inline unsigned grid_for(std::size_t n, unsigned block) {
  const std::size_t g = (n + block - 1) / block;
  return static_cast<unsigned>(g == 0 ? 1 : g);
}

// Grow-only device buffer. Allocation happens only when a call needs more
// capacity than any earlier call; steady-state solves allocate nothing.
struct DeviceBuffer {
  void* ptr = nullptr;
  std::size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void release() noexcept {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  void reserve(std::size_t need) {
    if (need <= bytes) return;
    release();
    CHUI_HIP(hipMalloc(&ptr, need));
    bytes = need;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Uploads of host tables synchronise the stream, which a stream recording a
// HIP graph does not allow: a capture-safe solve must find its tables built by
// an earlier solve with the same mesh and options.
inline void require_not_capturing(hipStream_t s, const char* what) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CHUI_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone)
    throw std::logic_error(std::string(what) +
                           " would be uploaded while a graph is recorded: run one solve with the same mesh and "
                           "options before capturing");
}

// Small device-to-host readbacks (counts, flags, bounds) through a pinned host
// block: the copy lands directly in host memory instead of going through the
// runtime's staging pass for pageable destinations.
struct PinnedReadback {
  static constexpr std::size_t kBytes = 256;
  void* ptr = nullptr;
  PinnedReadback() = default;
  PinnedReadback(const PinnedReadback&) = delete;
  PinnedReadback& operator=(const PinnedReadback&) = delete;
  ~PinnedReadback() {
    if (ptr) (void)hipHostFree(ptr);
  }
  // copies bytes (<= kBytes) from dev to host after the stream's prior work
  void read(void* host, const void* dev, std::size_t bytes, hipStream_t s) {
    if (bytes > kBytes) throw std::logic_error("pinned readback larger than its block");
    if (!ptr) CHUI_HIP(hipHostMalloc(&ptr, kBytes, hipHostMallocDefault));
    CHUI_HIP(hipMemcpyAsync(ptr, dev, bytes, hipMemcpyDeviceToHost, s));
    CHUI_HIP(hipStreamSynchronize(s));
    std::memcpy(host, ptr, bytes);
  }
};

// ---- device helpers -------------------------------------------------------
__device__ __forceinline__ uint64_t dbits(double v) { return static_cast<uint64_t>(__double_as_longlong(v)); }
__device__ __forceinline__ double bitsd(uint64_t b) { return __longlong_as_double(static_cast<long long>(b)); }

// Order-preserving map of any double onto uint64 (total order, -0 < +0).
__device__ __forceinline__ uint64_t order_key(double v) {
  const uint64_t b = dbits(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double order_key_inv(uint64_t k) {
  return bitsd((k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}
// Wave64 min/max of a 32-bit value through DPP lane moves (no LDS traffic):
// quad swaps, row rotations by 4 and 8, then the gfx9 row broadcasts of lanes
// 15 and 31; lane 63 ends with the result, returned as a wave-uniform value.
// All 64 lanes must be active.
__device__ __forceinline__ uint32_t dpp_min_step(uint32_t v, uint32_t o) { return o < v ? o : v; }
__device__ __forceinline__ uint32_t dpp_max_step(uint32_t v, uint32_t o) { return o > v ? o : v; }

#define CHUI_DPP(v, ctrl, rmask, ident) \
  static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(ident), static_cast<int>(v), ctrl, rmask, 0xf, false))

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = dpp_min_step(v, CHUI_DPP(v, 0xb1, 0xf, 0xFFFFFFFFu));   // quad_perm [1,0,3,2]
  v = dpp_min_step(v, CHUI_DPP(v, 0x4e, 0xf, 0xFFFFFFFFu));   // quad_perm [2,3,0,1]
  v = dpp_min_step(v, CHUI_DPP(v, 0x124, 0xf, 0xFFFFFFFFu));  // row_ror:4
  v = dpp_min_step(v, CHUI_DPP(v, 0x128, 0xf, 0xFFFFFFFFu));  // row_ror:8
  v = dpp_min_step(v, CHUI_DPP(v, 0x142, 0xa, 0xFFFFFFFFu));  // row_bcast:15
  v = dpp_min_step(v, CHUI_DPP(v, 0x143, 0xc, 0xFFFFFFFFu));  // row_bcast:31
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = dpp_max_step(v, CHUI_DPP(v, 0xb1, 0xf, 0u));
  v = dpp_max_step(v, CHUI_DPP(v, 0x4e, 0xf, 0u));
  v = dpp_max_step(v, CHUI_DPP(v, 0x124, 0xf, 0u));
  v = dpp_max_step(v, CHUI_DPP(v, 0x128, 0xf, 0u));
  v = dpp_max_step(v, CHUI_DPP(v, 0x142, 0xa, 0u));
  v = dpp_max_step(v, CHUI_DPP(v, 0x143, 0xc, 0u));
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
  v |= CHUI_DPP(v, 0xb1, 0xf, 0u);
  v |= CHUI_DPP(v, 0x4e, 0xf, 0u);
  v |= CHUI_DPP(v, 0x124, 0xf, 0u);
  v |= CHUI_DPP(v, 0x128, 0xf, 0u);
  v |= CHUI_DPP(v, 0x142, 0xa, 0u);
  v |= CHUI_DPP(v, 0x143, 0xc, 0u);
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}
__device__ __forceinline__ float readlane_f32(float v, uint32_t lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), static_cast<int>(lane)));
}
__device__ __forceinline__ double readlane_f64(double v, uint32_t lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xFFFFFFFFll), static_cast<int>(lane));
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), static_cast<int>(lane));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

}  // namespace chui

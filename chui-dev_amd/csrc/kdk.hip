// kdk.hip — FP64 drift/kick, deterministic FP64 mass/momentum reductions and
// the FP64 periodic Ewald accelerations used to measure force errors.
//
// Kick: v = drag * v + K * A    (gravity_runtime.cpp:3045-3058)
// Drift: x += v * D, canonical wrap into [0, L)  (reference_runtime_composition.cpp:59-108,
//        include/cosmosim/core/periodic_domain.hpp:8-16)
// Ewald: same sums as tests/support/periodic_ewald_reference.cpp:202-400; the
// reciprocal part goes through structure factors S(k) = sum_j m_j exp(i k.x_j)
// (sin(k.(x_j - x_i)) = Im(exp(-i k.x_i) S(k))), the real part through an
// LDS-tiled all-pairs loop over the requested images.
#include "kdk.hpp"

#include <cmath>
#include <mutex>

namespace chui {
namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kTwoOverSqrtPi = 1.12837916709551257389615890312154517;

__global__ void kick_kernel(double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                            const void* ax, const void* ay, const void* az, int is_float,
                            const uint32_t* __restrict__ active, std::size_t n, double kick, double drag) {
  const std::size_t s = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const std::size_t r = active ? active[s] : s;
  double a0, a1, a2;
  if (is_float) {
    a0 = static_cast<const float*>(ax)[s];
    a1 = static_cast<const float*>(ay)[s];
    a2 = static_cast<const float*>(az)[s];
  } else {
    a0 = static_cast<const double*>(ax)[s];
    a1 = static_cast<const double*>(ay)[s];
    a2 = static_cast<const double*>(az)[s];
  }
  vx[r] = drag * vx[r] + kick * a0;
  vy[r] = drag * vy[r] + kick * a1;
  vz[r] = drag * vz[r] + kick * a2;
}

__device__ __forceinline__ double canonical_position(double x, double L) {
  double w = fmod(x, L);
  if (w < 0.0) w += L;
  if (!(w < L)) w = 0.0;
  return w;
}

__global__ void drift_kernel(double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                             const double* __restrict__ vx, const double* __restrict__ vy,
                             const double* __restrict__ vz, std::size_t n, double d, double Lx, double Ly,
                             double Lz) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = canonical_position(x[i] + vx[i] * d, Lx);
  y[i] = canonical_position(y[i] + vy[i] * d, Ly);
  z[i] = canonical_position(z[i] + vz[i] * d, Lz);
}

// ---- CIC density contrast of unit-weight particles and its transpose ----------
// (the loss mesh of the differentiable forward; the stencil is the CIC of
// pm_solver.cpp:925-936 on its own n^3 mesh, cells row-major without padding)
__device__ __forceinline__ void cic_axis(double x, double L, int n, int (&i)[2], double (&w)[2], double& dw) {
  const double g = canonical_position(x, L) * (static_cast<double>(n) / L);
  const double base = floor(g);
  const long i0 = static_cast<long>(base);
  const double t = g - base;
  i[0] = static_cast<int>(((i0 % n) + n) % n);
  i[1] = (i[0] + 1) % n;
  w[0] = 1.0 - t;
  w[1] = t;
  dw = static_cast<double>(n) / L;  // d w1 / dx; d w0 / dx = -dw
}

__global__ void cic_density_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ z, std::size_t np, double L, int n, double scale,
                                   double* __restrict__ mesh) {
  const std::size_t p = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int ix[2], iy[2], iz[2];
  double wx[2], wy[2], wz[2], d;
  cic_axis(x[p], L, n, ix, wx, d);
  cic_axis(y[p], L, n, iy, wy, d);
  cic_axis(z[p], L, n, iz, wz, d);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        atomicAdd(&mesh[(static_cast<std::size_t>(ix[a]) * n + iy[b]) * n + iz[c]], scale * wx[a] * wy[b] * wz[c]);
}

__global__ void add_constant_kernel(double* __restrict__ v, std::size_t n, double c) {
  const std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] += c;
}

__global__ void cic_density_vjp_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                       const double* __restrict__ z, std::size_t np, double L, int n, double scale,
                                       const double* __restrict__ g, double* __restrict__ ox,
                                       double* __restrict__ oy, double* __restrict__ oz) {
  const std::size_t p = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int ix[2], iy[2], iz[2];
  double wx[2], wy[2], wz[2], dx, dy, dz;
  cic_axis(x[p], L, n, ix, wx, dx);
  cic_axis(y[p], L, n, iy, wy, dy);
  cic_axis(z[p], L, n, iz, wz, dz);
  const double sx[2] = {-dx, dx}, sy[2] = {-dy, dy}, sz[2] = {-dz, dz};
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const double v = g[(static_cast<std::size_t>(ix[a]) * n + iy[b]) * n + iz[c]];
        gx += sx[a] * wy[b] * wz[c] * v;
        gy += wx[a] * sy[b] * wz[c] * v;
        gz += wx[a] * wy[b] * sz[c] * v;
      }
  ox[p] = scale * gx;
  oy[p] = scale * gy;
  oz[p] = scale * gz;
}

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

__device__ double block_sum(double v) {
  __shared__ double part[kRedThreads / kWave];
  v = wave_sum(v);
  if (lane_id() == 0) part[threadIdx.x / kWave] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < kRedThreads / kWave; ++i) r += part[i];
  __syncthreads();
  return r;
}

// Fixed grid, fixed per-thread striding and a fixed tree of partial sums: the
// result does not depend on scheduling.
__global__ __launch_bounds__(kRedThreads) void mass_momentum_partial(const double* __restrict__ m,
                                                                     const double* __restrict__ vx,
                                                                     const double* __restrict__ vy,
                                                                     const double* __restrict__ vz, std::size_t n,
                                                                     double* __restrict__ partial) {
  double s[4] = {0, 0, 0, 0};
  for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    const double mm = m[i];
    s[0] += mm;
    s[1] += mm * vx[i];
    s[2] += mm * vy[i];
    s[3] += mm * vz[i];
  }
  for (int k = 0; k < 4; ++k) {
    const double b = block_sum(s[k]);
    if (threadIdx.x == 0) partial[k * gridDim.x + blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kRedThreads) void mass_momentum_final(const double* __restrict__ partial, int nb,
                                                                   double* __restrict__ out) {
  for (int k = 0; k < 4; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[k * nb + i];
    s = block_sum(s);
    if (threadIdx.x == 0) out[k] = s;
  }
}

// ---- Ewald -----------------------------------------------------------------------
struct EwaldArgs {
  double L[3];
  double G, alpha, eps;
  int real_images;
  int modes;
  uint32_t n_modes;   // half space (h > 0 lexicographically)
};

__device__ __forceinline__ void mode_of(uint32_t idx, int M, int& hx, int& hy, int& hz) {
  // enumerate the full cube [-M, M]^3 and keep the lexicographically positive half
  const int side = 2 * M + 1;
  const uint32_t half_start = static_cast<uint32_t>((side * side * side) / 2) + 1;  // index after (0,0,0)
  const uint32_t full = half_start + idx;
  hx = static_cast<int>(full / (side * side)) - M;
  hy = static_cast<int>((full / side) % side) - M;
  hz = static_cast<int>(full % side) - M;
}

constexpr int kSfThreads = 256;
constexpr uint32_t kSfChunk = 65536;

// Partial structure factors of one particle chunk; one thread per mode.
__global__ __launch_bounds__(kSfThreads) void structure_factor_kernel(EwaldArgs a, const double* __restrict__ x,
                                                                      const double* __restrict__ y,
                                                                      const double* __restrict__ z,
                                                                      const double* __restrict__ m, std::size_t n,
                                                                      double* __restrict__ S /* 2 per mode */) {
  const uint32_t mode = blockIdx.x * blockDim.x + threadIdx.x;
  const std::size_t p0 = static_cast<std::size_t>(blockIdx.y) * kSfChunk;
  const std::size_t p1 = p0 + kSfChunk < n ? p0 + kSfChunk : n;
  int hx = 0, hy = 0, hz = 0;
  const bool valid = mode < a.n_modes;
  if (valid) mode_of(mode, a.modes, hx, hy, hz);
  const double kx = 2.0 * kPi * hx / a.L[0], ky = 2.0 * kPi * hy / a.L[1], kz = 2.0 * kPi * hz / a.L[2];
  double re = 0.0, im = 0.0;
  for (std::size_t j = p0; j < p1; ++j) {
    const double ph = kx * x[j] + ky * y[j] + kz * z[j];
    double sn, cs;
    sincos(ph, &sn, &cs);
    re += m[j] * cs;
    im += m[j] * sn;
  }
  if (valid) {
    atomicAdd(&S[2 * mode], re);
    atomicAdd(&S[2 * mode + 1], im);
  }
}

// Reciprocal acceleration at each target: both h and -h contribute equally.
__global__ __launch_bounds__(256) void ewald_recip_kernel(EwaldArgs a, const double* __restrict__ x,
                                                          const double* __restrict__ y, const double* __restrict__ z,
                                                          const uint32_t* __restrict__ targets, std::size_t nt,
                                                          const double* __restrict__ S, double* __restrict__ ax,
                                                          double* __restrict__ ay, double* __restrict__ az) {
  const std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const uint32_t i = targets[t];
  const double V = a.L[0] * a.L[1] * a.L[2];
  const double pref = 4.0 * kPi * a.G / V;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (uint32_t md = 0; md < a.n_modes; ++md) {
    int hx, hy, hz;
    mode_of(md, a.modes, hx, hy, hz);
    const double kx = 2.0 * kPi * hx / a.L[0], ky = 2.0 * kPi * hy / a.L[1], kz = 2.0 * kPi * hz / a.L[2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double damp = exp(-k2 / (4.0 * a.alpha * a.alpha));
    double sn, cs;
    sincos(kx * x[i] + ky * y[i] + kz * z[i], &sn, &cs);
    // Im(exp(-i k.x_i) S) = cs * S_im - sn * S_re
    const double val = 2.0 * pref * damp * (cs * S[2 * md + 1] - sn * S[2 * md]) / k2;
    sx += val * kx;
    sy += val * ky;
    sz += val * kz;
  }
  ax[t] = sx;
  ay[t] = sy;
  az[t] = sz;
}

constexpr int kEwTile = 256;

// Real-space sum: one thread per target, sources streamed through LDS tiles.
__global__ __launch_bounds__(kEwTile) void ewald_real_kernel(EwaldArgs a, const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ m, std::size_t n,
                                                             const uint32_t* __restrict__ targets, std::size_t nt,
                                                             double* __restrict__ ax, double* __restrict__ ay,
                                                             double* __restrict__ az) {
  __shared__ double4 tile[kEwTile];
  const std::size_t t = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool valid = t < nt;
  const uint32_t self = valid ? targets[t] : 0xFFFFFFFFu;
  const double px = valid ? x[self] : 0.0, py = valid ? y[self] : 0.0, pz = valid ? z[self] : 0.0;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  const double e2 = a.eps * a.eps;
  for (std::size_t base = 0; base < n; base += kEwTile) {
    const std::size_t j = base + threadIdx.x;
    tile[threadIdx.x] = j < n ? make_double4(x[j], y[j], z[j], m[j]) : make_double4(0, 0, 0, 0);
    __syncthreads();
    const int cnt = static_cast<int>(n - base < kEwTile ? n - base : kEwTile);
    if (valid) {
      for (int q = 0; q < cnt; ++q) {
        const double4 s = tile[q];
        if (s.w == 0.0) continue;
        const bool is_self = (base + q) == self;
        double d[3] = {s.x - px, s.y - py, s.z - pz};
        for (int c = 0; c < 3; ++c) d[c] -= a.L[c] * rint(d[c] / a.L[c]);
        if (is_self) d[0] = d[1] = d[2] = 0.0;
        for (int ix = -a.real_images; ix <= a.real_images; ++ix)
          for (int iy = -a.real_images; iy <= a.real_images; ++iy)
            for (int iz = -a.real_images; iz <= a.real_images; ++iz) {
              if (is_self && ix == 0 && iy == 0 && iz == 0) continue;
              const double rx = d[0] + ix * a.L[0], ry = d[1] + iy * a.L[1], rz = d[2] + iz * a.L[2];
              const double r2 = rx * rx + ry * ry + rz * rz;
              const double r = sqrt(r2);
              const double ar = a.alpha * r;
              const double scr = erfc(ar) + kTwoOverSqrtPi * ar * exp(-ar * ar);
              double f = a.G * s.w * scr / (r2 * r);
              if (e2 > 0.0 && ix == 0 && iy == 0 && iz == 0) {
                const double s2 = r2 + e2;
                f += a.G * s.w * (1.0 / (s2 * sqrt(s2)) - 1.0 / (r2 * r));
              }
              sx += f * rx;
              sy += f * ry;
              sz += f * rz;
            }
      }
    }
    __syncthreads();
  }
  if (valid) {
    ax[t] += sx;
    ay[t] += sy;
    az[t] += sz;
  }
}

}  // namespace

void kick_device(double* vx, double* vy, double* vz, const void* ax, const void* ay, const void* az, int is_float,
                 const uint32_t* active, std::size_t n, double kick, double drag, hipStream_t s) {
  if (n == 0) return;
  kick_kernel<<<grid_for(n, 256), 256, 0, s>>>(vx, vy, vz, ax, ay, az, is_float, active, n, kick, drag);
  CHUI_LAUNCH_CHECK("kick_kernel");
}

void drift_device(double* x, double* y, double* z, const double* vx, const double* vy, const double* vz,
                  std::size_t n, double d, const double L[3], hipStream_t s) {
  if (n == 0) return;
  drift_kernel<<<grid_for(n, 256), 256, 0, s>>>(x, y, z, vx, vy, vz, n, d, L[0], L[1], L[2]);
  CHUI_LAUNCH_CHECK("drift_kernel");
}

void mass_momentum_device(const double* m, const double* vx, const double* vy, const double* vz, std::size_t n,
                          double* partial_scratch /* 4*1024 + 4 */, double out_host[4], hipStream_t s) {
  const int nb = kRedBlocks;
  mass_momentum_partial<<<nb, kRedThreads, 0, s>>>(m, vx, vy, vz, n, partial_scratch);
  CHUI_LAUNCH_CHECK("mass_momentum_partial");
  mass_momentum_final<<<1, kRedThreads, 0, s>>>(partial_scratch, nb, partial_scratch + 4 * nb);
  CHUI_LAUNCH_CHECK("mass_momentum_final");
  CHUI_HIP(hipMemcpyAsync(out_host, partial_scratch + 4 * nb, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  CHUI_HIP(hipStreamSynchronize(s));
}

// Allocation-free reduction: the partial sums live in a file-scope device
// array, one copy per GPU, instead of a buffer the context allocates. Host code
// never passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol); it passes the device address hipGetSymbolAddress returns,
// looked up once per device with that device current.
constexpr int kNoAllocPartials = 4 * kRedBlocks;
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("HIP device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return;
  int current = 0;
  CHUI_HIP(hipGetDevice(&current));
  if (current != device) throw std::logic_error("reduction scratch looked up with another device current");
  double* ptr = nullptr;
  CHUI_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch)));
  if (!ptr) throw DeviceError("hipGetSymbolAddress returned a null address for the reduction scratch");
  g_scratch_ptr[device] = ptr;
}

std::size_t noalloc_scratch_bytes() { return sizeof(double) * kNoAllocPartials; }

void mass_momentum_noalloc_device(int device, const double* m, const double* vx, const double* vy, const double* vz,
                                  std::size_t n, double* out_device, hipStream_t s) {
  double* partial = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device >= 0 && device < kMaxDevices) partial = g_scratch_ptr[device];
  }
  if (!partial) throw std::logic_error("reduction scratch of this device was not set up");
  // Same grid and summation tree as mass_momentum_device; every partial slot
  // is written before it is read, so the array needs no clearing.
  static_assert(4 * kRedBlocks <= kNoAllocPartials, "partials of the fixed grid must fit the scratch array");
  mass_momentum_partial<<<kRedBlocks, kRedThreads, 0, s>>>(m, vx, vy, vz, n, partial);
  CHUI_LAUNCH_CHECK("mass_momentum_partial");
  mass_momentum_final<<<1, kRedThreads, 0, s>>>(partial, kRedBlocks, out_device);
  CHUI_LAUNCH_CHECK("mass_momentum_final");
}

std::size_t ewald_scratch_bytes(int modes) {
  const std::size_t side = 2 * static_cast<std::size_t>(modes) + 1;
  const std::size_t half = (side * side * side) / 2;
  return (2 * half + 8) * sizeof(double);
}

void cic_density_device(const double* x, const double* y, const double* z, std::size_t np, double L, int n,
                        double* delta, hipStream_t s) {
  const std::size_t cells = static_cast<std::size_t>(n) * n * n;
  fill_f64(delta, 0.0, cells, s);
  if (np == 0) return;
  const double scale = static_cast<double>(cells) / static_cast<double>(np);  // 1 / mean count
  cic_density_kernel<<<grid_for(np, 256), 256, 0, s>>>(x, y, z, np, L, n, scale, delta);
  CHUI_LAUNCH_CHECK("cic_density_kernel");
  add_constant_kernel<<<grid_for(cells, 256), 256, 0, s>>>(delta, cells, -1.0);
  CHUI_LAUNCH_CHECK("add_constant_kernel");
}

void cic_density_vjp_device(const double* x, const double* y, const double* z, std::size_t np, double L, int n,
                            const double* g, double* ox, double* oy, double* oz, hipStream_t s) {
  if (np == 0) return;
  const double scale = static_cast<double>(n) * n * n / static_cast<double>(np);
  cic_density_vjp_kernel<<<grid_for(np, 256), 256, 0, s>>>(x, y, z, np, L, n, scale, g, ox, oy, oz);
  CHUI_LAUNCH_CHECK("cic_density_vjp_kernel");
}

void ewald_device(const double* x, const double* y, const double* z, const double* m, std::size_t n,
                  const uint32_t* targets, std::size_t nt, const double L[3], double G, double alpha,
                  int real_images, int modes, double eps, double* ax, double* ay, double* az, void* scratch,
                  hipStream_t s) {
  EwaldArgs a{};
  for (int c = 0; c < 3; ++c) a.L[c] = L[c];
  a.G = G;
  a.alpha = alpha;
  a.eps = eps;
  a.real_images = real_images;
  a.modes = modes;
  const std::size_t side = 2 * static_cast<std::size_t>(modes) + 1;
  a.n_modes = static_cast<uint32_t>((side * side * side) / 2);
  double* S = static_cast<double*>(scratch);
  CHUI_HIP(hipMemsetAsync(S, 0, 2 * static_cast<std::size_t>(a.n_modes) * sizeof(double), s));
  if (a.n_modes > 0 && n > 0) {
    dim3 grid(grid_for(a.n_modes, kSfThreads), static_cast<unsigned>((n + kSfChunk - 1) / kSfChunk));
    structure_factor_kernel<<<grid, kSfThreads, 0, s>>>(a, x, y, z, m, n, S);
    CHUI_LAUNCH_CHECK("structure_factor_kernel");
  }
  if (nt == 0) return;
  ewald_recip_kernel<<<grid_for(nt, 256), 256, 0, s>>>(a, x, y, z, targets, nt, S, ax, ay, az);
  CHUI_LAUNCH_CHECK("ewald_recip_kernel");
  ewald_real_kernel<<<grid_for(nt, kEwTile), kEwTile, 0, s>>>(a, x, y, z, m, n, targets, nt, ax, ay, az);
  CHUI_LAUNCH_CHECK("ewald_real_kernel");
}

}  // namespace chui

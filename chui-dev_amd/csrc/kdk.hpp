// kdk.hpp — FP64 integrator pieces, reductions and the Ewald accuracy oracle.
#pragma once

#include "common.hpp"
#include "primitives.hpp"

namespace chui {

void kick_device(double* vx, double* vy, double* vz, const void* ax, const void* ay, const void* az, int is_float,
                 const uint32_t* active, std::size_t n, double kick, double drag, hipStream_t s);
void drift_device(double* x, double* y, double* z, const double* vx, const double* vy, const double* vz,
                  std::size_t n, double d, const double L[3], hipStream_t s);
// partial_scratch: >= (4*1024 + 4) doubles of device memory.
void mass_momentum_device(const double* m, const double* vx, const double* vy, const double* vz, std::size_t n,
                          double* partial_scratch, double out_host[4], hipStream_t s);

// Allocation-free form: partial sums in the library's per-GPU device array
// (noalloc_scratch_init must have run once with `device` current), result in
// out_device[0..3] (device memory), stream-ordered, no synchronisation.
void noalloc_scratch_init(int device);
std::size_t noalloc_scratch_bytes();
void mass_momentum_noalloc_device(int device, const double* m, const double* vx, const double* vy, const double* vz,
                                  std::size_t n, double* out_device, hipStream_t s);

std::size_t ewald_scratch_bytes(int modes);
// delta(c) = count(c) * n^3 / np - 1 on an n^3 mesh (row-major, no padding),
// CIC weights, unit-weight particles.
void cic_density_device(const double* x, const double* y, const double* z, std::size_t np, double L, int n,
                        double* delta, hipStream_t s);
// out_p = d/dx_p sum_c g(c) delta(c) for the density above.
void cic_density_vjp_device(const double* x, const double* y, const double* z, std::size_t np, double L, int n,
                            const double* g, double* ox, double* oy, double* oz, hipStream_t s);
void ewald_device(const double* x, const double* y, const double* z, const double* m, std::size_t n,
                  const uint32_t* targets, std::size_t nt, const double L[3], double G, double alpha,
                  int real_images, int modes, double eps, double* ax, double* ay, double* az, void* scratch,
                  hipStream_t s);

}  // namespace chui

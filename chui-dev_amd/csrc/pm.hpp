// pm.hpp — periodic particle-mesh long-range solve on the device.
#pragma once

#include <vector>

#include "common.hpp"

namespace chui {

struct PmGeometry {
  int n[3];          // mesh (powers of two)
  int lg[3];         // log2(n)
  double L[3];       // box lengths
  std::size_t nzc;   // nz/2 + 1
  std::size_t row;   // padded real row stride = 2 * nzc
  std::size_t complex_cells() const { return static_cast<std::size_t>(n[0]) * n[1] * nzc; }
};

struct PmOptions {
  double G = 1.0;
  double split_scale = 0.0;
  int assignment = 0;      // 0 CIC, 1 TSC
  int deconvolution = 0;
  int fft_lines = 0;       // lines per block of the line passes (0: the most the LDS budget allows)
};

// Device mesh workspace for one precision. Three complex buffers; the real
// force meshes are written in place (padded rows) at the end of the solve.
struct PmWorkspace {
  DeviceBuffer buf[3];
  DeviceBuffer tables;   // k, derivative-k, window tables + twiddles
  // the tables' identity (mesh axes, box lengths, assignment, deconvolution)
  // and element offsets: they are uploaded again only when it changes
  double table_key[8] = {0, 0, 0, 0, 0, 0, -1, -1};
  std::size_t table_off[12] = {};
  std::size_t cells = 0;
  int elem_bytes = 0;
  // After a solve: real force meshes (padded stride row) per axis.
  void* force[3] = {nullptr, nullptr, nullptr};
  bool valid = false;
  // identity of the field (cadence reuse): mesh-level options, then the
  // caller's source generation, field version and force epoch (sequence,
  // scale factor)
  double key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint64_t source_generation = 0, field_version = 0, epoch_sequence = 0;
  double epoch_scale_factor = 0.0;
};

// Mesh-deposit work: subtrees whose particles (contiguous in tree order) all
// land on a small block of mesh cells. A unit with dims[0] == 0 has no such
// block and is deposited straight into the mesh. Each unit is split into
// chunks of kDepositChunk particles, one thread block per chunk;
// first_block[u] is the first block of unit u.
struct DepositUnit {
  uint32_t begin, count;
  int32_t lo[3];    // first cell of the block (wrapped into the mesh)
  int32_t dims[3];  // cells per axis
};
struct DepositUnits {
  const DepositUnit* units = nullptr;
  const uint32_t* first_block = nullptr;
  uint32_t count = 0, blocks = 0;
  // Capture-safe: {unit count, block count} on the device; the deposit then
  // launches `blocks` (a bound) blocks that loop over the real block count.
  const uint32_t* dev_counts = nullptr;
};
// What the unit builder needs to know about the mesh.
struct DepositTiling {
  int n[3];
  double inv_h[3];    // n / L
  int scheme;         // 0 CIC, 1 TSC
  uint32_t max_cells; // LDS tile capacity
  uint32_t chunk;
};
constexpr uint32_t kDepositChunk = 2048;
constexpr uint32_t kDepositSafeGrid = 2048;  // deposit blocks of a capture-safe solve (strided over the real count)
constexpr uint32_t kDepositTileBytes = 49152;

// Deposit (tree-ordered particles), forward r2c FFT, Green's function with the
// Gaussian long-range filter, spectral gradient and three c2r inverses.
// FP64: packed (x, y, z, m) caller doubles in tree order; uniform_mass != 0
// means every source carries mass um.
void pm_solve_fp64(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const double4* src, int uniform_mass,
                   double um, std::size_t n, const DepositUnits& units, hipStream_t s);
// Adjoint of the PM force field for the force VJP (FP64): deposits the three
// components of lam (comp[a] = tree-ordered (x, y, z, lam_a)), applies the
// transposed operator and leaves the real adjoint potential
// nu = sum_a S_a^T mu_a in buf[0] (padded rows), with S_a the operator that
// maps the deposited mass mesh to force component a.
void pm_adjoint_fp64(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const double4* const comp[3],
                     std::size_t n, const DepositUnits& units, hipStream_t s);
// Mixed: packed 32-bit fixed-point periodic coordinates + float mass bits,
// FP32 mesh and FFTs.
void pm_solve_mixed(PmWorkspace& w, const PmGeometry& g, const PmOptions& o, const uint4* src, int uniform_mass,
                    float um, std::size_t n, const DepositUnits& units, hipStream_t s);

// Stage timers (ms) of the last solve, filled when events are enabled.
struct PmTimings {
  double deposit_ms = 0, fft_ms = 0;
};

}  // namespace chui

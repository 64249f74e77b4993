// meshgen.hip -- the seeded jittered quad and hex meshes generated on the
// device, bit for bit the meshes dpg_mesh_quad / dpg_mesh_hex build on the
// host (the input generation of the reference's drivers, Mesh.cpp:591-614,
// for the structured meshes here).
//
// The host draws the jitter of the interior vertices from one xoshiro256**
// stream in vertex order (x fastest), `dim` uniforms per vertex. The state
// update of xoshiro256** is linear over GF(2): after n steps the state is
// T^n s0. The library keeps the columns of T^(2^b) (b < 48) in device memory;
// a thread reaches its vertex's first draw by applying the matrices of the set
// bits of n, then draws as the host does. Coordinates are formed with the
// host's operations in the host's order, without contraction into FMAs, so
// every double is the host's.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dpg_capi.h"
#include "errors.hpp"
#include "rng.hpp"

#define MG_CHECK(expr)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace dpg {
namespace {

constexpr int kJumpLevels = 48;  // jumps up to 2^48 - 1 steps
constexpr int kMaxDevices = 64;

// jump_cols[b][i] = T^(2^b) e_i, the 256-bit state as 4 words
__device__ uint64_t jump_cols[kJumpLevels][256][4];

std::mutex g_mu;
bool g_ready[kMaxDevices] = {};

void apply_host(const uint64_t (*M)[4], const uint64_t v[4], uint64_t r[4]) {
  r[0] = r[1] = r[2] = r[3] = 0;
  for (int i = 0; i < 256; ++i)
    if ((v[i >> 6] >> (i & 63)) & 1ull)
      for (int w = 0; w < 4; ++w) r[w] ^= M[i][w];
}

// the jump matrices of this device, built on the host once per device
void ensure_jump_tables() {
  int dev = -1;
  MG_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ready[dev]) return;
  std::vector<uint64_t> J((size_t)kJumpLevels * 256 * 4, 0);
  auto col = [&](int b, int i) { return &J[((size_t)b * 256 + i) * 4]; };
  for (int i = 0; i < 256; ++i) {
    uint64_t s[4] = {0, 0, 0, 0};
    s[i >> 6] = 1ull << (i & 63);
    Rng::step(s);
    for (int w = 0; w < 4; ++w) col(0, i)[w] = s[w];
  }
  for (int b = 1; b < kJumpLevels; ++b)
    for (int i = 0; i < 256; ++i)
      apply_host(reinterpret_cast<const uint64_t(*)[4]>(col(b - 1, 0)), col(b - 1, i), col(b, i));
  MG_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(jump_cols), J.data(), J.size() * sizeof(uint64_t)));
  g_ready[dev] = true;
}

// s <- T^n s
__device__ inline void jump_ahead(uint64_t s[4], uint64_t n) {
  for (int b = 0; n; ++b, n >>= 1) {
    if (!(n & 1ull)) continue;
    uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    for (int i = 0; i < 256; ++i) {
      const uint64_t m = 0ull - ((s[i >> 6] >> (i & 63)) & 1ull);
      const uint64_t* c = jump_cols[b][i];
      r0 ^= c[0] & m;
      r1 ^= c[1] & m;
      r2 ^= c[2] & m;
      r3 ^= c[3] & m;
    }
    s[0] = r0;
    s[1] = r1;
    s[2] = r2;
    s[3] = r3;
  }
}

__device__ inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ inline uint64_t next_draw(uint64_t s[4]) {
  const uint64_t r = rotl64(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return r;
}

struct Seed {
  uint64_t s[4];
};

// vertex coordinates of an nx x ny (x nz) grid on the unit square / cube,
// interior vertices jittered by U(-jitter, jitter) * h per coordinate
template <int DIM>
__global__ void jittered_vertices(Seed seed, int nx, int ny, int nz, double jitter, double hx, double hy, double hz,
                                  double* __restrict__ X) {
#pragma clang fp contract(off)
  const int64_t nvx = nx + 1, nvy = ny + 1, nvz = DIM == 3 ? nz + 1 : 1;
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvx * nvy * nvz) return;
  const int i = (int)(v % nvx), j = (int)((v / nvx) % nvy), k = DIM == 3 ? (int)(v / (nvx * nvy)) : 0;
  const double h[3] = {hx, hy, hz};
  double p[3] = {i * hx, j * hy, k * hz};
  bool interior = i > 0 && i < nx && j > 0 && j < ny;
  if (DIM == 3) interior = interior && k > 0 && k < nz;
  if (interior) {
    // interior vertices before this one, in the host's loop order
    const int64_t idx = DIM == 3 ? ((int64_t)(k - 1) * (ny - 1) + (j - 1)) * (nx - 1) + (i - 1)
                                 : (int64_t)(j - 1) * (nx - 1) + (i - 1);
    uint64_t s[4] = {seed.s[0], seed.s[1], seed.s[2], seed.s[3]};
    jump_ahead(s, (uint64_t)idx * DIM);
    const double lo = -jitter, hi = jitter;
    for (int d = 0; d < DIM; ++d) {
      const double u = (double)(next_draw(s) >> 11) * (1.0 / 9007199254740992.0);
      const double r = lo + (hi - lo) * u;
      p[d] += r * h[d];
    }
  }
  for (int d = 0; d < DIM; ++d) X[v * DIM + d] = p[d];
}

// nodes (C, V, DIM) in the cell vertex order of the host generators and the
// side parities (C, S) of the structured grid (+1 on the boundary and toward
// the higher cell number)
template <int DIM>
__global__ void grid_cells(int nx, int ny, int nz, const double* __restrict__ X, double* __restrict__ nodes,
                           int8_t* __restrict__ parities) {
  constexpr int V = DIM == 3 ? 8 : 4, S = DIM == 3 ? 6 : 4;
  const int64_t C = (int64_t)nx * ny * (DIM == 3 ? nz : 1);
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = DIM == 3 ? (int)(c / ((int64_t)nx * ny)) : 0;
  const int64_t nvx = nx + 1, nvy = ny + 1;
  // quads CCW; hexes in the shards Hexahedron<8> order
  const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int a = 0; a < V; ++a) {
    const int64_t vid = ((int64_t)(k + off[a][2]) * nvy + (j + off[a][1])) * nvx + (i + off[a][0]);
    for (int d = 0; d < DIM; ++d) nodes[(c * V + a) * DIM + d] = X[vid * DIM + d];
  }
  int8_t* p = parities + c * S;
  p[0] = (j == 0) ? 1 : -1;  // -y
  p[1] = 1;                  // +x
  p[2] = 1;                  // +y
  p[3] = (i == 0) ? 1 : -1;  // -x
  if (DIM == 3) {
    p[4] = (k == 0) ? 1 : -1;  // -z
    p[5] = 1;                  // +z
  }
}

void generate(int dim, int nx, int ny, int nz, double jitter, uint64_t seed, double* nodes, int8_t* parities,
              hipStream_t st) {
  if (nx < 1 || ny < 1 || nz < 1 || !nodes || !parities) throw std::invalid_argument("bad mesh arguments");
  if (!(jitter >= 0.0 && jitter < 0.5)) throw std::invalid_argument("jitter must be in [0, 0.5)");
  const int64_t nv = (int64_t)(nx + 1) * (ny + 1) * (dim == 3 ? nz + 1 : 1);
  const int64_t C = (int64_t)nx * ny * (dim == 3 ? nz : 1);
  if (nv * dim >= (1LL << 47)) throw std::invalid_argument("mesh too large for the jump tables");
  ensure_jump_tables();
  Rng r(seed);
  Seed s0;
  for (int w = 0; w < 4; ++w) s0.s[w] = r.state()[w];
  // the same spacings the host computes
  const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
  double* X = nullptr;
  MG_CHECK(hipMalloc((void**)&X, sizeof(double) * nv * dim));
  const unsigned gv = (unsigned)((nv + 255) / 256), gc = (unsigned)((C + 255) / 256);
  if (dim == 3) {
    jittered_vertices<3><<<gv, 256, 0, st>>>(s0, nx, ny, nz, jitter, hx, hy, hz, X);
    grid_cells<3><<<gc, 256, 0, st>>>(nx, ny, nz, X, nodes, parities);
  } else {
    jittered_vertices<2><<<gv, 256, 0, st>>>(s0, nx, ny, 1, jitter, hx, hy, 0.0, X);
    grid_cells<2><<<gc, 256, 0, st>>>(nx, ny, 1, X, nodes, parities);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(X);
  if (e != hipSuccess) throw std::runtime_error(std::string("mesh generation kernels: ") + hipGetErrorString(e));
}

}  // namespace
}  // namespace dpg

extern "C" {

int dpg_mesh_quad_device(int nx, int ny, double jitter, uint64_t seed, double* nodes, int8_t* parities,
                         void* stream) {
  return dpg::guarded([&] { dpg::generate(2, nx, ny, 1, jitter, seed, nodes, parities, (hipStream_t)stream); });
}

int dpg_mesh_hex_device(int nx, int ny, int nz, double jitter, uint64_t seed, double* nodes, int8_t* parities,
                        void* stream) {
  return dpg::guarded([&] { dpg::generate(3, nx, ny, nz, jitter, seed, nodes, parities, (hipStream_t)stream); });
}

}  // extern "C"

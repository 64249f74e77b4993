// inputs.cpp -- seeded synthetic inputs: portable RNG, jittered meshes, loads.
//
// RNG: splitmix64 seeding xoshiro256**, uniforms from the top 53 bits,
// normals by Box-Muller, so every host produces the same inputs.
// Meshes number cells x fastest; side parities follow the rule of
// GlobalDofAssignment.cpp:160-257 (+1 on the boundary, +1 when the cell ID is
// smaller than the neighbour's, -1 otherwise).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/dpg_capi.h"
#include "errors.hpp"

namespace dpg {

class Rng {
 public:
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s_[i] = splitmix(x);
  }
  uint64_t next() {
    uint64_t r = rotl(s_[1] * 5, 7) * 9;
    uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  double normal() {
    if (has_spare_) {
      has_spare_ = false;
      return spare_;
    }
    double u1 = 1.0 - uniform();  // (0,1]
    double u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    spare_ = r * std::sin(2.0 * M_PI * u2);
    has_spare_ = true;
    return r * std::cos(2.0 * M_PI * u2);
  }
  int integer(int lo, int hi) {  // inclusive
    return lo + (int)(uniform() * (hi - lo + 1));
  }

 private:
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  uint64_t s_[4];
  bool has_spare_ = false;
  double spare_ = 0.0;
};

}  // namespace dpg

extern "C" {

int dpg_rng_uniform(uint64_t seed, double lo, double hi, double* out, size_t n) {
  return dpg::guarded([&] {
    if (!out && n) throw std::invalid_argument("out is NULL");
    dpg::Rng r(seed);
    for (size_t i = 0; i < n; ++i) out[i] = r.uniform(lo, hi);
  });
}

int dpg_rng_normal(uint64_t seed, double* out, size_t n) {
  return dpg::guarded([&] {
    if (!out && n) throw std::invalid_argument("out is NULL");
    dpg::Rng r(seed);
    for (size_t i = 0; i < n; ++i) out[i] = r.normal();
  });
}

int dpg_mesh_quad(int nx, int ny, double jitter, uint64_t seed, double* nodes, int8_t* parities) {
  return dpg::guarded([&] {
    if (nx < 1 || ny < 1 || !nodes || !parities) throw std::invalid_argument("bad mesh arguments");
    if (!(jitter >= 0.0 && jitter < 0.5)) throw std::invalid_argument("jitter must be in [0, 0.5)");
    dpg::Rng r(seed);
    double hx = 1.0 / nx, hy = 1.0 / ny;
    std::vector<double> X((size_t)(nx + 1) * (ny + 1) * 2);
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        double* p = &X[((size_t)j * (nx + 1) + i) * 2];
        p[0] = i * hx;
        p[1] = j * hy;
        if (i > 0 && i < nx && j > 0 && j < ny) {
          p[0] += r.uniform(-jitter, jitter) * hx;
          p[1] += r.uniform(-jitter, jitter) * hy;
        }
      }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = (size_t)j * nx + i;
        const int vi[4][2] = {{i, j}, {i + 1, j}, {i + 1, j + 1}, {i, j + 1}};
        for (int v = 0; v < 4; ++v)
          for (int d = 0; d < 2; ++d) nodes[(c * 4 + v) * 2 + d] = X[((size_t)vi[v][1] * (nx + 1) + vi[v][0]) * 2 + d];
        // sides: 0 bottom (j-1), 1 right (i+1), 2 top (j+1), 3 left (i-1)
        parities[c * 4 + 0] = (j == 0) ? 1 : -1;
        parities[c * 4 + 1] = 1;
        parities[c * 4 + 2] = 1;
        parities[c * 4 + 3] = (i == 0) ? 1 : -1;
      }
  });
}

int dpg_mesh_hex(int nx, int ny, int nz, double jitter, uint64_t seed, double* nodes, int8_t* parities) {
  return dpg::guarded([&] {
    if (nx < 1 || ny < 1 || nz < 1 || !nodes || !parities) throw std::invalid_argument("bad mesh arguments");
    if (!(jitter >= 0.0 && jitter < 0.5)) throw std::invalid_argument("jitter must be in [0, 0.5)");
    dpg::Rng r(seed);
    double h[3] = {1.0 / nx, 1.0 / ny, 1.0 / nz};
    size_t nvx = nx + 1, nvy = ny + 1;
    std::vector<double> X(nvx * nvy * (nz + 1) * 3);
    for (int k = 0; k <= nz; ++k)
      for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
          double* p = &X[(((size_t)k * nvy + j) * nvx + i) * 3];
          p[0] = i * h[0];
          p[1] = j * h[1];
          p[2] = k * h[2];
          if (i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz)
            for (int d = 0; d < 3; ++d) p[d] += r.uniform(-jitter, jitter) * h[d];
        }
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          size_t c = ((size_t)k * ny + j) * nx + i;
          // shards Hexahedron<8> vertex order
          const int vi[8][3] = {{i, j, k},         {i + 1, j, k},         {i + 1, j + 1, k},     {i, j + 1, k},
                                {i, j, k + 1},     {i + 1, j, k + 1},     {i + 1, j + 1, k + 1}, {i, j + 1, k + 1}};
          for (int v = 0; v < 8; ++v)
            for (int d = 0; d < 3; ++d)
              nodes[(c * 8 + v) * 3 + d] = X[(((size_t)vi[v][2] * nvy + vi[v][1]) * nvx + vi[v][0]) * 3 + d];
          // faces: 0 -y, 1 +x, 2 +y, 3 -x, 4 -z, 5 +z
          int8_t* p = parities + c * 6;
          p[0] = (j == 0) ? 1 : -1;
          p[1] = 1;
          p[2] = 1;
          p[3] = (i == 0) ? 1 : -1;
          p[4] = (k == 0) ? 1 : -1;
          p[5] = 1;
        }
  });
}

int dpg_load_sine_modes(int spatial_dim, int ncomp, int nmodes, int kmax, uint64_t seed, dpg_load_desc* out) {
  return dpg::guarded([&] {
    if (!out || ncomp < 0 || ncomp > 3 || nmodes < 0 || nmodes > DPG_MAX_LOAD_MODES || kmax < 1 ||
        spatial_dim < 1 || spatial_dim > 3)
      throw std::invalid_argument("bad load arguments");
    std::memset(out, 0, sizeof(*out));
    dpg::Rng r(seed);
    out->ncomp = ncomp;
    for (int c = 0; c < ncomp; ++c) {
      out->nmodes[c] = nmodes;
      for (int m = 0; m < nmodes; ++m) {
        out->amp[c][m] = r.normal();
        for (int d = 0; d < spatial_dim; ++d) out->k[c][m][d] = r.integer(1, kmax);
      }
    }
  });
}

}  // extern "C"

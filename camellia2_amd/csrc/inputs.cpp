// inputs.cpp -- seeded synthetic inputs: portable RNG, jittered meshes, loads.
//
// RNG: splitmix64 seeding xoshiro256**, uniforms from the top 53 bits,
// normals by Box-Muller, so every host produces the same inputs.
// Meshes number cells x fastest; side parities follow the rule of
// GlobalDofAssignment.cpp:160-257 (+1 on the boundary, +1 when the cell ID is
// smaller than the neighbour's, -1 otherwise).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "../../include/dpg_capi.h"
#include "errors.hpp"
#include "rng.hpp"

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

int dpg_mesh_quad_graded(int n0, int levels, double jitter, uint64_t seed, double* nodes, int8_t* parities,
                         int32_t* level) {
  return dpg::guarded_status([&]() -> int {
    if (n0 < 1 || levels < 0 || levels > 20) throw std::invalid_argument("bad graded mesh arguments");
    if (!(jitter >= 0.0 && jitter < 0.25)) throw std::invalid_argument("jitter must be in [0, 0.25)");
    int64_t count = (int64_t)n0 * n0 + 3LL * n0 * ((1LL << levels) - 1);
    if (count > (1LL << 28)) throw std::invalid_argument("graded mesh too large");
    if (!nodes) return (int)count;
    if (!parities) throw std::invalid_argument("parities is NULL");
    // integer coordinates in units of the finest h: a level-l cell has edge U >> l
    const int64_t U = 1LL << levels, W = (int64_t)n0 * U;
    struct Cell {
      int lvl;
      int64_t x0, y0, sz;
    };
    std::vector<Cell> leaves, next;
    for (int j = 0; j < n0; ++j)
      for (int i = 0; i < n0; ++i) leaves.push_back({0, i * U, j * U, U});
    for (int l = 0; l < levels; ++l) {  // refine the column touching x = 1
      next.clear();
      for (const Cell& c : leaves) {
        if (c.lvl == l && c.x0 + c.sz == W) {
          int64_t h = c.sz / 2;
          for (int q = 0; q < 4; ++q) next.push_back({l + 1, c.x0 + (q & 1) * h, c.y0 + (q >> 1) * h, h});
        } else {
          next.push_back(c);
        }
      }
      leaves.swap(next);
    }
    std::sort(leaves.begin(), leaves.end(), [](const Cell& a, const Cell& b) {
      if (a.lvl != b.lvl) return a.lvl < b.lvl;
      if (a.y0 != b.y0) return a.y0 < b.y0;
      return a.x0 < b.x0;
    });
    if ((int64_t)leaves.size() != count) throw std::logic_error("graded mesh cell count");
    auto key = [](int lvl, int64_t ix, int64_t iy) { return ((uint64_t)lvl << 58) | ((uint64_t)ix << 29) | (uint64_t)iy; };
    std::unordered_map<uint64_t, int> id;
    for (size_t c = 0; c < leaves.size(); ++c)
      id[key(leaves[c].lvl, leaves[c].x0 / leaves[c].sz, leaves[c].y0 / leaves[c].sz)] = (int)c;
    // leaf containing a point given in quarter units of the finest h (-1 outside)
    auto locate = [&](int64_t qx, int64_t qy) -> int {
      if (qx < 0 || qy < 0 || qx >= 4 * W || qy >= 4 * W) return -1;
      for (int l = 0; l <= levels; ++l) {
        int64_t sz = 4 * (U >> l);
        auto it = id.find(key(l, qx / sz, qy / sz));
        if (it != id.end()) return it->second;
      }
      throw std::logic_error("graded mesh point location");
    };
    // vertices: smallest adjacent cell edge, and hanging (interior of a coarser edge)
    auto vkey = [&](int64_t x, int64_t y) { return (uint64_t)x * (uint64_t)(W + 1) + (uint64_t)y; };
    std::unordered_map<uint64_t, int64_t> hmin;
    std::unordered_map<uint64_t, std::pair<uint64_t, uint64_t>> hanging;  // -> endpoints of its coarse edge
    for (const Cell& c : leaves) {
      const int64_t vx[4] = {c.x0, c.x0 + c.sz, c.x0 + c.sz, c.x0}, vy[4] = {c.y0, c.y0, c.y0 + c.sz, c.y0 + c.sz};
      for (int v = 0; v < 4; ++v) {
        auto it = hmin.find(vkey(vx[v], vy[v]));
        if (it == hmin.end()) hmin[vkey(vx[v], vy[v])] = c.sz;
        else it->second = std::min(it->second, c.sz);
      }
    }
    for (const Cell& c : leaves) {
      if (c.sz < 2) continue;
      int64_t h = c.sz / 2;
      const int64_t mx[4] = {c.x0 + h, c.x0 + c.sz, c.x0 + h, c.x0}, my[4] = {c.y0, c.y0 + h, c.y0 + c.sz, c.y0 + h};
      const int64_t vx[4] = {c.x0, c.x0 + c.sz, c.x0 + c.sz, c.x0}, vy[4] = {c.y0, c.y0, c.y0 + c.sz, c.y0 + c.sz};
      for (int s = 0; s < 4; ++s)
        if (hmin.count(vkey(mx[s], my[s])))
          hanging[vkey(mx[s], my[s])] = {vkey(vx[s], vy[s]), vkey(vx[(s + 1) % 4], vy[(s + 1) % 4])};
    }
    // jitter of regular interior vertices, drawn in vertex-key order
    std::vector<uint64_t> keys;
    keys.reserve(hmin.size());
    for (auto& kv : hmin) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    std::unordered_map<uint64_t, std::pair<double, double>> disp;
    dpg::Rng r(seed);
    for (uint64_t k : keys) {
      int64_t x = (int64_t)(k / (uint64_t)(W + 1)), y = (int64_t)(k % (uint64_t)(W + 1));
      if (x == 0 || y == 0 || x == W || y == W || hanging.count(k)) continue;
      double h = (double)hmin[k] / (double)W;
      double dx = r.uniform(-jitter, jitter) * h, dy = r.uniform(-jitter, jitter) * h;
      disp[k] = {dx, dy};
    }
    // a hanging vertex stays on its coarse edge: the mean of the endpoints' moves
    // (the endpoints of a coarse edge are never hanging in this 2:1 mesh)
    for (auto& kv : hanging) {
      auto a = disp.find(kv.second.first), b = disp.find(kv.second.second);
      double dx = 0.0, dy = 0.0;
      if (a != disp.end()) {
        dx += 0.5 * a->second.first;
        dy += 0.5 * a->second.second;
      }
      if (b != disp.end()) {
        dx += 0.5 * b->second.first;
        dy += 0.5 * b->second.second;
      }
      if (dx != 0.0 || dy != 0.0) disp[kv.first] = {dx, dy};
    }
    for (size_t c = 0; c < leaves.size(); ++c) {
      const Cell& L = leaves[c];
      const int64_t vx[4] = {L.x0, L.x0 + L.sz, L.x0 + L.sz, L.x0}, vy[4] = {L.y0, L.y0, L.y0 + L.sz, L.y0 + L.sz};
      for (int v = 0; v < 4; ++v) {
        double x = (double)vx[v] / (double)W, y = (double)vy[v] / (double)W;
        auto it = disp.find(vkey(vx[v], vy[v]));
        if (it != disp.end()) {
          x += it->second.first;
          y += it->second.second;
        }
        nodes[(c * 4 + v) * 2 + 0] = x;
        nodes[(c * 4 + v) * 2 + 1] = y;
      }
      // probe a quarter of the way along each side, a quarter finest unit outside
      const int64_t qx = 4 * L.x0, qy = 4 * L.y0, qs = 4 * L.sz;
      const int64_t px[4] = {qx + qs / 4, qx + qs + 1, qx + qs / 4, qx - 1};
      const int64_t py[4] = {qy - 1, qy + qs / 4, qy + qs + 1, qy + qs / 4};
      for (int s = 0; s < 4; ++s) {
        int nb = locate(px[s], py[s]);
        parities[c * 4 + s] = (nb < 0 || nb > (int)c) ? 1 : -1;
      }
      if (level) level[c] = L.lvl;
    }
    return (int)count;
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

// skeleton.hip -- assembly of the condensed (trace and flux) system of a
// conforming mesh.
//
// After static condensation every cell contributes its Schur complement S_c
// (T x T over the cell's trace and flux dofs, in its dof order) and its
// condensed load g_c. The reference maps them to global dofs with the minimum
// rule (GDAMinimumRule::interpretLocalData, GDAMinimumRule.cpp:714-780, via
// LocalDofMapper) and inserts them into the global matrix (Solution.cpp:
// 1263-1293; CondensedDofInterpreter.cpp:1084,1194-1224). On a conforming mesh
// (no hanging sides, one polynomial order) that rule reduces to:
//   - an H^1 trace's vertex, edge and face nodes are shared by every cell that
//     touches the entity (so within one cell the same global dof appears on
//     several sides, and its contributions add up);
//   - an L^2 flux's nodes belong to one side, shared by its two cells (the
//     side parity is already inside the local matrices);
//   - every cell sees a shared entity in one canonical orientation: edges run
//     from the smaller to the larger vertex id; a face starts at its smallest
//     vertex, first axis toward the smaller of that vertex's two face
//     neighbours.
//
// Host: dpg_skeleton_numbering gives the global index of every trailing dof
// of every cell; dpg_skeleton_nodes the reference coordinates of those dofs'
// nodes. Device: dpg_skeleton_create sorts the (row, col) keys of all cell
// contributions (stable radix sort) into the CSR pattern and the order in
// which each nonzero sums its contributions; dpg_skeleton_assemble sums them,
// one thread per nonzero, in increasing (cell, local row, local column) order,
// so every run gives the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <rocprim/rocprim.hpp>
#include <string>
#include <vector>

#include "../../include/dpg_capi.h"
#include "element.hpp"
#include "errors.hpp"
#include "kernels.hpp"

#define SK_CHECK(expr)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace dpg {
namespace {

// the trailing (trace / flux) blocks of a layout, in dof order
std::vector<TrialBlock> side_blocks(const ElementLayout& L) {
  std::vector<TrialBlock> out;
  for (const TrialBlock& b : L.blocks)
    if (b.side >= 0) out.push_back(b);
  return out;
}

// nodes per direction of a side block's nodal basis, minus one
int side_degree(const ElementLayout& L, const TrialBlock& b) {
  return b.kind == DPG_TRACE ? L.trace_deg : L.flux_deg;
}

struct Numbering {
  int nglobal = 0;
  std::vector<int32_t> gdof;  // (C, T)
  // mesh entity of every global dof: vertices [0, nv), edges [nv, nv+ne), faces
  // (3D) [nv+ne, nv+ne+nf); a side's flux dofs share its entity with the trace
  // dofs interior to it
  std::vector<int32_t> entity;
  int nentities = 0;
};

Numbering number_skeleton(const ElementLayout& L, int C, const int32_t* cv) {
  const int V = L.rc.nverts, S = L.rc.nsides, dim = L.dim, T = L.T, F = L.F;
  const int nps = L.rc.nodes_per_side;
  if (L.topo == DPG_HEX && nps != 4) throw std::invalid_argument("hex faces must have 4 nodes");
  // compact the vertex ids: used ids in increasing order
  std::vector<int32_t> ids(cv, cv + (size_t)C * V);
  for (int32_t v : ids)
    if (v < 0) throw std::invalid_argument("negative vertex id");
  std::vector<int32_t> uniq = ids;
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  auto vid = [&](int c, int lv) {
    return (int)(std::lower_bound(uniq.begin(), uniq.end(), cv[(size_t)c * V + lv]) - uniq.begin());
  };
  for (int c = 0; c < C; ++c)
    for (int a = 0; a < V; ++a)
      for (int b = a + 1; b < V; ++b)
        if (cv[(size_t)c * V + a] == cv[(size_t)c * V + b]) throw std::invalid_argument("a cell repeats a vertex");
  const int nv = (int)uniq.size();
  // edges (2D: the sides; 3D: the edges of the faces) and faces, by sorted vertex ids
  std::map<std::pair<int, int>, int> edges;
  std::map<std::array<int, 4>, int> faces;
  auto edge_key = [](int a, int b) { return std::make_pair(std::min(a, b), std::max(a, b)); };
  for (int c = 0; c < C; ++c)
    for (int s = 0; s < S; ++s) {
      const int* sn = &L.rc.side_nodes[(size_t)s * nps];
      if (dim == 2) {
        edges.emplace(edge_key(vid(c, sn[0]), vid(c, sn[1])), 0);
      } else {
        std::array<int, 4> g{vid(c, sn[0]), vid(c, sn[1]), vid(c, sn[2]), vid(c, sn[3])};
        for (int k = 0; k < 4; ++k) edges.emplace(edge_key(g[k], g[(k + 1) % 4]), 0);
        std::sort(g.begin(), g.end());
        faces.emplace(g, 0);
      }
    }
  int ne = 0, nf = 0;
  for (auto& e : edges) e.second = ne++;
  for (auto& f : faces) f.second = nf++;
  const int nside_global = dim == 2 ? ne : nf;
  // global blocks per side variable, in trial order
  std::vector<TrialBlock> sb = side_blocks(L);
  std::map<int, int> var_base;  // var -> first global dof
  std::map<int, int> var_kind;
  int next = 0;
  for (const TrialBlock& b : sb) {
    if (var_base.count(b.var)) continue;
    var_base[b.var] = next;
    var_kind[b.var] = b.kind;
    const int n = side_degree(L, b);
    if (b.kind == DPG_TRACE) {
      next += nv + ne * (n - 1) + (dim == 3 ? nf * (n - 1) * (n - 1) : 0);
    } else {
      next += nside_global * (dim == 2 ? n + 1 : (n + 1) * (n + 1));
    }
  }
  Numbering out;
  out.nglobal = next;
  out.gdof.assign((size_t)C * T, -1);
  out.entity.assign((size_t)next, -1);
  out.nentities = nv + ne + (dim == 3 ? nf : 0);
  for (int c = 0; c < C; ++c)
    for (const TrialBlock& b : sb) {
      const int n = side_degree(L, b), base = var_base[b.var];
      const int* sn = &L.rc.side_nodes[(size_t)b.side * nps];
      int32_t* g = &out.gdof[(size_t)c * T + (b.offset - F)];
      if (dim == 2) {
        const int A = vid(c, sn[0]), B = vid(c, sn[1]);
        const int e = edges.at(edge_key(A, B));
        for (int i = 0; i <= n; ++i) {
          const int ci = A < B ? i : n - i;  // position from the smaller vertex id
          if (b.kind == DPG_TRACE) {
            if (i == 0) g[i] = base + A, out.entity[g[i]] = A;
            else if (i == n) g[i] = base + B, out.entity[g[i]] = B;
            else g[i] = base + nv + e * (n - 1) + (ci - 1), out.entity[g[i]] = nv + e;
          } else {
            g[i] = base + e * (n + 1) + ci;
            out.entity[g[i]] = nv + e;
          }
        }
      } else {
        const int G[4] = {vid(c, sn[0]), vid(c, sn[1]), vid(c, sn[2]), vid(c, sn[3])};
        std::array<int, 4> key{G[0], G[1], G[2], G[3]};
        std::sort(key.begin(), key.end());
        const int fi = faces.at(key);
        // corner grid positions in the face parametrisation (units of n)
        const int pos[4][2] = {{0, 0}, {n, 0}, {n, n}, {0, n}};
        int m = 0;
        for (int k = 1; k < 4; ++k)
          if (G[k] < G[m]) m = k;
        int un = (m + 1) % 4, vn = (m + 3) % 4;
        if (G[vn] < G[un]) std::swap(un, vn);
        const int ux = pos[un][0] - pos[m][0], uy = pos[un][1] - pos[m][1];
        const int vx = pos[vn][0] - pos[m][0], vy = pos[vn][1] - pos[m][1];
        for (int bb = 0; bb <= n; ++bb)
          for (int aa = 0; aa <= n; ++aa) {
            const int dx = aa - pos[m][0], dy = bb - pos[m][1];
            const int u = n > 0 ? (dx * ux + dy * uy) / n : 0, v = n > 0 ? (dx * vx + dy * vy) / n : 0;
            int32_t gg, ent;
            if (b.kind == DPG_FLUX) {
              gg = base + fi * (n + 1) * (n + 1) + u + v * (n + 1);
              ent = nv + ne + fi;
            } else {
              int corner = -1;
              for (int k = 0; k < 4; ++k)
                if (pos[k][0] == aa && pos[k][1] == bb) corner = k;
              if (corner >= 0) {
                gg = base + G[corner];
                ent = G[corner];
              } else if (aa == 0 || aa == n || bb == 0 || bb == n) {
                int k0, k1, t;  // face edge (k0, k1), distance t from corner k0
                if (bb == 0) {
                  k0 = 0, k1 = 1, t = aa;
                } else if (aa == n) {
                  k0 = 1, k1 = 2, t = bb;
                } else if (bb == n) {
                  k0 = 2, k1 = 3, t = n - aa;
                } else {
                  k0 = 3, k1 = 0, t = n - bb;
                }
                const int e = edges.at(edge_key(G[k0], G[k1]));
                const int cd = G[k0] < G[k1] ? t : n - t;
                gg = base + nv + e * (n - 1) + (cd - 1);
                ent = nv + e;
              } else {
                gg = base + nv + ne * (n - 1) + fi * (n - 1) * (n - 1) + (u - 1) + (v - 1) * (n - 1);
                ent = nv + ne + fi;
              }
            }
            g[aa + bb * (n + 1)] = gg;
            out.entity[gg] = ent;
          }
      }
    }
  for (int32_t v : out.gdof)
    if (v < 0 || v >= out.nglobal) throw std::logic_error("skeleton numbering left a dof unassigned");
  for (int32_t v : out.entity)
    if (v < 0 || v >= out.nentities) throw std::logic_error("skeleton numbering left a dof without entity");
  return out;
}

// ------------------------------------------------------------------ device
__global__ void expand_keys(const int32_t* __restrict__ gdof, int T, int64_t nglobal, uint64_t E,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const uint64_t TT = (uint64_t)T * T;
  const uint64_t c = p / TT, rem = p - c * TT, i = rem / T, j = rem - i * T;
  keys[p] = (uint64_t)gdof[c * T + i] * (uint64_t)nglobal + (uint64_t)gdof[c * T + j];
  idx[p] = (uint32_t)p;
}

// the same for cells with different numbers of dofs: cell c holds dofs
// [off[c], off[c+1]) and its contributions are its n_c x n_c block, row-major,
// from position off2[c] (off2 = prefix sums of n_c^2)
__global__ void expand_keys_ragged(const int32_t* __restrict__ off, const int32_t* __restrict__ dofs,
                                   const uint64_t* __restrict__ off2, int ncells, int64_t nglobal, uint64_t E,
                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  int lo = 0, hi = ncells - 1;  // the cell with off2[c] <= p < off2[c+1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (off2[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  const int c = lo, nc = off[c + 1] - off[c];
  const uint64_t rem = p - off2[c], i = rem / nc, j = rem - i * nc;
  keys[p] = (uint64_t)dofs[off[c] + i] * (uint64_t)nglobal + (uint64_t)dofs[off[c] + j];
  idx[p] = (uint32_t)p;
}

__global__ void expand_rows(const int32_t* __restrict__ gdof, uint64_t E, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ idx) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  keys[p] = (uint64_t)gdof[p];
  idx[p] = (uint32_t)p;
}

// CSR columns and row pointers from the sorted unique keys row*nglobal + col
__global__ void csr_from_keys(const uint64_t* __restrict__ uk, uint32_t nnz, int64_t nglobal,
                              int32_t* __restrict__ col, int64_t* __restrict__ row_ptr) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e > nnz) return;
  const int64_t r = e < nnz ? (int64_t)(uk[e] / (uint64_t)nglobal) : nglobal;
  const int64_t rp = e > 0 ? (int64_t)(uk[e - 1] / (uint64_t)nglobal) : -1;
  if (e < nnz) col[e] = (int32_t)(uk[e] - (uint64_t)r * (uint64_t)nglobal);
  for (int64_t q = rp + 1; q <= r; ++q) row_ptr[q] = (int64_t)e;
}

// out[e] = sum of src[perm[q]] for q in [seg[e], seg[e+1]), in order
__global__ void ordered_segment_sum(const double* __restrict__ src, const uint32_t* __restrict__ perm,
                                    const uint32_t* __restrict__ seg, uint32_t n, double* __restrict__ out) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double acc = 0.0;
  for (uint32_t q = seg[e]; q < seg[e + 1]; ++q) acc = acc + src[perm[q]];
  out[e] = acc;
}

// ----------------------------------------------------------- global solve
// y[row] = (A x)[row] for free rows, 0 for fixed rows; one wavefront per row,
// reduced across the wavefront in a fixed order
__global__ __launch_bounds__(256) void spmv_free_rows(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                      const double* __restrict__ val, const double* __restrict__ x,
                                                      const int32_t* __restrict__ fixed, int64_t n,
                                                      double* __restrict__ y) {
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;  // whole wavefronts: n rows per 64-lane groups
  double acc = 0.0;
  if (!fixed[row])
    for (int64_t q = rp[row] + lane; q < rp[row + 1]; q += 64) acc = fma(val[q], x[col[q]], acc);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) y[row] = acc;
}

__global__ void csr_diagonal(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                             const double* __restrict__ val, int64_t n, double* __restrict__ d) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int64_t lo = rp[row], hi = rp[row + 1];
  while (lo < hi) {  // columns are sorted within a row
    const int64_t mid = (lo + hi) / 2;
    if (col[mid] < row) lo = mid + 1;
    else hi = mid;
  }
  d[row] = (lo < rp[row + 1] && col[lo] == row) ? val[lo] : 0.0;
}

constexpr int kRedBlocks = 1024;  // fixed grid: the same summation order on every call

// part[b] = sum over the grid-stride slice of block b of a[i] * b[i] (op 0),
// or of max |a[i]| (op 1); tree-reduced in shared memory
__global__ __launch_bounds__(256) void reduce_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                      int64_t n, int op, double* __restrict__ part) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < n; i += (int64_t)gridDim.x * 256)
    acc = op == 0 ? fma(a[i], b[i], acc) : fmax(acc, fabs(a[i]));
  sh[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] = op == 0 ? sh[t] + sh[t + w] : fmax(sh[t], sh[t + w]);
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void reduce_final(const double* __restrict__ part, int np, int op,
                                                    double* __restrict__ out) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int i = t; i < np; i += 256) acc = op == 0 ? acc + part[i] : fmax(acc, part[i]);
  sh[t] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] = op == 0 ? sh[t] + sh[t + w] : fmax(sh[t], sh[t + w]);
    __syncthreads();
  }
  if (t == 0) out[0] = sh[0];
}

// r = b - q on free rows (0 on fixed), z = r / d, p = z
__global__ void cg_start(const double* __restrict__ b, const double* __restrict__ q, const double* __restrict__ d,
                         const int32_t* __restrict__ fixed, int64_t n, double* __restrict__ r, double* __restrict__ z,
                         double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ri = fixed[i] ? 0.0 : b[i] - q[i];
  const double zi = fixed[i] ? 0.0 : ri / d[i];
  r[i] = ri;
  z[i] = zi;
  p[i] = zi;
}

// x += alpha p, r -= alpha q (p and q are 0 on fixed rows)
__global__ void cg_step(double alpha, const double* __restrict__ p, const double* __restrict__ q, int64_t n,
                        double* __restrict__ x, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = fma(alpha, p[i], x[i]);
  r[i] = fma(-alpha, q[i], r[i]);
}

// z = r / d on free rows, 0 on fixed ones
__global__ void cg_precondition(const double* __restrict__ r, const double* __restrict__ d,
                                const int32_t* __restrict__ fixed, int64_t n, double* __restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  z[i] = fixed[i] ? 0.0 : r[i] / d[i];
}

// p = z + beta p
__global__ void cg_direction(double beta, const double* __restrict__ z, int64_t n, double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = fma(beta, p[i], z[i]);
}

// Block-Jacobi over mesh entities. Block b holds the global dofs
// dofs[ptr[b] .. ptr[b+1]) (nb <= 64); its dense diagonal block of the matrix,
// row-major at off2[b], rows and columns of fixed dofs replaced by the identity.
// One workgroup of 64 threads per block; entries by binary search in the
// sorted columns of the row.
__global__ __launch_bounds__(64) void block_extract(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                    const double* __restrict__ val, const int32_t* __restrict__ fixed,
                                                    const int32_t* __restrict__ ptr, const int32_t* __restrict__ dofs,
                                                    const int64_t* __restrict__ off2, double* __restrict__ blk) {
  const int b = blockIdx.x, p0 = ptr[b], nb = ptr[b + 1] - p0;
  double* out = blk + off2[b];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb, di = dofs[p0 + i], dj = dofs[p0 + j];
    double v;
    if (fixed[di] || fixed[dj]) {
      v = i == j ? 1.0 : 0.0;
    } else {
      int64_t lo = rp[di], hi = rp[di + 1];
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (col[mid] < dj) lo = mid + 1;
        else hi = mid;
      }
      v = (lo < rp[di + 1] && col[lo] == dj) ? val[lo] : 0.0;
    }
    out[e] = v;
  }
}

// In place: every block <- its inverse, by Cholesky L L^T in LDS (lane = row),
// the columns of L^-1 by substitution (lane = column) and (L^-1)^T L^-1 (lane =
// row). A non-positive pivot sets *bad.
__global__ __launch_bounds__(64) void block_invert(const int32_t* __restrict__ ptr, const int64_t* __restrict__ off2,
                                                   int nbmax, double* __restrict__ blk, int32_t* __restrict__ bad) {
  extern __shared__ double sh[];
  const int b = blockIdx.x, nb = ptr[b + 1] - ptr[b], t = threadIdx.x;
  double* a = sh;                        // nb x nb, row-major
  double* li = sh + (size_t)nbmax * nbmax;  // L^-1, row-major
  double* g = blk + off2[b];
  for (int e = t; e < nb * nb; e += blockDim.x) a[e] = g[e];
  __syncthreads();
  for (int k = 0; k < nb; ++k) {
    if (t == k) {
      const double d = a[k * nb + k];
      if (!(d > 0.0)) atomicExch(bad, 1);
      a[k * nb + k] = d > 0.0 ? sqrt(d) : 1.0;
    }
    __syncthreads();
    if (t > k && t < nb) a[t * nb + k] /= a[k * nb + k];
    __syncthreads();
    if (t > k && t < nb)
      for (int j = k + 1; j <= t; ++j) a[t * nb + j] = fma(-a[t * nb + k], a[j * nb + k], a[t * nb + j]);
    __syncthreads();
  }
  if (t < nb) {  // column t of L^-1: L x = e_t
    for (int j = 0; j < t; ++j) li[j * nb + t] = 0.0;
    li[t * nb + t] = 1.0 / a[t * nb + t];
    for (int j = t + 1; j < nb; ++j) {
      double acc = 0.0;
      for (int k = t; k < j; ++k) acc = fma(a[j * nb + k], li[k * nb + t], acc);
      li[j * nb + t] = -acc / a[j * nb + j];
    }
  }
  __syncthreads();
  if (t < nb)  // row t of (L^-1)^T L^-1
    for (int j = 0; j < nb; ++j) {
      double acc = 0.0;
      for (int k = t > j ? t : j; k < nb; ++k) acc = fma(li[k * nb + t], li[k * nb + j], acc);
      g[t * nb + j] = acc;
    }
}

// z = (block inverse) r on the free dofs, 0 on the fixed ones; one thread per
// position of the block-ordered dof list
__global__ void block_apply(const int32_t* __restrict__ ptr, const int32_t* __restrict__ dofs,
                            const int32_t* __restrict__ pos_block, const int64_t* __restrict__ off2,
                            const double* __restrict__ inv, const int32_t* __restrict__ fixed,
                            const double* __restrict__ r, int64_t npos, double* __restrict__ z) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npos) return;
  const int b = pos_block[p], p0 = ptr[b], nb = ptr[b + 1] - p0, i = (int)(p - p0), di = dofs[p];
  if (fixed[di]) {
    z[di] = 0.0;
    return;
  }
  const double* row = inv + off2[b] + (int64_t)i * nb;
  double acc = 0.0;
  for (int j = 0; j < nb; ++j) acc = fma(row[j], r[dofs[p0 + j]], acc);
  z[di] = acc;
}

__global__ void copy_f64(const double* __restrict__ a, int64_t n, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void mark_above(const double* __restrict__ err, int n, const double* __restrict__ emax, double rel,
                           int32_t* __restrict__ flags, double* __restrict__ fcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = err[i] > rel * emax[0] ? 1 : 0;
  flags[i] = f;
  fcount[i] = f;
}

// --------------------------------------------- two-level (p-multigrid) solve
// Coarse blocks of every cell: Sc = P^T D S D P, with D the free mask of the
// cell's fine dofs (rows and columns of prescribed dofs dropped) and P the
// local prolongation (T x Tc), the same for every cell. One workgroup per
// cell, W = D S D P in LDS; Sc is computed for a >= b and mirrored, so it is
// symmetric bit for bit.
__global__ __launch_bounds__(256) void coarse_blocks(const double* __restrict__ S, int T, int Tc,
                                                     const double* __restrict__ P, const int32_t* __restrict__ gdof,
                                                     const int32_t* __restrict__ fixed, double* __restrict__ Sc) {
  extern __shared__ double lds[];
  double* W = lds;                   // T x Tc
  double* m = lds + (size_t)T * Tc;  // free flags (0 / 1) of the cell's fine dofs
  const int c = blockIdx.x, t = threadIdx.x;
  const double* Sin = S + (size_t)c * T * T;
  for (int i = t; i < T; i += 256) m[i] = fixed[gdof[(size_t)c * T + i]] ? 0.0 : 1.0;
  __syncthreads();
  for (int e = t; e < T * Tc; e += 256) {
    const int i = e / Tc, a = e % Tc;
    double acc = 0.0;
    if (m[i] != 0.0)
      for (int j = 0; j < T; ++j) acc = fma(Sin[(size_t)i * T + j] * m[j], P[(size_t)j * Tc + a], acc);
    W[e] = acc;
  }
  __syncthreads();
  double* out = Sc + (size_t)c * Tc * Tc;
  for (int e = t; e < Tc * Tc; e += 256) {
    const int a = e / Tc, b = e % Tc, hi = a > b ? a : b, lo = a > b ? b : a;
    double acc = 0.0;
    for (int i = 0; i < T; ++i) acc = fma(P[(size_t)i * Tc + hi], W[(size_t)i * Tc + lo], acc);
    out[e] = acc;
  }
}

// y[i] = sum over row i of val * x[col] (y[i] += ... with add), 0 where
// mask[i] is set (mask may be NULL); one thread per row
__global__ void csr_apply(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                          const double* __restrict__ val, const double* __restrict__ x,
                          const int32_t* __restrict__ mask, int64_t n, int add, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  if (!(mask && mask[i]))
    for (int k = rp[i]; k < rp[i + 1]; ++k) acc = fma(val[k], x[col[k]], acc);
  y[i] = add ? y[i] + acc : acc;
}

// coarse dofs whose row of the coarse matrix vanishes (all their fine support
// prescribed) are held at 0 like prescribed dofs
__global__ void zero_diag_mask(const double* __restrict__ d, int64_t n, int32_t* __restrict__ mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mask[i] = d[i] > 0.0 ? 0 : 1;
}

// conjugate gradients on the coarse matrix with the scalars kept on the
// device (s[0] = r.z, s[1] = p.q, s[2] = the new r.z): no host round trips
__global__ void dcg_step(const double* __restrict__ s, const double* __restrict__ p, const double* __restrict__ q,
                         int64_t n, double* __restrict__ x, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = s[1] > 0.0 ? s[0] / s[1] : 0.0;
  x[i] = fma(alpha, p[i], x[i]);
  r[i] = fma(-alpha, q[i], r[i]);
}

__global__ void dcg_direction(const double* __restrict__ s, const double* __restrict__ z, int64_t n,
                              double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = s[0] > 0.0 ? s[2] / s[0] : 0.0;
  p[i] = fma(beta, p[i], z[i]);
}

__global__ void dcg_shift(double* __restrict__ s) { s[0] = s[2]; }

// r - y on the free rows, 0 on the prescribed ones
__global__ void sub_free(const double* __restrict__ r, const double* __restrict__ y, const int32_t* __restrict__ fixed,
                         int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = fixed[i] ? 0.0 : r[i] - y[i];
}

__global__ void fill_f64(double* __restrict__ a, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// a += b
__global__ void add_f64(const double* __restrict__ b, int64_t n, double* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

// a *= s
__global__ void scale_f64(double* __restrict__ a, int64_t n, double s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] *= s;
}

// a start vector for power iteration: a fixed pseudo-random value in [-1, 1)
// on the free rows (Weyl sequence of the row index), 0 on the prescribed ones
__global__ void seed_free(const int32_t* __restrict__ fixed, int64_t n, double* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = (uint32_t)((uint64_t)(i + 1) * 2654435761ull);
  v[i] = fixed[i] ? 0.0 : (double)h * (2.0 / 4294967296.0) - 1.0;
}

template <class T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  explicit DevArray(size_t count = 0) { resize(count); }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() {
    if (p) (void)hipFree(p);
  }
  void resize(size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    if (count) SK_CHECK(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
  }
};

unsigned key_bits(uint64_t maxkey) {
  unsigned b = 1;
  while (b < 64 && (maxkey >> b) != 0) ++b;
  return b;
}

// stable sort of (key, index) pairs, then the segments of equal keys:
// perm = indices in key order, uk = unique keys, seg = segment starts (runs + 1)
void sort_and_segment(DevArray<uint64_t>& keys, DevArray<uint32_t>& idx, uint64_t E, uint64_t maxkey,
                      DevArray<uint32_t>& perm, DevArray<uint64_t>& uk, DevArray<uint32_t>& seg, uint32_t& runs,
                      hipStream_t st) {
  DevArray<uint64_t> sk(E);
  perm.resize(E);
  const unsigned bits = key_bits(maxkey);
  size_t tb = 0;
  SK_CHECK(rocprim::radix_sort_pairs(nullptr, tb, keys.p, sk.p, idx.p, perm.p, (size_t)E, 0, bits, st));
  DevArray<unsigned char> tmp(tb);
  SK_CHECK(rocprim::radix_sort_pairs(tmp.p, tb, keys.p, sk.p, idx.p, perm.p, (size_t)E, 0, bits, st));
  keys.resize(0);
  idx.resize(0);
  uk.resize(E);
  DevArray<uint32_t> counts(E + 1), nruns(1);  // the scan below reads one past the runs
  tb = 0;
  SK_CHECK(rocprim::run_length_encode(nullptr, tb, sk.p, (unsigned int)E, uk.p, counts.p, nruns.p, st));
  DevArray<unsigned char> tmp2(tb);
  SK_CHECK(rocprim::run_length_encode(tmp2.p, tb, sk.p, (unsigned int)E, uk.p, counts.p, nruns.p, st));
  SK_CHECK(hipMemcpyAsync(&runs, nruns.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SK_CHECK(hipStreamSynchronize(st));
  seg.resize((size_t)runs + 1);
  tb = 0;
  SK_CHECK(rocprim::exclusive_scan(nullptr, tb, counts.p, seg.p, (uint32_t)0, (size_t)runs + 1,
                                   rocprim::plus<uint32_t>(), st));
  DevArray<unsigned char> tmp3(tb);
  // counts[runs] is past the runs: make it 0 so that seg[runs] = E
  SK_CHECK(hipMemsetAsync(counts.p + runs, 0, sizeof(uint32_t), st));
  SK_CHECK(rocprim::exclusive_scan(tmp3.p, tb, counts.p, seg.p, (uint32_t)0, (size_t)runs + 1,
                                   rocprim::plus<uint32_t>(), st));
  SK_CHECK(hipStreamSynchronize(st));
}

// reduction of n values on `st` (op 0: dot of a and b; op 1: max |a|), read back
double reduce(const double* a, const double* b, int64_t n, int op, DevArray<double>& part, DevArray<double>& one,
              hipStream_t st) {
  reduce_partial<<<kRedBlocks, 256, 0, st>>>(a, b, n, op, part.p);
  reduce_final<<<1, 256, 0, st>>>(part.p, kRedBlocks, op, one.p);
  SK_CHECK(hipGetLastError());
  double v = 0.0;
  SK_CHECK(hipMemcpyAsync(&v, one.p, sizeof(double), hipMemcpyDeviceToHost, st));
  SK_CHECK(hipStreamSynchronize(st));
  return v;
}

unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

// reference coordinates (T, dim) of the nodes of a layout's trailing dofs
std::vector<double> side_node_coords(const ElementLayout& L) {
  const int dim = L.dim;
  std::vector<double> xi((size_t)L.T * dim);
  for (const TrialBlock& b : side_blocks(L)) {
    const int n = side_degree(L, b);
    const std::vector<double> t = n > 0 ? gll_points(n) : std::vector<double>{0.0};  // degree 0: the side centre
    const double* p = &L.rc.side_param[(size_t)b.side * 3 * dim];
    for (int k = 0; k < b.card; ++k) {
      const int a = k % (n + 1), bb = k / (n + 1);
      double* x = &xi[(size_t)(b.offset - L.F + k) * dim];
      for (int d = 0; d < dim; ++d) x[d] = p[d] + p[dim + d] * t[a] + (dim == 3 ? p[2 * dim + d] * t[bb] : 0.0);
    }
  }
  return xi;
}

// Local prolongation (T x Tc, row-major) from the lowest-order layout Lc to L:
// a coarse trace dof is a vertex of its side and its function on the side is
// the cell's vertex function prod_d (1 + v_d xi_d) / 2 (reference cell
// [-1,1]^d), evaluated at the fine nodes of the same variable and side; a
// coarse flux dof is the constant 1 on its side.
std::vector<double> local_prolongation(const ElementLayout& L, const ElementLayout& Lc) {
  const int dim = L.dim, T = L.T, Tc = Lc.T;
  const std::vector<double> xf = side_node_coords(L), xc = side_node_coords(Lc);
  std::vector<double> P((size_t)T * Tc, 0.0);
  for (const TrialBlock& b : side_blocks(L))
    for (const TrialBlock& bc : side_blocks(Lc)) {
      if (b.var != bc.var || b.side != bc.side) continue;
      for (int i = 0; i < b.card; ++i)
        for (int j = 0; j < bc.card; ++j) {
          const int fi = b.offset - L.F + i, cj = bc.offset - Lc.F + j;
          double w = 1.0;
          if (b.kind == DPG_TRACE)
            for (int d = 0; d < dim; ++d) w *= 0.5 * (1.0 + xc[(size_t)cj * dim + d] * xf[(size_t)fi * dim + d]);
          P[(size_t)fi * Tc + cj] = w;
        }
    }
  return P;
}

}  // namespace
}  // namespace dpg

struct dpg_skeleton {
  int ncells = 0, T = 0;
  int64_t nglobal = 0;
  uint32_t nnz = 0;
  dpg::DevArray<uint32_t> perm, seg;  // matrix: contributions in summation order, segment starts
  dpg::DevArray<uint32_t> rperm, rseg;  // load vector
  dpg::DevArray<int64_t> row_ptr;
  dpg::DevArray<int32_t> col;
  // block-Jacobi blocks (dpg_skeleton_set_blocks); none: point Jacobi
  int nblocks = 0, nbmax = 0;
  int64_t blk_total = 0;  // sum of the squared block sizes
  dpg::DevArray<int32_t> blk_ptr, blk_dofs, pos_block;
  dpg::DevArray<int64_t> blk_off2;
};

// the lowest-order level of the condensed system (dpg_coarse_create)
struct dpg_coarse {
  int ncells = 0, T = 0, Tc = 0;
  int64_t nfine = 0, ncoarse = 0;
  dpg::DevArray<int32_t> gdof;  // (C, T) fine global dofs of every cell
  dpg::DevArray<double> Pl;     // (T, Tc) local prolongation
  dpg::DevArray<int32_t> p_rp, p_col, pt_rp, pt_col;  // global P (fine rows) and P^T (coarse rows), CSR
  dpg::DevArray<double> p_val, pt_val;
  std::unique_ptr<dpg_skeleton> csk;  // coarse pattern and summation order
};

namespace dpg {
namespace {

// CSR pattern and summation orders of sum_c P_c^T S_c P_c for cells with the
// global dofs dofs[off[c] .. off[c+1]) (host arrays); T > 0 when every cell has
// T dofs (the contributions are then (c, i, j) in order), 0 for the ragged form.
std::unique_ptr<dpg_skeleton> build_skeleton(int ncells, const int32_t* off, const int32_t* dofs, int nglobal, int T,
                                             hipStream_t st) {
  const int64_t ndof = off[ncells];
  uint64_t E = 0;
  std::vector<uint64_t> off2((size_t)ncells + 1, 0);
  for (int c = 0; c < ncells; ++c) {
    const uint64_t nc = (uint64_t)(off[c + 1] - off[c]);
    off2[(size_t)c + 1] = off2[c] + nc * nc;
  }
  E = off2[ncells];
  if (E > 0xFFFFFFFFull) throw std::invalid_argument("the number of contributions must be below 2^32");
  for (int64_t q = 0; q < ndof; ++q)
    if (dofs[q] < 0 || dofs[q] >= nglobal) throw std::invalid_argument("global dof index out of range");
  std::unique_ptr<dpg_skeleton> sk(new dpg_skeleton());
  sk->ncells = ncells;
  sk->T = T;
  sk->nglobal = nglobal;
  DevArray<int32_t> dg((size_t)ndof), doff((size_t)ncells + 1);
  DevArray<uint64_t> doff2((size_t)ncells + 1);
  SK_CHECK(hipMemcpyAsync(dg.p, dofs, sizeof(int32_t) * dg.n, hipMemcpyHostToDevice, st));
  SK_CHECK(hipMemcpyAsync(doff.p, off, sizeof(int32_t) * doff.n, hipMemcpyHostToDevice, st));
  SK_CHECK(hipMemcpyAsync(doff2.p, off2.data(), sizeof(uint64_t) * doff2.n, hipMemcpyHostToDevice, st));
  {  // matrix pattern and summation order
    DevArray<uint64_t> keys(E), uk;
    DevArray<uint32_t> idx(E);
    if (T > 0)
      expand_keys<<<(unsigned)((E + 255) / 256), 256, 0, st>>>(dg.p, T, nglobal, E, keys.p, idx.p);
    else
      expand_keys_ragged<<<(unsigned)((E + 255) / 256), 256, 0, st>>>(doff.p, dg.p, doff2.p, ncells, nglobal, E,
                                                                       keys.p, idx.p);
    SK_CHECK(hipGetLastError());
    const uint64_t maxkey = (uint64_t)nglobal * (uint64_t)nglobal - 1;
    sort_and_segment(keys, idx, E, maxkey, sk->perm, uk, sk->seg, sk->nnz, st);
    sk->col.resize(sk->nnz);
    sk->row_ptr.resize((size_t)nglobal + 1);
    csr_from_keys<<<(unsigned)(((uint64_t)sk->nnz + 1 + 255) / 256), 256, 0, st>>>(uk.p, sk->nnz, nglobal, sk->col.p,
                                                                                  sk->row_ptr.p);
    SK_CHECK(hipGetLastError());
  }
  {  // load vector: every global dof has at least one contribution
    const uint64_t E2 = (uint64_t)ndof;
    DevArray<uint64_t> keys(E2), uk;
    DevArray<uint32_t> idx(E2);
    expand_rows<<<(unsigned)((E2 + 255) / 256), 256, 0, st>>>(dg.p, E2, keys.p, idx.p);
    SK_CHECK(hipGetLastError());
    uint32_t runs = 0;
    sort_and_segment(keys, idx, E2, (uint64_t)nglobal - 1, sk->rperm, uk, sk->rseg, runs, st);
    if ((int64_t)runs != nglobal) throw std::invalid_argument("some global dofs belong to no cell");
  }
  SK_CHECK(hipStreamSynchronize(st));
  return sk;
}

}  // namespace
}  // namespace dpg

extern "C" {

int dpg_skeleton_numbering(const dpg_form_desc* form, const dpg_bucket_desc* bucket, int ncells,
                           const int32_t* cell_vertices, int32_t* gdof) {
  return dpg::guarded_status([&] {
    if (!form || !bucket || ncells < 0 || (ncells > 0 && (!cell_vertices || !gdof)))
      throw std::invalid_argument("NULL argument or negative cell count");
    dpg::ElementLayout L = dpg::make_layout(*form, *bucket);
    if (L.T <= 0) throw std::invalid_argument("the form has no trace or flux variables");
    if (ncells == 0) return 0;
    dpg::Numbering nb = dpg::number_skeleton(L, ncells, cell_vertices);
    std::memcpy(gdof, nb.gdof.data(), nb.gdof.size() * sizeof(int32_t));
    return nb.nglobal;
  });
}

int dpg_skeleton_entities(const dpg_form_desc* form, const dpg_bucket_desc* bucket, int ncells,
                          const int32_t* cell_vertices, int32_t* entity) {
  return dpg::guarded_status([&] {
    if (!form || !bucket || ncells < 0 || (ncells > 0 && (!cell_vertices || !entity)))
      throw std::invalid_argument("NULL argument or negative cell count");
    dpg::ElementLayout L = dpg::make_layout(*form, *bucket);
    if (L.T <= 0) throw std::invalid_argument("the form has no trace or flux variables");
    if (ncells == 0) return 0;
    dpg::Numbering nb = dpg::number_skeleton(L, ncells, cell_vertices);
    std::memcpy(entity, nb.entity.data(), nb.entity.size() * sizeof(int32_t));
    return nb.nentities;
  });
}

int dpg_skeleton_nodes(const dpg_form_desc* form, const dpg_bucket_desc* bucket, double* xi) {
  return dpg::guarded([&] {
    if (!form || !bucket || !xi) throw std::invalid_argument("NULL argument");
    dpg::ElementLayout L = dpg::make_layout(*form, *bucket);
    const std::vector<double> x = dpg::side_node_coords(L);
    std::memcpy(xi, x.data(), x.size() * sizeof(double));
  });
}

int dpg_skeleton_create(int ncells, int T, const int32_t* gdof, int nglobal, void* stream, dpg_skeleton** out) {
  return dpg::guarded([&] {
    if (!out || ncells <= 0 || T <= 0 || nglobal <= 0 || !gdof) throw std::invalid_argument("bad arguments");
    *out = nullptr;
    std::vector<int32_t> off((size_t)ncells + 1);
    for (int c = 0; c <= ncells; ++c) off[c] = c * T;
    if ((int64_t)ncells * T > 0x7fffffffLL) throw std::invalid_argument("ncells * T must be below 2^31");
    *out = dpg::build_skeleton(ncells, off.data(), gdof, nglobal, T, (hipStream_t)stream).release();
  });
}

int dpg_skeleton_create_ragged(int ncells, const int32_t* offsets, const int32_t* dofs, int nglobal, void* stream,
                               dpg_skeleton** out) {
  return dpg::guarded([&] {
    if (!out || ncells <= 0 || nglobal <= 0 || !offsets || !dofs) throw std::invalid_argument("bad arguments");
    *out = nullptr;
    if (offsets[0] != 0) throw std::invalid_argument("offsets[0] must be 0");
    for (int c = 0; c < ncells; ++c)
      if (offsets[c + 1] <= offsets[c]) throw std::invalid_argument("every cell needs at least one dof");
    *out = dpg::build_skeleton(ncells, offsets, dofs, nglobal, 0, (hipStream_t)stream).release();
  });
}

int dpg_skeleton_info(const dpg_skeleton* sk, int64_t* nnz, int64_t* nglobal, const int64_t** row_ptr,
                      const int32_t** col) {
  return dpg::guarded([&] {
    if (!sk) throw std::invalid_argument("NULL skeleton");
    if (nnz) *nnz = sk->nnz;
    if (nglobal) *nglobal = sk->nglobal;
    if (row_ptr) *row_ptr = sk->row_ptr.p;
    if (col) *col = sk->col.p;
  });
}

int dpg_skeleton_pattern(const dpg_skeleton* sk, int64_t* row_ptr, int32_t* col, void* stream) {
  return dpg::guarded([&] {
    if (!sk || !row_ptr || !col) throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    SK_CHECK(hipMemcpyAsync(row_ptr, sk->row_ptr.p, sizeof(int64_t) * ((size_t)sk->nglobal + 1),
                            hipMemcpyDeviceToDevice, st));
    SK_CHECK(hipMemcpyAsync(col, sk->col.p, sizeof(int32_t) * (size_t)sk->nnz, hipMemcpyDeviceToDevice, st));
  });
}

int dpg_skeleton_assemble(const dpg_skeleton* sk, const double* schur, const double* schur_rhs, double* values,
                          double* rhs, void* stream) {
  return dpg::guarded([&] {
    if (!sk) throw std::invalid_argument("NULL skeleton");
    hipStream_t st = (hipStream_t)stream;
    if (values) {
      if (!schur) throw std::invalid_argument("values needs schur");
      dpg::ordered_segment_sum<<<(unsigned)(((uint64_t)sk->nnz + 255) / 256), 256, 0, st>>>(
          schur, sk->perm.p, sk->seg.p, sk->nnz, values);
      SK_CHECK(hipGetLastError());
    }
    if (rhs) {
      if (!schur_rhs) throw std::invalid_argument("rhs needs schur_rhs");
      dpg::ordered_segment_sum<<<(unsigned)((sk->nglobal + 255) / 256), 256, 0, st>>>(
          schur_rhs, sk->rperm.p, sk->rseg.p, (uint32_t)sk->nglobal, rhs);
      SK_CHECK(hipGetLastError());
    }
  });
}

int dpg_skeleton_set_blocks(dpg_skeleton* sk, const int32_t* entity, int nentities, void* stream) {
  return dpg::guarded([&] {
    if (!sk || nentities < 0 || (nentities > 0 && !entity)) throw std::invalid_argument("bad arguments");
    hipStream_t st = (hipStream_t)stream;
    sk->nblocks = sk->nbmax = 0;
    sk->blk_total = 0;
    if (nentities == 0) return;  // back to point Jacobi
    const int64_t n = sk->nglobal;
    // counting sort of the dofs by entity (ascending dof within an entity); empty entities dropped
    std::vector<int32_t> count((size_t)nentities + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
      if (entity[i] < 0 || entity[i] >= nentities) throw std::invalid_argument("entity id out of range");
      ++count[(size_t)entity[i] + 1];
    }
    std::vector<int32_t> start((size_t)nentities + 1, 0);
    for (int e = 0; e < nentities; ++e) start[(size_t)e + 1] = start[e] + count[(size_t)e + 1];
    std::vector<int32_t> dofs((size_t)n), fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < n; ++i) dofs[(size_t)fill[(size_t)entity[i]]++] = (int32_t)i;
    std::vector<int32_t> ptr(1, 0), pos_block((size_t)n);
    std::vector<int64_t> off2(1, 0);
    int nbmax = 0;
    for (int e = 0; e < nentities; ++e) {
      const int nb = start[(size_t)e + 1] - start[e];
      if (nb == 0) continue;
      if (nb > 64) throw std::invalid_argument("an entity holds more than 64 dofs (block-Jacobi limit)");
      const int b = (int)ptr.size() - 1;
      for (int k = start[e]; k < start[(size_t)e + 1]; ++k) pos_block[(size_t)k] = b;
      ptr.push_back(start[(size_t)e + 1]);
      off2.push_back(off2.back() + (int64_t)nb * nb);
      nbmax = std::max(nbmax, nb);
    }
    sk->nblocks = (int)ptr.size() - 1;
    sk->nbmax = nbmax;
    sk->blk_total = off2.back();
    sk->blk_ptr.resize(ptr.size());
    sk->blk_dofs.resize(dofs.size());
    sk->pos_block.resize(pos_block.size());
    sk->blk_off2.resize(off2.size());
    SK_CHECK(hipMemcpyAsync(sk->blk_ptr.p, ptr.data(), sizeof(int32_t) * ptr.size(), hipMemcpyHostToDevice, st));
    SK_CHECK(hipMemcpyAsync(sk->blk_dofs.p, dofs.data(), sizeof(int32_t) * dofs.size(), hipMemcpyHostToDevice, st));
    SK_CHECK(hipMemcpyAsync(sk->pos_block.p, pos_block.data(), sizeof(int32_t) * pos_block.size(),
                            hipMemcpyHostToDevice, st));
    SK_CHECK(hipMemcpyAsync(sk->blk_off2.p, off2.data(), sizeof(int64_t) * off2.size(), hipMemcpyHostToDevice, st));
    SK_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
  });
}

int dpg_skeleton_solve(const dpg_skeleton* sk, const double* values, const double* rhs, const int32_t* fixed,
                       double* x, double tol, int maxit, int* iters, double* relres, void* stream) {
  return dpg::guarded([&] {
    if (!sk || !values || !rhs || !fixed || !x) throw std::invalid_argument("NULL argument");
    if (!(tol >= 0.0) || maxit < 0) throw std::invalid_argument("tol must be >= 0 and maxit >= 0");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = sk->nglobal;
    dpg::DevArray<double> r(n), z(n), p(n), q(n), d(n), part(dpg::kRedBlocks), one(1);
    const unsigned bn = dpg::blocks_for(n), bw = dpg::blocks_for(n * 64);
    dpg::csr_diagonal<<<bn, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, n, d.p);
    dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, x, fixed, n, q.p);
    dpg::cg_start<<<bn, 256, 0, st>>>(rhs, q.p, d.p, fixed, n, r.p, z.p, p.p);
    SK_CHECK(hipGetLastError());
    // block-Jacobi: the inverses of the entity blocks, z = B^-1 r instead of r / diag
    const bool blocks = sk->nblocks > 0;
    dpg::DevArray<double> binv(blocks ? (size_t)sk->blk_total : 0);
    auto precondition = [&]() {
      if (blocks)
        dpg::block_apply<<<bn, 256, 0, st>>>(sk->blk_ptr.p, sk->blk_dofs.p, sk->pos_block.p, sk->blk_off2.p, binv.p,
                                             fixed, r.p, n, z.p);
      else
        dpg::cg_precondition<<<bn, 256, 0, st>>>(r.p, d.p, fixed, n, z.p);
    };
    if (blocks) {
      dpg::DevArray<int32_t> bad(1);
      dpg::launch_fill_i32(bad.p, 1, 0, st);
      dpg::block_extract<<<sk->nblocks, 64, 0, st>>>(sk->row_ptr.p, sk->col.p, values, fixed, sk->blk_ptr.p,
                                                     sk->blk_dofs.p, sk->blk_off2.p, binv.p);
      dpg::block_invert<<<sk->nblocks, 64, 2 * sizeof(double) * sk->nbmax * sk->nbmax, st>>>(
          sk->blk_ptr.p, sk->blk_off2.p, sk->nbmax, binv.p, bad.p);
      SK_CHECK(hipGetLastError());
      int32_t hb = 0;
      SK_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      SK_CHECK(hipStreamSynchronize(st));
      if (hb) throw std::runtime_error("a diagonal block of the free system is not positive definite");
      precondition();
      dpg::copy_f64<<<bn, 256, 0, st>>>(z.p, n, p.p);
      SK_CHECK(hipGetLastError());
    }
    const double rr0 = dpg::reduce(r.p, r.p, n, 0, part, one, st);
    double rz = dpg::reduce(r.p, z.p, n, 0, part, one, st);
    double rr = rr0;
    int it = 0;
    while (it < maxit && rr > tol * tol * rr0 && rr0 > 0.0) {
      dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, p.p, fixed, n, q.p);
      const double pq = dpg::reduce(p.p, q.p, n, 0, part, one, st);
      if (!(pq > 0.0)) throw std::runtime_error("the free block of the system is not positive definite");
      const double alpha = rz / pq;
      dpg::cg_step<<<bn, 256, 0, st>>>(alpha, p.p, q.p, n, x, r.p);
      rr = dpg::reduce(r.p, r.p, n, 0, part, one, st);
      ++it;
      if (rr <= tol * tol * rr0) break;
      precondition();
      const double rz_new = dpg::reduce(r.p, z.p, n, 0, part, one, st);
      dpg::cg_direction<<<bn, 256, 0, st>>>(rz_new / rz, z.p, n, p.p);
      SK_CHECK(hipGetLastError());
      rz = rz_new;
    }
    if (iters) *iters = it;
    if (relres) *relres = rr0 > 0.0 ? std::sqrt(rr / rr0) : 0.0;
  });
}

int dpg_mark_cells(const double* err, int ncells, double rel_threshold, int32_t* flags, void* stream) {
  return dpg::guarded_status([&] {
    if (ncells < 0) throw std::invalid_argument("ncells < 0");
    if (ncells == 0) return 0;
    if (!err || !flags) throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    dpg::DevArray<double> part(dpg::kRedBlocks), emax(1), fc(ncells), one(1);
    dpg::reduce_partial<<<dpg::kRedBlocks, 256, 0, st>>>(err, err, ncells, 1, part.p);
    dpg::reduce_final<<<1, 256, 0, st>>>(part.p, dpg::kRedBlocks, 1, emax.p);
    dpg::mark_above<<<dpg::blocks_for(ncells), 256, 0, st>>>(err, ncells, emax.p, rel_threshold, flags, fc.p);
    SK_CHECK(hipGetLastError());
    return (int)dpg::reduce(fc.p, fc.p, ncells, 0, part, one, st);  // flags are 0 or 1: the count
  });
}

int dpg_skeleton_destroy(dpg_skeleton* sk) {
  return dpg::guarded([&] { delete sk; });
}

int dpg_coarse_create(const dpg_form_desc* form, const dpg_bucket_desc* bucket, int ncells,
                      const int32_t* cell_vertices, void* stream, dpg_coarse** out) {
  return dpg::guarded([&] {
    if (!form || !bucket || !out || ncells <= 0 || !cell_vertices) throw std::invalid_argument("bad arguments");
    *out = nullptr;
    if (bucket->topology != DPG_QUAD && bucket->topology != DPG_HEX)
      throw std::invalid_argument("two-level preconditioning is built for quads and hexes");
    hipStream_t st = (hipStream_t)stream;
    const dpg::ElementLayout L = dpg::make_layout(*form, *bucket);
    dpg_bucket_desc cb = *bucket;
    cb.h1_order = 1;  // traces linear on the sides, fluxes constant
    const dpg::ElementLayout Lc = dpg::make_layout(*form, cb, true);
    if (L.T <= 0) throw std::invalid_argument("the form has no trace or flux variables");
    const int C = ncells, T = L.T, Tc = Lc.T;
    const dpg::Numbering nf = dpg::number_skeleton(L, C, cell_vertices);
    const dpg::Numbering nc = dpg::number_skeleton(Lc, C, cell_vertices);
    const std::vector<double> Pl = dpg::local_prolongation(L, Lc);
    // global P: every fine dof's row from the first cell that holds it (the
    // coarse functions are conforming, so every cell gives the same row: checked)
    std::vector<int32_t> first((size_t)nf.nglobal, -1);  // (cell * T + local) of the first holder
    for (int c = 0; c < C; ++c)
      for (int i = 0; i < T; ++i) {
        const int32_t g = nf.gdof[(size_t)c * T + i];
        if (first[g] < 0) {
          first[g] = c * T + i;
          continue;
        }
        const int c0 = first[g] / T, i0 = first[g] % T;
        for (int j = 0; j < Tc; ++j) {
          const double v = Pl[(size_t)i * Tc + j];
          if (v == 0.0) continue;
          const int32_t gc = nc.gdof[(size_t)c * Tc + j];
          bool found = false;
          for (int j0 = 0; j0 < Tc; ++j0)
            if (nc.gdof[(size_t)c0 * Tc + j0] == gc && std::fabs(Pl[(size_t)i0 * Tc + j0] - v) <= 1e-13) found = true;
          if (!found) throw std::logic_error("coarse functions disagree between the cells of a shared dof");
        }
      }
    std::vector<int32_t> rp((size_t)nf.nglobal + 1, 0), col;
    std::vector<double> val;
    for (int64_t g = 0; g < nf.nglobal; ++g) {
      const int c = first[g] / T, i = first[g] % T;
      for (int j = 0; j < Tc; ++j) {
        const double v = Pl[(size_t)i * Tc + j];
        if (v == 0.0) continue;
        col.push_back(nc.gdof[(size_t)c * Tc + j]);
        val.push_back(v);
      }
      rp[(size_t)g + 1] = (int32_t)col.size();
    }
    // P^T by counting
    std::vector<int32_t> trp((size_t)nc.nglobal + 1, 0), tcol(col.size());
    std::vector<double> tval(col.size());
    for (int32_t j : col) ++trp[(size_t)j + 1];
    for (int64_t j = 0; j < nc.nglobal; ++j) trp[(size_t)j + 1] += trp[(size_t)j];
    std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
    for (int64_t g = 0; g < nf.nglobal; ++g)
      for (int k = rp[g]; k < rp[(size_t)g + 1]; ++k) {
        const int32_t q = fill[(size_t)col[k]]++;
        tcol[q] = (int32_t)g;
        tval[q] = val[k];
      }
    std::unique_ptr<dpg_coarse> co(new dpg_coarse());
    co->ncells = C;
    co->T = T;
    co->Tc = Tc;
    co->nfine = nf.nglobal;
    co->ncoarse = nc.nglobal;
    auto up = [&](auto& dst, const auto& src) {
      dst.resize(src.size());
      SK_CHECK(hipMemcpyAsync(dst.p, src.data(), sizeof(src[0]) * src.size(), hipMemcpyHostToDevice, st));
    };
    up(co->gdof, nf.gdof);
    up(co->Pl, Pl);
    up(co->p_rp, rp);
    up(co->p_col, col);
    up(co->p_val, val);
    up(co->pt_rp, trp);
    up(co->pt_col, tcol);
    up(co->pt_val, tval);
    dpg_skeleton* csk = nullptr;
    if (dpg_skeleton_create(C, Tc, nc.gdof.data(), nc.nglobal, stream, &csk) != 0)
      throw std::runtime_error(std::string("coarse skeleton: ") + dpg_last_error());
    co->csk.reset(csk);
    SK_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
    *out = co.release();
  });
}

int dpg_coarse_info(const dpg_coarse* co, int64_t* nglobal_coarse, int64_t* nnz_coarse, int* T_coarse) {
  return dpg::guarded([&] {
    if (!co) throw std::invalid_argument("NULL coarse level");
    if (nglobal_coarse) *nglobal_coarse = co->ncoarse;
    if (nnz_coarse) *nnz_coarse = co->csk->nnz;
    if (T_coarse) *T_coarse = co->Tc;
  });
}

int dpg_coarse_destroy(dpg_coarse* co) {
  return dpg::guarded([&] { delete co; });
}

int dpg_skeleton_solve_twolevel(const dpg_skeleton* sk, const dpg_coarse* co, const double* schur,
                                const double* values, const double* rhs, const int32_t* fixed, double* x, double tol,
                                int maxit, int coarse_iters, int* iters, double* relres, void* stream) {
  return dpg::guarded([&] {
    if (!sk || !co || !schur || !values || !rhs || !fixed || !x) throw std::invalid_argument("NULL argument");
    if (!(tol >= 0.0) || maxit < 0 || coarse_iters < 0) throw std::invalid_argument("bad tolerance or counts");
    if (sk->ncells != co->ncells || sk->T != co->T || sk->nglobal != co->nfine)
      throw std::invalid_argument("the coarse level belongs to another mesh or element type");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = sk->nglobal, nc = co->ncoarse;
    const int C = co->ncells, T = co->T, Tc = co->Tc;
    const size_t lds = sizeof(double) * ((size_t)T * Tc + T);
    if (lds > 64 * 1024) throw std::invalid_argument("coarse blocks: the element type is too large");
    const dpg_skeleton& cs = *co->csk;
    // coarse matrix A_c = P^T A_ff P from the cells' Schur blocks
    dpg::DevArray<double> Sc((size_t)C * Tc * Tc), vc(cs.nnz), dc(nc);
    dpg::DevArray<int32_t> cmask(nc);
    dpg::coarse_blocks<<<C, 256, lds, st>>>(schur, T, Tc, co->Pl.p, co->gdof.p, fixed, Sc.p);
    dpg::ordered_segment_sum<<<(unsigned)(((uint64_t)cs.nnz + 255) / 256), 256, 0, st>>>(Sc.p, cs.perm.p, cs.seg.p,
                                                                                          cs.nnz, vc.p);
    const unsigned bc = dpg::blocks_for(nc), bcw = dpg::blocks_for(nc * 64);
    dpg::csr_diagonal<<<bc, 256, 0, st>>>(cs.row_ptr.p, cs.col.p, vc.p, nc, dc.p);
    dpg::zero_diag_mask<<<bc, 256, 0, st>>>(dc.p, nc, cmask.p);
    SK_CHECK(hipGetLastError());
    // fine smoother: block Jacobi over the entities (or point Jacobi)
    dpg::DevArray<double> r(n), z(n), p(n), q(n), d(n), t1(n), t2(n), part(dpg::kRedBlocks), one(1);
    const unsigned bn = dpg::blocks_for(n), bw = dpg::blocks_for(n * 64);
    dpg::csr_diagonal<<<bn, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, n, d.p);
    const bool blocks = sk->nblocks > 0;
    dpg::DevArray<double> binv(blocks ? (size_t)sk->blk_total : 0);
    if (blocks) {
      dpg::DevArray<int32_t> bad(1);
      dpg::launch_fill_i32(bad.p, 1, 0, st);
      dpg::block_extract<<<sk->nblocks, 64, 0, st>>>(sk->row_ptr.p, sk->col.p, values, fixed, sk->blk_ptr.p,
                                                     sk->blk_dofs.p, sk->blk_off2.p, binv.p);
      dpg::block_invert<<<sk->nblocks, 64, 2 * sizeof(double) * sk->nbmax * sk->nbmax, st>>>(
          sk->blk_ptr.p, sk->blk_off2.p, sk->nbmax, binv.p, bad.p);
      SK_CHECK(hipGetLastError());
      int32_t hb = 0;
      SK_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      SK_CHECK(hipStreamSynchronize(st));
      if (hb) throw std::runtime_error("a diagonal block of the free system is not positive definite");
    }
    auto smooth = [&](const double* rin, double* zout) {
      if (blocks)
        dpg::block_apply<<<bn, 256, 0, st>>>(sk->blk_ptr.p, sk->blk_dofs.p, sk->pos_block.p, sk->blk_off2.p, binv.p,
                                             fixed, rin, n, zout);
      else
        dpg::cg_precondition<<<bn, 256, 0, st>>>(rin, d.p, fixed, n, zout);
    };
    // Damping of the smoother. The symmetric two-level preconditioner below is
    // positive definite only where the smoother S converges on its own,
    // omega * lambda_max(S A) < 2. Undamped, the entity blocks reach
    // lambda_max(S A) ~ 4 on quads and ~ 6 on hexes (a block couples to its
    // neighbours through every cell it touches), point Jacobi 10-50. The
    // largest eigenvalue is estimated by power iteration on S A (norm ratios,
    // 30 steps) and omega, its reciprocal, is folded into the block inverses or
    // the diagonal.
    {
      dpg::DevArray<double> v(n), w(n), y(n);
      dpg::seed_free<<<bn, 256, 0, st>>>(fixed, n, v.p);
      double lam = 0.0;
      for (int k = 0; k < 30; ++k) {
        dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, v.p, fixed, n, w.p);
        smooth(w.p, y.p);
        const double yy = dpg::reduce(y.p, y.p, n, 0, part, one, st), vv = dpg::reduce(v.p, v.p, n, 0, part, one, st);
        if (!(vv > 0.0) || !(yy > 0.0) || !std::isfinite(yy)) break;
        lam = std::sqrt(yy / vv);
        dpg::copy_f64<<<bn, 256, 0, st>>>(y.p, n, v.p);
        dpg::scale_f64<<<bn, 256, 0, st>>>(v.p, n, 1.0 / std::sqrt(yy));
      }
      if (lam > 0.0 && std::isfinite(lam)) {
        if (blocks)
          dpg::scale_f64<<<dpg::blocks_for(sk->blk_total), 256, 0, st>>>(binv.p, sk->blk_total, 1.0 / lam);
        else
          dpg::scale_f64<<<bn, 256, 0, st>>>(d.p, n, lam);
      }
      SK_CHECK(hipGetLastError());
      SK_CHECK(hipStreamSynchronize(st));  // v, w, y go out of scope
    }
    // coarse solve: coarse_iters steps of Jacobi-preconditioned CG from 0, scalars on the device
    dpg::DevArray<double> rc(nc), ec(nc), zc(nc), pc(nc), qc(nc), sc(4);
    auto dot_dev = [&](const double* a, const double* b2, int64_t len, double* slot) {
      dpg::reduce_partial<<<dpg::kRedBlocks, 256, 0, st>>>(a, b2, len, 0, part.p);
      dpg::reduce_final<<<1, 256, 0, st>>>(part.p, dpg::kRedBlocks, 0, slot);
    };
    auto coarse_solve = [&]() {  // ec ~ A_c^-1 rc (rc is overwritten)
      dpg::cg_precondition<<<bc, 256, 0, st>>>(rc.p, dc.p, cmask.p, nc, zc.p);
      dpg::copy_f64<<<bc, 256, 0, st>>>(zc.p, nc, pc.p);
      dot_dev(rc.p, zc.p, nc, sc.p + 0);
      for (int k = 0; k < coarse_iters; ++k) {
        dpg::spmv_free_rows<<<bcw, 256, 0, st>>>(cs.row_ptr.p, cs.col.p, vc.p, pc.p, cmask.p, nc, qc.p);
        dot_dev(pc.p, qc.p, nc, sc.p + 1);
        dpg::dcg_step<<<bc, 256, 0, st>>>(sc.p, pc.p, qc.p, nc, ec.p, rc.p);
        dpg::cg_precondition<<<bc, 256, 0, st>>>(rc.p, dc.p, cmask.p, nc, zc.p);
        dot_dev(rc.p, zc.p, nc, sc.p + 2);
        dpg::dcg_direction<<<bc, 256, 0, st>>>(sc.p, zc.p, nc, pc.p);
        dpg::dcg_shift<<<1, 1, 0, st>>>(sc.p);
      }
    };
    // two-level preconditioner, symmetric: smooth, coarse correction, smooth
    auto precondition = [&](const double* rin, double* zout) {
      smooth(rin, zout);
      dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, zout, fixed, n, t1.p);
      dpg::sub_free<<<bn, 256, 0, st>>>(rin, t1.p, fixed, n, t2.p);
      dpg::csr_apply<<<bc, 256, 0, st>>>(co->pt_rp.p, co->pt_col.p, co->pt_val.p, t2.p, cmask.p, nc, 0, rc.p);
      dpg::fill_f64<<<bc, 256, 0, st>>>(ec.p, nc, 0.0);
      coarse_solve();
      dpg::csr_apply<<<bn, 256, 0, st>>>(co->p_rp.p, co->p_col.p, co->p_val.p, ec.p, fixed, n, 1, zout);
      dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, zout, fixed, n, t1.p);
      dpg::sub_free<<<bn, 256, 0, st>>>(rin, t1.p, fixed, n, t2.p);
      smooth(t2.p, t1.p);
      dpg::add_f64<<<bn, 256, 0, st>>>(t1.p, n, zout);
      SK_CHECK(hipGetLastError());
    };
    // flexible conjugate gradients (the coarse solve is inexact)
    dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, x, fixed, n, q.p);
    dpg::sub_free<<<bn, 256, 0, st>>>(rhs, q.p, fixed, n, r.p);
    const double rr0 = dpg::reduce(r.p, r.p, n, 0, part, one, st);
    double rr = rr0;
    int it = 0;
    if (rr0 > 0.0) {
      precondition(r.p, z.p);
      dpg::copy_f64<<<bn, 256, 0, st>>>(z.p, n, p.p);
      while (it < maxit && rr > tol * tol * rr0) {
        dpg::spmv_free_rows<<<bw, 256, 0, st>>>(sk->row_ptr.p, sk->col.p, values, p.p, fixed, n, q.p);
        const double pq = dpg::reduce(p.p, q.p, n, 0, part, one, st);
        if (!(pq > 0.0)) throw std::runtime_error("the free block of the system is not positive definite");
        const double pr = dpg::reduce(p.p, r.p, n, 0, part, one, st);
        dpg::cg_step<<<bn, 256, 0, st>>>(pr / pq, p.p, q.p, n, x, r.p);
        rr = dpg::reduce(r.p, r.p, n, 0, part, one, st);
        ++it;
        if (rr <= tol * tol * rr0) break;
        precondition(r.p, z.p);
        const double zq = dpg::reduce(z.p, q.p, n, 0, part, one, st);
        dpg::cg_direction<<<bn, 256, 0, st>>>(-zq / pq, z.p, n, p.p);
        SK_CHECK(hipGetLastError());
      }
    }
    if (iters) *iters = it;
    if (relres) *relres = rr0 > 0.0 ? std::sqrt(rr / rr0) : 0.0;
  });
}

}  // extern "C"

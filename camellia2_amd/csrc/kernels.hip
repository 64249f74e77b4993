// kernels.hip -- gfx950 kernels of the DPG local-assembly pipeline.
//
// Per batch of cells of one element type the pipeline is
//   point_kernel      per (cell, cubature point): Jacobian, detJ, J^-1,
//                     normals, measures, coefficients -> small per-point
//                     weight tensors W[group][test var][prim]
//   operand_kernel    per (cell, point, test dof): sum_r W[..][r] * R[q][r][i]
//                     (physical operator values times weights; R are the
//                     reference prims shared by all cells)
//   gemm_tn_f64       batched FP64 MFMA contraction C(i,j) = a*sum_k X(k,i)Y(k,j) + b*C
//                     (Gram matrix, B blocks, and the Cholesky trailing updates)
//   panel_f64         per 64-column panel: LDS Cholesky of the diagonal block
//                     and the triangular solve of the rows below it
//   small kernels     load vector, POEQU scaling, residual row, extraction
//
// All per-cell matrices live in one column-major "augmented" matrix
//   A = [[S G S, S B, S l, S r], [.., 0]]   (lower triangle used)
// of size D = N + M + 2; a partial Cholesky over its first N columns leaves
// -[B l r]^T G^-1 [B l r] in the trailing block, i.e. -K, -rhs and -eta^2.
#include <hip/hip_runtime.h>

#include <utility>

#include <cmath>
#include <cstdint>

#include "../../include/dpg_capi.h"
#include "device_common.hpp"
#include "kernels.hpp"

namespace dpg {

typedef double d4 __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------- geometry
template <int DIM>
struct PointGeo {
  double J[DIM][DIM], Jinv[DIM][DIM], detJ;
};

template <int DIM>
__device__ inline void point_geometry(const double* __restrict__ X, const double* __restrict__ dN, int V,
                                      PointGeo<DIM>& g) {
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int e = 0; e < DIM; ++e) g.J[d][e] = 0.0;
  for (int v = 0; v < V; ++v) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double xv = X[v * DIM + d];
#pragma unroll
      for (int e = 0; e < DIM; ++e) g.J[d][e] = fma(xv, dN[v * DIM + e], g.J[d][e]);
    }
  }
  if constexpr (DIM == 2) {
    g.detJ = g.J[0][0] * g.J[1][1] - g.J[0][1] * g.J[1][0];
    double id = 1.0 / g.detJ;
    g.Jinv[0][0] = g.J[1][1] * id;
    g.Jinv[0][1] = -g.J[0][1] * id;
    g.Jinv[1][0] = -g.J[1][0] * id;
    g.Jinv[1][1] = g.J[0][0] * id;
  } else {
    const double(*A)[DIM] = g.J;
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    g.detJ = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    double id = 1.0 / g.detJ;
    g.Jinv[0][0] = c00 * id;
    g.Jinv[1][0] = c01 * id;
    g.Jinv[2][0] = c02 * id;
    g.Jinv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
    g.Jinv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
    g.Jinv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
    g.Jinv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    g.Jinv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    g.Jinv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  }
}

// coefficients t[r] expressing physical operator `op` of a test dof in its
// reference prims (HGRAD: v, dv/dxi; HDIV: tau_hat, div_hat) -- the transforms
// of BasisEvaluation.cpp:391-650 (grad: J^-T, HDIV value: J/detJ, div: 1/detJ)
template <int DIM>
__device__ inline void op_to_prims(int op, const PointGeo<DIM>& g, const double* n, double* t) {
#pragma unroll
  for (int r = 0; r <= DIM; ++r) t[r] = 0.0;
  if (op == DPG_OP_VALUE) {
    t[0] = 1.0;
  } else if (op >= DPG_OP_DX && op < DPG_OP_DX + DIM) {
    // the direction selected among compile-time indices: a run-time index into
    // g would put the whole geometry in scratch memory
    const int d = op - DPG_OP_DX;
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd)
      if (dd == d) {
#pragma unroll
        for (int e = 0; e < DIM; ++e) t[1 + e] = g.Jinv[e][dd];
      }
  } else if (op >= DPG_OP_X && op < DPG_OP_X + DIM) {
    const int d = op - DPG_OP_X;
    double id = 1.0 / g.detJ;
#pragma unroll
    for (int dd = 0; dd < DIM; ++dd)
      if (dd == d) {
#pragma unroll
        for (int e = 0; e < DIM; ++e) t[e] = g.J[dd][e] * id;
      }
  } else if (op == DPG_OP_DIV) {
    t[DIM] = 1.0 / g.detJ;
  } else if (op == DPG_OP_DOT_NORMAL && n) {
    double id = 1.0 / g.detJ;
#pragma unroll
    for (int e = 0; e < DIM; ++e) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) s += n[d] * g.J[d][e];
      t[e] = s * id;
    }
  }
}

__device__ inline double coef_value(const DSummand& s, const double* cc, const double* n) {
  double v = s.scale;
  if (s.slot >= 0) {
    double a = cc[s.slot];
    v *= (s.fn == DPG_FN_SQRT) ? sqrt(a) : (s.fn == DPG_FN_INV ? 1.0 / a : a);
  }
  if (s.normal >= 0) {
    // the component selected among compile-time indices: a run-time index into
    // the caller's normal would put it in scratch memory
    double nv = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d == s.normal && n) nv = n[d];
    v *= nv;
  }
  return v;
}

template <int DIM>
__device__ inline double load_value(const dpg_load_desc& L, int comp, const double* x) {
  if (comp >= L.ncomp) return 0.0;
  double f = L.constant[comp];
  for (int m = 0; m < L.nmodes[comp]; ++m) {
    double p = L.amp[comp][m];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p *= sin(L.k[comp][m][d] * M_PI * x[d]);
    f += p;
  }
  return f;
}

// ------------------------------------------------------------- point kernel
// kind 0: Gram points, weight sqrt(w detJ), groups = IP terms
// kind 1: stiffness points, weight w detJ, groups = field blocks (+ rhs into Wrhs)
// kind 2: side points, weight w |J t|, unit outward normal, groups = side variables
template <int DIM>
__global__ __launch_bounds__(128) void point_kernel(int kind, const DevProgram* __restrict__ prog,
                                                    DevTables tab, int npts, int Qs, int ncells,
                                                    const double* __restrict__ nodes,
                                                    const int8_t* __restrict__ parities,
                                                    const double* __restrict__ coefs, int ncoef, int nsides,
                                                    dpg_load_desc load, double* __restrict__ W,
                                                    double* __restrict__ Wrhs) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ncells * npts) return;
  int c = gid / npts, p = gid % npts;
  const int V = prog->nverts;
  const int npr = DIM + 1;
  const int nvar = prog->nvar;
  const double* X = nodes + (size_t)c * V * DIM;
  const double* cc = coefs ? coefs + (size_t)c * ncoef : nullptr;
  const double* dN = kind == 0 ? tab.dN_ip : (kind == 1 ? tab.dN_vol : tab.dN_side);
  PointGeo<DIM> g;
  point_geometry<DIM>(X, dN + (size_t)p * V * DIM, V, g);
  double n[3] = {0.0, 0.0, 0.0};
  double fac;
  int s = 0;
  if (kind == 0) {
    fac = sqrt(tab.w_ip[p] * g.detJ);
  } else if (kind == 1) {
    fac = tab.w_vol[p] * g.detJ;
  } else {
    s = p / Qs;
    double T[2][3];
    for (int a = 0; a < DIM - 1; ++a)
      for (int d = 0; d < DIM; ++d) {
        double acc = 0.0;
        for (int e = 0; e < DIM; ++e) acc += g.J[d][e] * tab.side_tan[((size_t)s * 2 + a) * DIM + e];
        T[a][d] = acc;
      }
    if constexpr (DIM == 2) {
      n[0] = T[0][1];
      n[1] = -T[0][0];
    } else {
      n[0] = T[0][1] * T[1][2] - T[0][2] * T[1][1];
      n[1] = T[0][2] * T[1][0] - T[0][0] * T[1][2];
      n[2] = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    }
    double len = 0.0;
    for (int d = 0; d < DIM; ++d) len += n[d] * n[d];
    len = sqrt(len);
    for (int d = 0; d < DIM; ++d) n[d] /= len;
    fac = tab.w_side[p] * len;
  }
  int ng;
  const int32_t* gbeg;
  if (kind == 0) {
    ng = prog->n_lt;
    gbeg = prog->lt_beg;
  } else if (kind == 1) {
    ng = prog->n_vb;
    gbeg = prog->vb_beg;
  } else {
    ng = prog->n_sb;
    gbeg = prog->sb_beg;
  }
  const int slice = ng * nvar * npr;
  double* Wp = W + ((size_t)c * npts + p) * slice;
  for (int e = 0; e < slice; ++e) Wp[e] = 0.0;
  double t[DIM + 1];
  // n stays zero off the sides (normal components and DOT_NORMAL then give 0,
  // as with no normal at all); always passing the array, rather than choosing
  // between it and a null pointer, keeps it in registers
  for (int gi = 0; gi < ng; ++gi) {
    double gfac = fac;
    if (kind == 2 && prog->sb_flux[gi]) gfac *= (double)parities[(size_t)c * nsides + s];
    for (int si = gbeg[gi]; si < gbeg[gi + 1]; ++si) {
      const DSummand& sm = prog->s[si];
      double cv = coef_value(sm, cc, n) * gfac;
      op_to_prims<DIM>(sm.op, g, n, t);
      double* w = Wp + ((size_t)gi * nvar + sm.var) * npr;
      for (int r = 0; r < npr; ++r) w[r] += cv * t[r];
    }
  }
  if (kind == 1 && Wrhs) {
    double x[3] = {0.0, 0.0, 0.0};
    const double* Nv = tab.Nv_vol + (size_t)p * V;
    for (int v = 0; v < V; ++v)
      for (int d = 0; d < DIM; ++d) x[d] += Nv[v] * X[v * DIM + d];
    double* Wr = Wrhs + ((size_t)c * npts + p) * nvar * npr;
    for (int e = 0; e < nvar * npr; ++e) Wr[e] = 0.0;
    for (int si = prog->rhs_beg[0]; si < prog->rhs_beg[1]; ++si) {
      const DSummand& sm = prog->s[si];
      double cv = load_value<DIM>(load, sm.aux, x) * sm.scale * fac;
      op_to_prims<DIM>(sm.op, g, nullptr, t);
      for (int r = 0; r < npr; ++r) Wr[sm.var * npr + r] += cv * t[r];
    }
  }
}

// ----------------------------------------------------------- operand kernel
// out[c*sc + b*sb + p*sp + i] = sum_r W[c][p][b][var(i)][r] * R[p][r][i]
// One workgroup per (256 dofs, kOperandPoints points, cell): the weights of its
// points are staged once, then every thread writes its dof's values point by point.
template <typename T, int NPR>
__global__ __launch_bounds__(256) void operand_kernel(const double* __restrict__ W, int ng, int nvar,
                                                      const double* __restrict__ R, const int32_t* __restrict__ vod,
                                                      int N, int npts, T* __restrict__ out, int64_t sc, int64_t sb,
                                                      int64_t sp) {
  extern __shared__ double Ws[];
  const int p0 = blockIdx.y * kOperandPoints, c = blockIdx.z;
  const int np = min(kOperandPoints, npts - p0);
  const int slice = ng * nvar * NPR;
  const double* Wp = W + ((size_t)c * npts + p0) * slice;
  for (int e = threadIdx.x; e < np * slice; e += blockDim.x) Ws[e] = Wp[e];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int var = vod[i];
  for (int pp = 0; pp < np; ++pp) {
    const int p = p0 + pp;
    double r[NPR];
#pragma unroll
    for (int k = 0; k < NPR; ++k) r[k] = R[((size_t)p * NPR + k) * N + i];
    T* o = out + (size_t)c * sc + (size_t)p * sp + i;
    for (int b = 0; b < ng; ++b) {
      const double* w = Ws + ((size_t)pp * ng * nvar + (size_t)b * nvar + var) * NPR;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < NPR; ++k) acc = fma(w[k], r[k], acc);
      o[(size_t)b * sb] = (T)acc;
    }
  }
}

// load vector: A[c][row_l + i*lda] = sum_q sum_r Wrhs[c][q][var(i)][r] R_vol[q][r][i]
template <int NPR>
__global__ __launch_bounds__(256) void rhs_kernel(const double* __restrict__ Wrhs, int nvar,
                                                  const double* __restrict__ R, const int32_t* __restrict__ vod,
                                                  int N, int Q, double* __restrict__ A, int64_t lda, int64_t sA,
                                                  int row_l, const double* __restrict__ s, int64_t ss) {
  int c = blockIdx.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int var = vod[i];
  double acc = 0.0;
  for (int q = 0; q < Q; ++q) {
    const double* w = Wrhs + ((size_t)c * Q + q) * nvar * NPR + var * NPR;
#pragma unroll
    for (int k = 0; k < NPR; ++k) acc = fma(w[k], R[((size_t)q * NPR + k) * N + i], acc);
  }
  A[(size_t)c * sA + row_l + (size_t)i * lda] = s ? acc * s[(size_t)c * ss + i] : acc;
}

// Gram operand (one thread per test dof j, loop over the points in a fixed
// order): out[c][q*n_lt + lt][j] = sum_r Wip[c][q][lt][var j][r] R[q][r][j]
// (the weighted values of IP term lt; G = out^T out), and the DPOEQU factors
// s[c][j] = 1/sqrt(G_jj) with G_jj = sum over rows of out[..][j]^2.
template <int NPR>
__global__ __launch_bounds__(256) void ip_operand_kernel(const double* __restrict__ Wip, int n_lt, int nvar,
                                                         const double* __restrict__ R,
                                                         const int32_t* __restrict__ vod, int N, int QG, int QGp,
                                                         double* __restrict__ out, int64_t sc, int64_t ld,
                                                         double* __restrict__ s, int64_t ss) {
  extern __shared__ double Wl[];
  const int wslice = n_lt * nvar * NPR;
  const int c = blockIdx.y, t = threadIdx.x;
  const int j = blockIdx.x * blockDim.x + t;
  const bool active = j < N;
  const int var = active ? vod[j] : 0;
  double diag = 0.0;
  double* oc = out + (size_t)c * sc + j;
  for (int q = 0; q < QG; ++q) {
    __syncthreads();
    const double* Wp = Wip + ((size_t)c * QG + q) * wslice;
    for (int e = t; e < wslice; e += blockDim.x) Wl[e] = Wp[e];
    __syncthreads();
    if (!active) continue;
    double r[NPR];
#pragma unroll
    for (int k = 0; k < NPR; ++k) r[k] = R[((size_t)q * NPR + k) * N + j];
    for (int lt = 0; lt < n_lt; ++lt) {
      const double* w = Wl + (lt * nvar + var) * NPR;
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < NPR; ++k) acc = fma(w[k], r[k], acc);
      oc[((size_t)lt * QGp + q) * ld] = acc;  // term-major rows: term lt, point q
      diag = fma(acc, acc, diag);
    }
  }
  if (!active) return;
  for (int lt = 0; lt < n_lt; ++lt)
    for (int q = QG; q < QGp; ++q) oc[((size_t)lt * QGp + q) * ld] = 0.0;
  s[(size_t)c * ss + j] = 1.0 / sqrt(diag);
}

// lane `k`'s value of a double (k per lane), through the LDS crossbar
// (ds_bpermute: no LDS allocation, result in vector registers)
__device__ inline double wave_bcast(double v, int k) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_ds_bpermute(k << 2, (int)(b & 0xffffffffLL));
  int hi = __builtin_amdgcn_ds_bpermute(k << 2, (int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// the value of lane (lane & 15) + 16 qs of a double in every lane: row qs of
// the wave's four 16-lane rows broadcast to all rows by two row swaps,
// v_permlane16_swap (rows 0,1 and 2,3) then v_permlane32_swap (halves); qs is
// a constant after unrolling, so the choices fold away
__device__ inline int row_bcast32(int x, int qs) {
  const auto a = __builtin_amdgcn_permlane16_swap(x, x, false, false);  // [0]: rows 0,0,2,2  [1]: 1,1,3,3
  const int y = (qs & 1) ? a[1] : a[0];
  const auto b = __builtin_amdgcn_permlane32_swap(y, y, false, false);  // [0]: rows 0,1,0,1  [1]: 2,3,2,3
  return (qs & 2) ? b[1] : b[0];
}
__device__ inline double row_bcast(double v, int qs) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = row_bcast32((int)(b & 0xffffffffLL), qs), hi = row_bcast32((int)(b >> 32), qs);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// lane `k`'s value of a double for a wave-uniform k, read into scalar
// registers (v_readlane: no trip through the LDS crossbar)
__device__ inline double wave_read(double v, int k) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), k);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// ------------------------------------------------------------------- GEMM
// C(i,j) = alpha * sum_k X(k,i) Y(k,j) + beta * C(i,j), X(k,i) = X[k*ldx + i],
// C col-major (C[i + j*ldc]); 64x64 tiles, 4 waves, v_mfma_f64_16x16x4.
constexpr int GT = 64, GKC = 16;  // 64x64 tiles, k-chunks of 16

__device__ inline int xcd_remap(int b, int nb) {
  // consecutive logical tiles on one XCD (blocks are dealt round-robin to 8 XCDs)
  int per = nb / 8, rem = nb % 8, x = b % 8, idx = b / 8;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + idx;
}

__device__ inline void tile_coords(int tile, int lower, int tiles_m, int& ti, int& tj) {
  if (lower) {
    int t = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= tile) ++t;
    while (t * (t + 1) / 2 > tile) --t;
    ti = t;
    tj = tile - t * (t + 1) / 2;
  } else {
    ti = tile % tiles_m;
    tj = tile / tiles_m;
  }
}

// TT x TT output tile per workgroup of (TT/WT)^2 waves; each wave owns a WT x WT
// sub-tile (WT/16 x WT/16 MFMA blocks). TT = 64 (WT 32) for large operands,
// TT = 32 for buckets whose dimensions would leave most of a 64-tile empty.
// (A 128-tile of four 64 x 64 waves was measured slower on every role: its
// ~230 registers per lane leave two waves per SIMD, too few to hide the loads.)
// k runs in chunks of GKC rows staged through LDS; each thread stages one fixed
// column of X and of Y (rows t/TT + r*NT/TT of a chunk), from a clamped,
// always valid address, zeroed outside the matrix (no branches per element).
// Global loads run two chunks ahead of the MFMAs (two register sets).
template <int TT, int WT = 32>
__device__ inline void gemm_tn_f64_body(const GemmArgs& g) {
  constexpr int WB = WT / 16;                     // MFMA blocks per wave and dimension
  constexpr int NT = (TT / WT) * (TT / WT) * 64;  // threads
  constexpr int PER = GKC * TT / NT;              // staged elements per thread and operand
  constexpr int RS = NT / TT;                     // k-row step between a thread's elements
  constexpr int LD = TT + 4;
  int b = xcd_remap(blockIdx.x, gridDim.x);
  int cell = b / g.ntiles, tile = b % g.ntiles;
  if (cell >= g.ncells) return;
  int ti, tj;
  tile_coords(tile, g.lower, g.tiles_m, ti, tj);
  const int i0 = ti * TT, j0 = tj * TT;
  double* __restrict__ C = g.C + (size_t)cell * g.sc;
  // two buffers for the 64-tile (chunk i is read from one while chunk i+1 is
  // written to the other); one for the 32-tile of one wave, whose small
  // workgroups would otherwise be limited by LDS rather than by registers
  constexpr int NBUF = TT >= 64 ? 2 : 1;
  __shared__ double Xs[NBUF][GKC][LD];
  __shared__ double Ys[NBUF][GKC][LD];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w / (TT / WT), wc = w % (TT / WT);
  d4 acc[WB][WB];
#pragma unroll
  for (int a = 0; a < WB; ++a)
#pragma unroll
    for (int bb = 0; bb < WB; ++bb) acc[a][bb] = d4{0.0, 0.0, 0.0, 0.0};
  const int kend = g.k;
  const int kk0 = t / TT, ii = t % TT;
  const bool vx = i0 + ii < g.m, vy = j0 + ii < g.n;
  const double* __restrict__ xp = g.X + (size_t)cell * g.sx + (vx ? i0 + ii : 0) + (size_t)kk0 * g.ldx;
  const double* __restrict__ yp = g.Y + (size_t)cell * g.sy + (vy ? j0 + ii : 0) + (size_t)kk0 * g.ldy;
  // k-chunks of segments that are zero for this tile are skipped: walk the
  // segments [s*kseg, (s+1)*kseg) with their mask bits (no division)
  const uint64_t km = g.kmask ? g.kmask[tile] : ~0ull;
  const int kseg = g.kmask ? (int)g.kseg : kend;
  int seg = 0, segend = kseg;
  auto skip_zero_segments = [&](int k0) {
    while (k0 < kend && !((km >> seg) & 1ull)) {
      k0 = segend;
      ++seg;
      segend += kseg;
    }
    return k0;
  };
  auto next_chunk = [&](int k0) {
    if (k0 >= kend) return k0;
    k0 += GKC;
    if (k0 >= segend) {
      ++seg;
      segend += kseg;
      k0 = skip_zero_segments(k0);
    }
    return k0;
  };
  // raw loads only: the zeroing happens when the chunk is staged, so that no
  // wait for the loads is needed before then
  const int ldx = (int)g.ldx, ldy = (int)g.ldy;  // per-cell offsets fit in 32 bits
  auto fetch = [&](double (&xr)[PER], double (&yr)[PER], int k0) {
    if (k0 + GKC <= kend) {
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        xr[r] = xp[(k0 + r * RS) * ldx];
        yr[r] = yp[(k0 + r * RS) * ldy];
      }
    } else {  // last, partial chunk (k not a multiple of GKC): clamp the row
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int kr = min(k0 + kk0 + r * RS, kend - 1) - kk0;
        xr[r] = xp[kr * ldx];
        yr[r] = yp[kr * ldy];
      }
    }
  };
  // weight of the segment the walker is in (kw: merged equal terms), 1 without
  auto segw = [&]() { return (g.kw && seg < g.nseg) ? g.kw[(size_t)tile * g.nseg + seg] : 1.0; };
  auto stage = [&](const double (&xr)[PER], const double (&yr)[PER], int k0, double wx, int buf) {
    buf &= NBUF - 1;
    if constexpr (NBUF == 1) __syncthreads();  // every wave is done reading the buffer
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const bool vk = k0 + kk0 + r * RS < kend;
      Xs[buf][kk0 + r * RS][ii] = (vx && vk) ? xr[r] * wx : 0.0;
      Ys[buf][kk0 + r * RS][ii] = (vy && vk) ? yr[r] : 0.0;
    }
  };
  auto multiply = [&](int buf) {
    buf &= NBUF - 1;
#pragma unroll
    for (int k4 = 0; k4 < GKC; k4 += 4) {
      const int kr = k4 + (lane >> 4), cc = lane & 15;
      double xa[WB], yb[WB];
#pragma unroll
      for (int q = 0; q < WB; ++q) {
        xa[q] = Xs[buf][kr][wr * WT + q * 16 + cc];
        yb[q] = Ys[buf][kr][wc * WT + q * 16 + cc];
      }
#pragma unroll
      for (int bj = 0; bj < WB; ++bj)
#pragma unroll
        for (int bi = 0; bi < WB; ++bi)
          acc[bj][bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(yb[bj], xa[bi], acc[bj][bi], 0, 0, 0);
    }
  };
  double xa_[PER], ya_[PER], xb_[PER], yb_[PER];
  int kc = skip_zero_segments(0);  // chunk to multiply next
  double wa = segw();              // each register set carries its chunk's segment weight
  int kn = next_chunk(kc);         // the one after
  double wb = segw();
  if (kc < kend) fetch(xa_, ya_, kc);
  if (kn < kend) fetch(xb_, yb_, kn);
  // software pipeline, unrolled by two so that the register sets and the LDS
  // buffers alternate without copies: while chunk i is multiplied from one
  // buffer, chunk i+1 goes from registers to the other and chunk i+2 is loaded
  // into the registers chunk i+1 left; one barrier per chunk
  // a wave whose WT x WT sub-tile holds no stored element (all rows >= m, all
  // columns >= n, or strictly above the diagonal of a lower tile) stages its
  // share but issues no MFMAs, leaving the SIMD's matrix pipe to other waves
  const int si0 = i0 + wr * WT, sj0 = j0 + wc * WT;
  const bool busy = si0 < g.m && sj0 < g.n && ((!g.lower && !g.skip_upper) || si0 + WT - 1 >= sj0);
  if (kc < kend) {
    stage(xa_, ya_, kc, wa, 0);
    int kf = next_chunk(kn);  // buffer 0: kc; set b: kn; set a: kf
    wa = segw();
    if (kf < kend) fetch(xa_, ya_, kf);
    __syncthreads();
    while (true) {
      if (busy) multiply(0);
      if (kn >= kend) break;
      stage(xb_, yb_, kn, wb, 1);
      int kg = next_chunk(kf);
      wb = segw();
      if (kg < kend) fetch(xb_, yb_, kg);
      __syncthreads();  // buffer 1: kn; set a: kf; set b: kg
      if (busy) multiply(1);
      if (kf >= kend) break;
      stage(xa_, ya_, kf, wa, 0);
      int kh = next_chunk(kg);
      wa = segw();
      if (kh < kend) fetch(xa_, ya_, kh);
      __syncthreads();  // buffer 0: kf; set b: kg; set a: kh
      kn = kg;
      kf = kh;
    }
  }
  // v_mfma_f64_16x16x4 result layout: D[a][b] in lane b + 16*(a%4), element a/4,
  // with a the A-operand (Y, C column j) and b the B-operand (X, C row i) index.
  // This thread's rows are i0 + wr*WT + bi*16 + (lane & 15) and its columns
  // j0 + wc*WT + bj*16 + (lane >> 4) + 4v. The epilogue runs in WB parts (bj):
  // each issues all its loads (scales, C when beta != 0) from clamped addresses
  // before using any, one memory round trip per half instead of one per element.
  int ir[WB];
  bool iv[WB];
  double sr[WB];
#pragma unroll
  for (int bi = 0; bi < WB; ++bi) {
    sr[bi] = 1.0;
    ir[bi] = i0 + wr * WT + bi * 16 + (lane & 15);
    iv[bi] = ir[bi] < g.m;
    if (g.srow) sr[bi] = g.srow[(size_t)cell * g.ss + (iv[bi] ? ir[bi] : 0)];
  }
#pragma unroll
  for (int bj = 0; bj < WB; ++bj) {
    int jc[4];
    bool jv[4];
    double sc[4] = {1.0, 1.0, 1.0, 1.0}, cold[WB][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      jc[v] = j0 + wc * WT + bj * 16 + (lane >> 4) + 4 * v;
      jv[v] = jc[v] < g.n;
      if (g.scol) sc[v] = g.scol[(size_t)cell * g.ss + (jv[v] ? jc[v] : 0)];
    }
    if (g.beta != 0.0) {
#pragma unroll
      for (int bi = 0; bi < WB; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          cold[bi][v] = C[(size_t)(iv[bi] ? ir[bi] : 0) + (size_t)(jv[v] ? jc[v] : 0) * g.ldc];
    }
    // one wait for all of them (vmcnt 0, expcnt and lgkmcnt untouched): without it
    // the compiler re-waits in every conditional store below, stores included
    __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
    for (int bi = 0; bi < WB; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = ir[bi], j = jc[v];
        if (iv[bi] && jv[v] && ((!g.lower && !g.skip_upper) || i >= j)) {
          double val = g.alpha * acc[bj][bi][v];  // (alpha acc) s_i s_j: multiplying by 1 is exact
          val *= sr[bi];
          val *= sc[v];
          if (g.beta != 0.0) val = fma(g.beta, cold[bi][v], val);
          C[(size_t)i + (size_t)j * g.ldc] = val;
        }
      }
  }
}

// 128 x 64 output tile (rows of C x columns) per workgroup of 4 waves for the
// tall panel updates of the factorisation (m = rows below the panel, n = the
// panel's 64 columns): wave w owns rows 32w..32w+31 and all 64 columns, 2 x 4
// MFMA blocks, so an 8-row k-chunk staged between two barriers feeds 16 MFMAs
// per wave and each X row is staged once for the whole panel. Chunks of 8 rows
// rather than the square tile's 16 keep the kernel within 128 VGPRs, 4 waves
// per SIMD (3% faster on the tall panels than 16-row chunks at 3 waves, which
// hold the operands of twice the k-rows in registers). A panel of at most 32
// columns (the short last panel of a factorisation) takes the 128 x 32 form of
// the same tile (TN = 32, 2 x 2 MFMA blocks per wave) rather than half-empty
// blocks. Same staging scheme as above (fixed columns per thread, clamped
// addresses, zeroing at LDS write, two chunks in flight) and the same k order
// per output element, so the results are those of the square tiles bit for
// bit. No k-segment masks, no lower-triangle tile walk (dense m x n).
//
// SOLVE (rows below a panel's diagonal block): the updated tile is then solved
// in registers against the diagonal factor, X = C L^-T, before it is stored,
// so the panel is written once instead of written, read and written again.
// The solve is blocked by the 16-column MFMA blocks: inside a block the
// substitution runs column by column as the reference DTRSM does (x_s = c_s /
// L_ss, then c_m -= L_ms x_s for the later columns m of the block), with x_s
// moved between the four lane groups that hold a row's columns by two row
// swaps (row_bcast); the later blocks then take X_block L[later, block]^T as MFMAs
// (a rank-4 update per k-slice, the B operand being the solved accumulator
// element itself). L stays in LDS (the staging buffers, reused).
constexpr int WKC = 8;  // k-rows per staged chunk of the 128 x 64 tile
template <bool SOLVE, int TN>
__device__ inline void gemm_tn_f64_wide_body(const GemmArgs& g) {
  constexpr int TM = 128, NT = 256, NB = TN / 16;  // NB column blocks of 16
  constexpr int PX = WKC * TM / NT, RX = NT / TM;  // X: 4 elements per thread, k-rows t/128 + 2r
  constexpr int PY = WKC * TN / NT, RY = NT / TN;  // Y: 2 (TN 64) or 1 elements per thread, k-rows t/TN + RY r
  constexpr int LXS = TM + 4, LYS = TN + 4, LP = 68;  // LDS row pitches (L: one column of the factor per row)
  constexpr int NSTAGE = 2 * WKC * (LXS + LYS), NL = SOLVE ? TN * LP : 0;  // two staging buffers; L's first TN columns
  int b = xcd_remap(blockIdx.x, gridDim.x);
  int cell = b / g.ntiles, tile = b % g.ntiles;
  if (cell >= g.ncells) return;
  const int ti = tile % g.tiles_m, tj = tile / g.tiles_m;
  const int i0 = ti * TM, j0 = tj * TN;
  double* __restrict__ C = g.C + (size_t)cell * g.sc;
  __shared__ double smem[NSTAGE > NL ? NSTAGE : NL];
  double (*Xs)[WKC][LXS] = reinterpret_cast<double (*)[WKC][LXS]>(smem);
  double (*Ys)[WKC][LYS] = reinterpret_cast<double (*)[WKC][LYS]>(smem + 2 * WKC * LXS);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  d4 acc[NB][2];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) acc[a][bb] = d4{0.0, 0.0, 0.0, 0.0};
  const int kend = g.k;
  const int kx0 = t / TM, ix = t % TM, ky0 = t / TN, iy = t % TN;
  const bool vx = i0 + ix < g.m, vy = j0 + iy < g.n;
  const double* __restrict__ xp = g.X + (size_t)cell * g.sx + (vx ? i0 + ix : 0) + (size_t)kx0 * g.ldx;
  const double* __restrict__ yp = g.Y + (size_t)cell * g.sy + (vy ? j0 + iy : 0) + (size_t)ky0 * g.ldy;
  const int ldx = (int)g.ldx, ldy = (int)g.ldy;
  auto fetch = [&](double (&xr)[PX], double (&yr)[PY], int k0) {
    if (k0 + WKC <= kend) {
#pragma unroll
      for (int r = 0; r < PX; ++r) xr[r] = xp[(k0 + r * RX) * ldx];
#pragma unroll
      for (int r = 0; r < PY; ++r) yr[r] = yp[(k0 + r * RY) * ldy];
    } else {
#pragma unroll
      for (int r = 0; r < PX; ++r) xr[r] = xp[(min(k0 + kx0 + r * RX, kend - 1) - kx0) * ldx];
#pragma unroll
      for (int r = 0; r < PY; ++r) yr[r] = yp[(min(k0 + ky0 + r * RY, kend - 1) - ky0) * ldy];
    }
  };
  auto stage = [&](const double (&xr)[PX], const double (&yr)[PY], int k0, int buf) {
#pragma unroll
    for (int r = 0; r < PX; ++r) Xs[buf][kx0 + r * RX][ix] = (vx && k0 + kx0 + r * RX < kend) ? xr[r] : 0.0;
#pragma unroll
    for (int r = 0; r < PY; ++r) Ys[buf][ky0 + r * RY][iy] = (vy && k0 + ky0 + r * RY < kend) ? yr[r] : 0.0;
  };
  auto multiply = [&](int buf) {
#pragma unroll
    for (int k4 = 0; k4 < WKC; k4 += 4) {
      const int kr = k4 + (lane >> 4), cc = lane & 15;
      double xa[2], yb[NB];
      xa[0] = Xs[buf][kr][w * 32 + cc];
      xa[1] = Xs[buf][kr][w * 32 + 16 + cc];
#pragma unroll
      for (int bj = 0; bj < NB; ++bj) yb[bj] = Ys[buf][kr][bj * 16 + cc];
#pragma unroll
      for (int bj = 0; bj < NB; ++bj)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
          acc[bj][bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(yb[bj], xa[bi], acc[bj][bi], 0, 0, 0);
    }
  };
  double xa_[PX], ya_[PY], xb_[PX], yb_[PY];
  int kc = 0, kn = WKC;
  if (kc < kend) fetch(xa_, ya_, kc);
  if (kn < kend) fetch(xb_, yb_, kn);
  const int si0 = i0 + w * 32;
  const bool busy = si0 < g.m && j0 < g.n && (!g.skip_upper || si0 + 31 >= j0);
  // two LDS buffers and two register sets, one barrier per chunk (as in the
  // square tile)
  if (kc < kend) {
    stage(xa_, ya_, kc, 0);
    int kf = kn + WKC;  // buffer 0: kc; set b: kn; set a: kf
    if (kf < kend) fetch(xa_, ya_, kf);
    __syncthreads();
    while (true) {
      if (busy) multiply(0);
      if (kn >= kend) break;
      stage(xb_, yb_, kn, 1);
      const int kg = kf + WKC;
      if (kg < kend) fetch(xb_, yb_, kg);
      __syncthreads();  // buffer 1: kn; set a: kf; set b: kg
      if (busy) multiply(1);
      if (kf >= kend) break;
      stage(xa_, ya_, kf, 0);
      const int kh = kg + WKC;
      if (kh < kend) fetch(xa_, ya_, kh);
      __syncthreads();  // buffer 0: kf; set b: kg; set a: kh
      kn = kg;
      kf = kh;
    }
  }
  // epilogue as in the square tile, by 16-column blocks (bj)
  int ir[2];
  bool iv[2];
  double sr[2] = {1.0, 1.0};
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
    ir[bi] = si0 + bi * 16 + (lane & 15);
    iv[bi] = ir[bi] < g.m;
    if (!SOLVE && g.srow) sr[bi] = g.srow[(size_t)cell * g.ss + (iv[bi] ? ir[bi] : 0)];
  }
  if constexpr (SOLVE) {
    // the updated tile C = beta C + alpha acc (0 outside the matrix), in place
    // of the accumulators
#pragma unroll
    for (int bj = 0; bj < NB; ++bj) {
      int jc[4];
      bool jv[4];
      double cold[2][4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        jc[v] = j0 + bj * 16 + (lane >> 4) + 4 * v;
        jv[v] = jc[v] < g.n;
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          cold[bi][v] = C[(size_t)(iv[bi] ? ir[bi] : 0) + (size_t)(jv[v] ? jc[v] : 0) * g.ldc];
      __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          acc[bj][bi][v] = (iv[bi] && jv[v]) ? fma(g.beta, cold[bi][v], g.alpha * acc[bj][bi][v]) : 0.0;
    }
    // the diagonal factor into LDS (the staging buffers are free once every
    // wave has left the k loop): row j of Ls = column j of L
    __syncthreads();
    double* Ls = smem;
    const double* __restrict__ Lc = g.Lt + (size_t)cell * 64 * 64;
    for (int e = t; e < TN * 64; e += NT) Ls[(e >> 6) * LP + (e & 63)] = Lc[e];
    __syncthreads();
    const int q = lane >> 4, rowl = lane & 15;
#pragma unroll
    for (int bj0 = 0; bj0 < NB; ++bj0) {
      // substitution over the 16 columns of block bj0; column s of the block is
      // element s/4 of lane group s%4
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int vs = s >> 2, qs = s & 3, js = bj0 * 16 + s;
        const double dinv = Ls[js * LP + js];
        double lm[4];
#pragma unroll
        for (int v = vs; v < 4; ++v) lm[v] = Ls[js * LP + bj0 * 16 + q + 4 * v];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          const double xs = row_bcast(acc[bj0][bi][vs], qs) * dinv;
          acc[bj0][bi][vs] = q == qs ? xs : (q > qs ? fma(-lm[vs], xs, acc[bj0][bi][vs]) : acc[bj0][bi][vs]);
#pragma unroll
          for (int v = vs + 1; v < 4; ++v) acc[bj0][bi][v] = fma(-lm[v], xs, acc[bj0][bi][v]);
        }
      }
      // the later blocks: C[:, bj1] -= X[:, bj0] L[bj1, bj0]^T, a rank-4 MFMA
      // per k-slice v0 (columns bj0*16 + 4 v0 + k, k = lane group)
#pragma unroll
      for (int bj1 = bj0 + 1; bj1 < NB; ++bj1)
#pragma unroll
        for (int v0 = 0; v0 < 4; ++v0) {
          const double la = -Ls[(bj0 * 16 + 4 * v0 + q) * LP + bj1 * 16 + rowl];
#pragma unroll
          for (int bi = 0; bi < 2; ++bi)
            acc[bj1][bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(la, acc[bj0][bi][v0], acc[bj1][bi], 0, 0, 0);
        }
    }
#pragma unroll
    for (int bj = 0; bj < NB; ++bj)
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = ir[bi], j = j0 + bj * 16 + (lane >> 4) + 4 * v;
          if (iv[bi] && j < g.n) C[(size_t)i + (size_t)j * g.ldc] = acc[bj][bi][v];
        }
    return;
  }
#pragma unroll
  for (int bj = 0; bj < NB; ++bj) {
    int jc[4];
    bool jv[4];
    double sc[4] = {1.0, 1.0, 1.0, 1.0}, cold[2][4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      jc[v] = j0 + bj * 16 + (lane >> 4) + 4 * v;
      jv[v] = jc[v] < g.n;
      if (g.scol) sc[v] = g.scol[(size_t)cell * g.ss + (jv[v] ? jc[v] : 0)];
    }
    if (g.beta != 0.0) {
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          cold[bi][v] = C[(size_t)(iv[bi] ? ir[bi] : 0) + (size_t)(jv[v] ? jc[v] : 0) * g.ldc];
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = ir[bi], j = jc[v];
        if (iv[bi] && jv[v] && (!g.skip_upper || i >= j)) {
          double val = g.alpha * acc[bj][bi][v];
          val *= sr[bi];
          val *= sc[v];
          if (g.beta != 0.0) val = fma(g.beta, cold[bi][v], val);
          C[(size_t)i + (size_t)j * g.ldc] = val;
        }
      }
  }
}

// one kernel symbol per role and tile so that profiles separate them
__global__ __launch_bounds__(256) void gemm_gram_f64(GemmArgs g) { gemm_tn_f64_body<64>(g); }
__global__ __launch_bounds__(256) void gemm_bmat_f64(GemmArgs g) { gemm_tn_f64_body<64>(g); }
__global__ __launch_bounds__(256) void gemm_update_f64(GemmArgs g) { gemm_tn_f64_body<64>(g); }
__global__ __launch_bounds__(64) void gemm_gram_f64_t32(GemmArgs g) { gemm_tn_f64_body<32>(g); }
__global__ __launch_bounds__(64) void gemm_bmat_f64_t32(GemmArgs g) { gemm_tn_f64_body<32>(g); }
__global__ __launch_bounds__(64) void gemm_update_f64_t32(GemmArgs g) { gemm_tn_f64_body<32>(g); }
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void gemm_update_f64_w128(GemmArgs g) {
  gemm_tn_f64_wide_body<false, 64>(g);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void update_solve_f64(GemmArgs g) {
  gemm_tn_f64_wide_body<true, 64>(g);
}
// panels of at most 32 columns (the short last panel): half the MFMA blocks
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void update_solve_f64_n32(GemmArgs g) {
  gemm_tn_f64_wide_body<true, 32>(g);
}

// ------------------------------------------------- sum-factorised Gram
// One workgroup per (block pair, cell). G_AB[i, j] with i = (i1, i2, i3),
// j = (j1, j2, j3), x fastest, is
//   sum_pairs sum_q M_p(q) F_fa1(i1,q1) F_fb1(j1,q1) F_fa2(i2,q2) F_fb2(j2,q2) F_fa3(i3,q3) F_fb3(j3,q3)
// evaluated direction by direction for each slab (i3, j3) (3D) or (i2, j2) (2D):
//   T1[p][q2][q1]      = sum_q3 F_fa3(i3,q3) F_fb3(j3,q3) M_p(q1,q2,q3)
//   T2[cls][i2,j2][q1] = sum_{p in cls} sum_q2 F_fa2(i2,q2) F_fb2(j2,q2) T1[p][q2][q1]
//   G[(i1,i2),(j1,j2)] = sum_cls sum_q1 F_cfa(i1,q1) F_cfb(j1,q1) T2[cls][i2,j2][q1]
// where cls groups the pairs by their direction-1 families. M_p is formed
// from the point weights W in LDS first. Writes are unscaled, or S G S when
// the DPOEQU factors s are given (the products of the scaling pass, in its
// order: (g s_i) s_j). A block pair with no pairs writes its zeros. QN > 0: the points per
// direction as a compile-time constant (the point sums unrolled, their loads
// issued together; the same FMAs in the same order as with QN = 0, which takes
// q at run time).
//
// MODE 3 (gram_sf_diag_kernel): G's diagonal only (the diagonal block pairs'
// diagonal slabs and entries, each the same FMAs as in the full pass), so that
// the DPOEQU factors exist before the full pass, which then stores S G S.
//
// MODE 1 (the volume part of B, bmat_sf_kernel): side A is a trial field block
// (rows N + its offset), side B a test block; the pair's weight is one point
// weight of the stiffness program, M_p(q) = W[q][a] (a = (group * nvar + test
// var) * npr + prim), and the stores carry the DPOEQU factor of the test column.
//
// MODE 2 (the side part of B, bmat_side_sf_kernel): side A is a trace or flux
// block on one side, side B a test block restricted to that side; the kernel's
// directions are the side's coordinates and last the normal one, whose single
// point fills the first q^(DIM-1) of the q^DIM point slots (the others zero);
// each block pair brings its side's 1D tables (foff) and point weights (woff),
// and its columns' strides (sB) keep the test block's reference order.
template <int DIM, int QN, int MODE>
__device__ inline void sf_body(const SfBlockPair* __restrict__ bps, const SfPair* __restrict__ pairs,
                               const double* __restrict__ F, int nF, int nfam, int q_rt,
                               const double* __restrict__ Wp, int n_lt, int na, int QG, double* __restrict__ A,
                               int64_t lda, int64_t sA, const double* __restrict__ s, int64_t ss) {
  const int q = QN > 0 ? QN : q_rt;
  extern __shared__ double sm[];
  // workgroups are dealt round-robin to the XCDs in launch order (x fastest):
  // remapped so that each XCD takes every block pair of a run of cells rather
  // than a fixed few block pairs of every cell (which leaves XCDs unevenly
  // loaded whenever 8 divides the number of block pairs, and splits a cell's
  // point weights over eight L2s)
  const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int bpi = lin % gridDim.x, c = lin / gridDim.x;
  const SfBlockPair& bp = bps[bpi];  // read in place: a local copy, indexed at run time, would live in scratch
  const int t = threadIdx.x, NT = blockDim.x;
  const int np = bp.npairs, ncls = bp.ncls;
  const SfPair* pr = pairs + bp.pair0;
  if (np == 0) {
    // no term couples the two blocks: their entries are zeros (written, as the
    // dense contraction writes them, over whatever the last call left there)
    double* Az = A + (size_t)c * sA;
    const int nAt = bp.nA[0] * bp.nA[1] * bp.nA[2], nB01 = bp.nB[0] * bp.nB[1];
    for (int e = t; e < nAt * nB01 * bp.nB[2]; e += NT) {
      const int ia = e % nAt, jb = e / nAt;
      const int j1 = jb % bp.nB[0], j2 = (jb / bp.nB[0]) % bp.nB[1], j3 = jb / nB01;
      const int row = bp.offA + ia, col = bp.offB + j1 * bp.sB[0] + j2 * bp.sB[1] + j3 * bp.sB[2];
      if (bp.diag && row < col) continue;
      Az[(size_t)row + (size_t)col * lda] = 0.0;
    }
    return;  // the whole workgroup: np is the block pair's
  }
  double* Ms = sm;                 // np * QG
  double* Fs = Ms + np * QG;       // nfam * nF * q
  double* T1 = Fs + nfam * nF * q;  // 3D: np*q*q   2D: ncls*q
  double* T2 = T1 + (DIM == 3 ? np * q * q : ncls * q);  // 3D: ncls * nA2*nB2 * q
  if constexpr (MODE == 0 || MODE == 3) {
    const double* W = Wp + (size_t)c * QG * n_lt * na + bp.woff;
    for (int e = t; e < np * QG; e += NT) {
      int p = e / QG, qq = e - p * QG;
      const double* w = W + (size_t)qq * n_lt * na;
      int a = pr[p].a, b = pr[p].b;
      double acc = 0.0;
      for (int lt = 0; lt < n_lt; ++lt) acc = fma(w[lt * na + a], w[lt * na + b], acc);
      Ms[e] = acc;
    }
  } else if constexpr (MODE == 1) {
    const double* W = Wp + (size_t)c * QG * na + bp.woff;
    for (int e = t; e < np * QG; e += NT) {
      int p = e / QG, qq = e - p * QG;
      Ms[e] = W[(size_t)qq * na + pr[p].a];
    }
  } else {
    // a side: its points fill the first q^(DIM-1) of the q^DIM slots (the last
    // direction is the normal one, one point); n_lt holds the sides' points per cell
    const int qf = DIM == 3 ? q * q : q;
    const double* W = Wp + (size_t)c * n_lt * na + bp.woff;
    for (int e = t; e < np * QG; e += NT) {
      int p = e / QG, qq = e - p * QG;
      Ms[e] = qq < qf ? W[(size_t)qq * na + pr[p].a] : 0.0;
    }
  }
  const double* sc = s ? s + (size_t)c * ss : nullptr;
  if (MODE == 3 && !bp.diag) return;  // the whole workgroup
  for (int e = t; e < nfam * nF * q; e += NT) Fs[e] = F[bp.foff + e];
  __syncthreads();
  double* Ac = A + (size_t)c * sA;
  auto fam = [&](int f, int i) { return Fs + (f * nF + i) * q; };
  // the last direction for a BI x BJ block of outputs, rows r0 + i1 + {0..BI),
  // columns c0 + (j1 + {0..BJ)) * sB[0]: G[i, j] = sum_cls sum_q1 F_cfa(i1,q1)
  // F_cfb(j1,q1) tp[cls * cstride + q1]. Each output sees the same FMAs in the
  // same order as alone; a block shares its loads of tp and of each family row.
  // Outputs past the block's edge (odd sizes) or above a diagonal block's
  // diagonal are computed but not stored. 2 x 2 for G (LDS already limits its
  // workgroups per CU), 1 x 1 for B, whose small workgroups are limited by
  // registers.
  auto final_block = [&](auto bi_c, auto bj_c, int i1, int j1, int r0, int c0, const double* tp0, int cstride) {
    constexpr int BI = decltype(bi_c)::value, BJ = decltype(bj_c)::value;
    const int cs = bp.sB[0];
    int ii[BI], jj[BJ];
#pragma unroll
    for (int d = 0; d < BI; ++d) ii[d] = min(i1 + d, bp.nA[0] - 1);
#pragma unroll
    for (int d = 0; d < BJ; ++d) jj[d] = min(j1 + d, bp.nB[0] - 1);
    if (bp.diag && r0 + ii[BI - 1] < c0 + j1 * cs) return;  // all above the diagonal
    double acc[BI][BJ];
#pragma unroll
    for (int da = 0; da < BI; ++da)
#pragma unroll
      for (int db = 0; db < BJ; ++db) acc[da][db] = 0.0;
    for (int cl = 0; cl < ncls; ++cl) {
      const double* fa[BI];
      const double* fb[BJ];
#pragma unroll
      for (int d = 0; d < BI; ++d) fa[d] = fam(bp.clsfa[cl], ii[d]);
#pragma unroll
      for (int d = 0; d < BJ; ++d) fb[d] = fam(bp.clsfb[cl], jj[d]);
      const double* tp = tp0 + cl * cstride;
#pragma unroll
      for (int q1 = 0; q1 < q; ++q1) {
        const double tq = tp[q1];
#pragma unroll
        for (int db = 0; db < BJ; ++db)
#pragma unroll
          for (int da = 0; da < BI; ++da) acc[da][db] = fma(fa[da][q1] * fb[db][q1], tq, acc[da][db]);
      }
    }
#pragma unroll
    for (int da = 0; da < BI; ++da)
#pragma unroll
      for (int db = 0; db < BJ; ++db) {
        const int row = r0 + i1 + da, col = c0 + (j1 + db) * cs;
        if (i1 + da >= bp.nA[0] || j1 + db >= bp.nB[0] || (bp.diag && row < col)) continue;
        // G: S G S as the scaling pass computes it ((g s_i) s_j); B: g s_j
        double v = acc[da][db];
        if (sc) v = MODE == 0 ? v * sc[row] * sc[col] : v * sc[col];
        Ac[(size_t)row + (size_t)col * lda] = v;
      }
  };
  using one_c = std::integral_constant<int, 1>;
  using two_c = std::integral_constant<int, 2>;
  if constexpr (DIM == 3) {
    const int q2n = q * q;
    const int n2 = bp.nA[1] * bp.nB[1];
    const int nIA = bp.nA[0] * bp.nA[1];
    for (int i3 = 0; i3 < bp.nA[2]; ++i3)
      for (int j3 = 0; j3 < bp.nB[2]; ++j3) {
        if ((bp.diag && j3 > i3) || (MODE == 3 && j3 != i3)) continue;
        for (int e = t; e < np * q2n; e += NT) {
          int p = e / q2n, r = e - p * q2n;
          const double* fa = fam(pr[p].fa[2], i3);
          const double* fb = fam(pr[p].fb[2], j3);
          const double* m = Ms + p * QG + r;
          double acc = 0.0;
#pragma unroll
          for (int q3 = 0; q3 < q; ++q3) acc = fma(fa[q3] * fb[q3], m[q3 * q2n], acc);
          T1[e] = acc;
        }
        __syncthreads();
        // (the diagonal pass needs only the entries with j2 = i2)
        const int n2e = MODE == 3 ? bp.nA[1] : n2;
        for (int e2 = t; e2 < ncls * n2e * q; e2 += NT) {
          int cl = e2 / (n2e * q), r = e2 - cl * n2e * q;
          int ij = r / q, q1 = r - ij * q;
          if (MODE == 3) ij *= bp.nA[1] + 1;  // (i2, i2)
          const int e = (cl * n2 + ij) * q + q1;
          int i2 = ij % bp.nA[1], j2 = ij / bp.nA[1];
          double acc = 0.0;
          for (int p = 0; p < np; ++p) {
            if (pr[p].cls != cl) continue;
            const double* fa = fam(pr[p].fa[1], i2);
            const double* fb = fam(pr[p].fb[1], j2);
            const double* tp = T1 + p * q2n + q1;
#pragma unroll
            for (int q2 = 0; q2 < q; ++q2) acc = fma(fa[q2] * fb[q2], tp[q2 * q], acc);
          }
          T2[e] = acc;
        }
        __syncthreads();
        const int hA = (bp.nA[0] + 1) >> 1, hB = (bp.nB[0] + 1) >> 1;
        if constexpr (MODE == 3) {
          for (int e = t; e < nIA; e += NT) {  // il = jl
            const int i1 = e % bp.nA[0], i2 = e / bp.nA[0];
            final_block(one_c{}, one_c{}, i1, i1, bp.offA + bp.nA[0] * i2 + nIA * i3,
                        bp.offB + bp.sB[1] * i2 + bp.sB[2] * j3, T2 + (i2 + bp.nA[1] * i2) * q, n2 * q);
          }
        } else if constexpr (MODE == 0) {
          for (int e = t; e < hA * bp.nA[1] * hB * bp.nB[1]; e += NT) {
            int r = e;
            const int a = r % hA;
            r /= hA;
            const int i2 = r % bp.nA[1];
            r /= bp.nA[1];
            const int b = r % hB, j2 = r / hB;
            final_block(two_c{}, two_c{}, 2 * a, 2 * b, bp.offA + bp.nA[0] * i2 + nIA * i3,
                        bp.offB + bp.sB[1] * j2 + bp.sB[2] * j3, T2 + (i2 + bp.nA[1] * j2) * q, n2 * q);
          }
        } else {
          for (int e = t; e < nIA * bp.nB[0] * bp.nB[1]; e += NT) {
            int r = e;
            const int i1 = r % bp.nA[0];
            r /= bp.nA[0];
            const int i2 = r % bp.nA[1];
            r /= bp.nA[1];
            const int j1 = r % bp.nB[0], j2 = r / bp.nB[0];
            final_block(one_c{}, one_c{}, i1, j1, bp.offA + bp.nA[0] * i2 + nIA * i3,
                        bp.offB + bp.sB[1] * j2 + bp.sB[2] * j3, T2 + (i2 + bp.nA[1] * j2) * q, n2 * q);
          }
        }
        __syncthreads();
      }
  } else {
    for (int i2 = 0; i2 < bp.nA[1]; ++i2)
      for (int j2 = 0; j2 < bp.nB[1]; ++j2) {
        if ((bp.diag && j2 > i2) || (MODE == 3 && j2 != i2)) continue;
        for (int e = t; e < ncls * q; e += NT) {
          int cl = e / q, q1 = e - cl * q;
          double acc = 0.0;
          for (int p = 0; p < np; ++p) {
            if (pr[p].cls != cl) continue;
            const double* fa = fam(pr[p].fa[1], i2);
            const double* fb = fam(pr[p].fb[1], j2);
            const double* m = Ms + p * QG + q1;
#pragma unroll
            for (int q2 = 0; q2 < q; ++q2) acc = fma(fa[q2] * fb[q2], m[q2 * q], acc);
          }
          T1[e] = acc;
        }
        __syncthreads();
        const int hA = (bp.nA[0] + 1) >> 1, hB = (bp.nB[0] + 1) >> 1;
        if constexpr (MODE == 3) {
          for (int e = t; e < bp.nA[0]; e += NT)
            final_block(one_c{}, one_c{}, e, e, bp.offA + bp.nA[0] * i2, bp.offB + bp.sB[1] * j2, T1, q);
        } else if constexpr (MODE == 0) {
          for (int e = t; e < hA * hB; e += NT) {
            const int b = e / hA, a = e - b * hA;
            final_block(two_c{}, two_c{}, 2 * a, 2 * b, bp.offA + bp.nA[0] * i2, bp.offB + bp.sB[1] * j2, T1, q);
          }
        } else {
          for (int e = t; e < bp.nA[0] * bp.nB[0]; e += NT) {
            const int j1 = e / bp.nA[0], i1 = e - j1 * bp.nA[0];
            final_block(one_c{}, one_c{}, i1, j1, bp.offA + bp.nA[0] * i2, bp.offB + bp.sB[1] * j2, T1, q);
          }
        }
        __syncthreads();
      }
  }
}

template <int DIM, int QN>
__global__ __launch_bounds__(256) void gram_sf_kernel(const SfBlockPair* __restrict__ bps,
                                                      const SfPair* __restrict__ pairs,
                                                      const double* __restrict__ F, int nF, int q_rt,
                                                      const double* __restrict__ Wip, int n_lt, int na, int QG,
                                                      double* __restrict__ A, int64_t lda, int64_t sA,
                                                      const double* __restrict__ s, int64_t ss) {
  sf_body<DIM, QN, 0>(bps, pairs, F, nF, 3, q_rt, Wip, n_lt, na, QG, A, lda, sA, s, ss);
}
template <int DIM, int QN>
__global__ __launch_bounds__(256) void gram_sf_diag_kernel(const SfBlockPair* __restrict__ bps,
                                                           const SfPair* __restrict__ pairs,
                                                           const double* __restrict__ F, int nF, int q_rt,
                                                           const double* __restrict__ Wip, int n_lt, int na, int QG,
                                                           double* __restrict__ A, int64_t lda, int64_t sA) {
  sf_body<DIM, QN, 3>(bps, pairs, F, nF, 3, q_rt, Wip, n_lt, na, QG, A, lda, sA, nullptr, 0);
}
template <int DIM, int QN>
__global__ __launch_bounds__(256) void bmat_side_sf_kernel(const SfBlockPair* __restrict__ bps,
                                                           const SfPair* __restrict__ pairs,
                                                           const double* __restrict__ F, int nF, int nfam, int q_rt,
                                                           const double* __restrict__ Wside, int npts, int na,
                                                           double* __restrict__ A, int64_t lda, int64_t sA,
                                                           const double* __restrict__ s, int64_t ss) {
  const int q = QN > 0 ? QN : q_rt;
  sf_body<DIM, QN, 2>(bps, pairs, F, nF, nfam, q_rt, Wside, npts, na, DIM == 3 ? q * q * q : q * q, A, lda, sA, s,
                      ss);
}
template <int DIM, int QN>
__global__ __launch_bounds__(256) void bmat_sf_kernel(const SfBlockPair* __restrict__ bps,
                                                      const SfPair* __restrict__ pairs,
                                                      const double* __restrict__ F, int nF, int q_rt,
                                                      const double* __restrict__ Wvol, int na, int Q,
                                                      double* __restrict__ A, int64_t lda, int64_t sA,
                                                      const double* __restrict__ s, int64_t ss) {
  sf_body<DIM, QN, 1>(bps, pairs, F, nF, 4, q_rt, Wvol, 1, na, Q, A, lda, sA, s, ss);
}

// --------------------------------------------------- mixed-precision GEMM
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int MKC = 32;  // k per chunk

template <int MODE>
struct MixedOps;

template <>
struct MixedOps<DPG_PREC_FP32> {
  static constexpr int kSplit = 0;
};
template <>
struct MixedOps<DPG_PREC_BF16X3> {
  static constexpr int kSplit = 1;
  typedef __bf16 T;
  typedef bf16x8 V;
  static __device__ inline f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <>
struct MixedOps<DPG_PREC_FP16X3> {
  static constexpr int kSplit = 1;
  typedef _Float16 T;
  typedef f16x8 V;
  static __device__ inline f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

template <int MODE>
__device__ inline void gemm_mixed_body(const GemmArgs& g) {
  int b = xcd_remap(blockIdx.x, gridDim.x);
  int cell = b / g.ntiles, tile = b % g.ntiles;
  if (cell >= g.ncells) return;
  int ti, tj;
  tile_coords(tile, g.lower, g.tiles_m, ti, tj);
  const int i0 = ti * GT, j0 = tj * GT;
  const double* __restrict__ X = g.X + (size_t)cell * g.sx;
  const double* __restrict__ Y = g.Y + (size_t)cell * g.sy;
  double* __restrict__ C = g.C + (size_t)cell * g.sc;
  const double* px = g.px ? g.px + (size_t)cell * g.ss : nullptr;
  const double* py = g.py ? g.py + (size_t)cell * g.ss : nullptr;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  // per-thread column scale factors of the 8 elements it stages (its columns are fixed)
  double sxv[8], syv[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    int ii = (t + 256 * r) & 63;
    sxv[r] = (px && i0 + ii < g.m) ? px[i0 + ii] : 1.0;
    syv[r] = (py && j0 + ii < g.n) ? py[j0 + ii] : 1.0;
  }
  double xr[8], yr[8];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int e = t + 256 * r, kk = e >> 6, ii = e & 63, kg = k0 + kk;
      xr[r] = (kg < g.k && i0 + ii < g.m) ? X[(size_t)kg * g.ldx + i0 + ii] * sxv[r] : 0.0;
      yr[r] = (kg < g.k && j0 + ii < g.n) ? Y[(size_t)kg * g.ldy + j0 + ii] * syv[r] : 0.0;
    }
  };
  f4 acc[2][2], accx[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb) {
      acc[a][bb] = f4{0.f, 0.f, 0.f, 0.f};
      accx[a][bb] = f4{0.f, 0.f, 0.f, 0.f};
    }
  if constexpr (MixedOps<MODE>::kSplit == 0) {
    __shared__ float Xs[MKC][GT + 4];
    __shared__ float Ys[MKC][GT + 4];
    fetch(0);
    for (int k0 = 0; k0 < g.k; k0 += MKC) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        int e = t + 256 * r;
        Xs[e >> 6][e & 63] = (float)xr[r];
        Ys[e >> 6][e & 63] = (float)yr[r];
      }
      __syncthreads();
      if (k0 + MKC < g.k) fetch(k0 + MKC);
#pragma unroll
      for (int k4 = 0; k4 < MKC; k4 += 4) {
        const int kr = k4 + (lane >> 4), cc = lane & 15;
        float xa[2], yb[2];
        xa[0] = Xs[kr][wr * 32 + cc];
        xa[1] = Xs[kr][wr * 32 + 16 + cc];
        yb[0] = Ys[kr][wc * 32 + cc];
        yb[1] = Ys[kr][wc * 32 + 16 + cc];
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int bi = 0; bi < 2; ++bi)
            acc[bj][bi] = __builtin_amdgcn_mfma_f32_16x16x4f32(yb[bj], xa[bi], acc[bj][bi], 0, 0, 0);
      }
    }
  } else {
    typedef typename MixedOps<MODE>::T T;
    typedef typename MixedOps<MODE>::V V;
    // row-major tiles (row, k) so that a lane reads its 8 consecutive k in one access
    __shared__ __attribute__((aligned(16))) T Xh[GT][MKC + 8];
    __shared__ __attribute__((aligned(16))) T Xl[GT][MKC + 8];
    __shared__ __attribute__((aligned(16))) T Yh[GT][MKC + 8];
    __shared__ __attribute__((aligned(16))) T Yl[GT][MKC + 8];
    fetch(0);
    for (int k0 = 0; k0 < g.k; k0 += MKC) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        int e = t + 256 * r, kk = e >> 6, ii = e & 63;
        T xh = (T)xr[r], yh = (T)yr[r];
        Xh[ii][kk] = xh;
        Xl[ii][kk] = (T)(xr[r] - (double)xh);
        Yh[ii][kk] = yh;
        Yl[ii][kk] = (T)(yr[r] - (double)yh);
      }
      __syncthreads();
      if (k0 + MKC < g.k) fetch(k0 + MKC);
      const int kq = 8 * (lane >> 4), cc = lane & 15;
      V xah[2], xal[2], ybh[2], ybl[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        xah[q] = *reinterpret_cast<const V*>(&Xh[wr * 32 + q * 16 + cc][kq]);
        xal[q] = *reinterpret_cast<const V*>(&Xl[wr * 32 + q * 16 + cc][kq]);
        ybh[q] = *reinterpret_cast<const V*>(&Yh[wc * 32 + q * 16 + cc][kq]);
        ybl[q] = *reinterpret_cast<const V*>(&Yl[wc * 32 + q * 16 + cc][kq]);
      }
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          acc[bj][bi] = MixedOps<MODE>::mma(ybh[bj], xah[bi], acc[bj][bi]);
          accx[bj][bi] = MixedOps<MODE>::mma(ybh[bj], xal[bi], accx[bj][bi]);
          accx[bj][bi] = MixedOps<MODE>::mma(ybl[bj], xah[bi], accx[bj][bi]);
        }
    }
  }
  // f32 16x16 result layout: D[a][b] in lane b + 16*(a/4), element a%4
#pragma unroll
  for (int bj = 0; bj < 2; ++bj)
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int i = i0 + wr * 32 + bi * 16 + (lane & 15);
        int j = j0 + wc * 32 + bj * 16 + 4 * (lane >> 4) + v;
        if (i < g.m && j < g.n && (!g.lower || i >= j)) {
          double* cp = C + (size_t)i + (size_t)j * g.ldc;
          double val = g.alpha * ((double)acc[bj][bi][v] + (double)accx[bj][bi][v]);
          if (g.srow) val *= g.srow[(size_t)cell * g.ss + i];
          if (g.scol) val *= g.scol[(size_t)cell * g.ss + j];
          if (g.beta != 0.0) val = fma(g.beta, *cp, val);
          *cp = val;
        }
      }
}

__global__ __launch_bounds__(256) void gemm_gram_f32(GemmArgs g) { gemm_mixed_body<DPG_PREC_FP32>(g); }
__global__ __launch_bounds__(256) void gemm_bmat_f32(GemmArgs g) { gemm_mixed_body<DPG_PREC_FP32>(g); }
__global__ __launch_bounds__(256) void gemm_gram_bf16x3(GemmArgs g) { gemm_mixed_body<DPG_PREC_BF16X3>(g); }
__global__ __launch_bounds__(256) void gemm_bmat_bf16x3(GemmArgs g) { gemm_mixed_body<DPG_PREC_BF16X3>(g); }
__global__ __launch_bounds__(256) void gemm_gram_f16x3(GemmArgs g) { gemm_mixed_body<DPG_PREC_FP16X3>(g); }
__global__ __launch_bounds__(256) void gemm_bmat_f16x3(GemmArgs g) { gemm_mixed_body<DPG_PREC_FP16X3>(g); }

// ------------------------------------------- int8 digit split of the Gram
// Producer: two lanes per test dof j (rows N..NR-1 write zero digits), lane h
// of the pair taking points 16h..16h+15 of every 32-point k-chunk, so that a
// wave stores whole 32-byte rows (1 KB contiguous per plane). Pass 1 sums the
// squares of the operand column (the DPOEQU factor s_j; the pair's halves are
// added with one shuffle); pass 2 recomputes the values, equilibrates them
// (|x| <= 1), splits each into kI8Digits signed digits by exact steps (64x,
// then rint and 128 * remainder) and stores 16 bytes per plane.
template <int NPR>
__global__ __launch_bounds__(256) void ip_digits_kernel(const double* __restrict__ Wip, int n_lt, int nvar,
                                                        const double* __restrict__ R,
                                                        const int32_t* __restrict__ vod, int N, int QG, int QGp,
                                                        int NR, int8_t* __restrict__ dig, int64_t sd,
                                                        double* __restrict__ s, int64_t ss) {
  constexpr int S = kI8Digits, QB = 16, QC = 32;
  extern __shared__ double Wl[];  // QC points x (n_lt, nvar, NPR)
  const int wslice = n_lt * nvar * NPR;
  const int c = blockIdx.y, t = threadIdx.x, h = t & 1;
  const int j = blockIdx.x * (blockDim.x / 2) + (t >> 1);
  const bool active = j < N;
  const int var = active ? vod[j] : 0;
  double diag = 0.0;
  for (int q0 = 0; q0 < QG; q0 += QC) {
    const int nq = min(QC, QG - q0);
    __syncthreads();
    const double* Wp = Wip + ((size_t)c * QG + q0) * wslice;
    for (int e = t; e < nq * wslice; e += blockDim.x) Wl[e] = Wp[e];
    __syncthreads();
    if (!active) continue;
    for (int e = QB * h; e < min(nq, QB * h + QB); ++e) {
      double r[NPR];
#pragma unroll
      for (int k = 0; k < NPR; ++k) r[k] = R[((size_t)(q0 + e) * NPR + k) * N + j];
      for (int lt = 0; lt < n_lt; ++lt) {
        const double* w = Wl + ((e * n_lt + lt) * nvar + var) * NPR;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < NPR; ++k) acc = fma(w[k], r[k], acc);
        diag = fma(acc, acc, diag);
      }
    }
  }
  diag += __shfl_xor(diag, 1);  // the pair's two halves (every lane takes part)
  const double sj = 1.0 / sqrt(diag);
  if (active && h == 0) s[(size_t)c * ss + j] = sj;
  const double sx = (active && diag > 0.0) ? sj : 0.0;  // an all-zero column stays zero
  const int nch = n_lt * QGp / QC;
  int8_t* dc = dig + (size_t)c * sd + (size_t)j * 32 + 16 * h;
  for (int q0 = 0; q0 < QGp; q0 += QC) {
    const int nq = max(0, min(QC, QG - q0));
    __syncthreads();
    const double* Wp = Wip + ((size_t)c * QG + q0) * wslice;
    for (int e = t; e < nq * wslice; e += blockDim.x) Wl[e] = Wp[e];
    __syncthreads();
    if (j >= NR) continue;
    const int nh = max(0, min(QB, nq - QB * h));  // this lane's valid points
    double r[QB][NPR];
#pragma unroll
    for (int e = 0; e < QB; ++e)
#pragma unroll
      for (int k = 0; k < NPR; ++k)
        r[e][k] = (active && e < nh) ? R[((size_t)(q0 + QB * h + e) * NPR + k) * N + j] : 0.0;
    for (int lt = 0; lt < n_lt; ++lt) {
      uint32_t pk[S][4];
#pragma unroll
      for (int p = 0; p < S; ++p)
#pragma unroll
        for (int u = 0; u < 4; ++u) pk[p][u] = 0u;
#pragma unroll
      for (int e = 0; e < QB; ++e) {
        double v = 0.0;
        if (e < nh) {  // points past QG are padding (and their LDS slots are not loaded)
          const double* w = Wl + (((QB * h + e) * n_lt + lt) * nvar + var) * NPR;
#pragma unroll
          for (int k = 0; k < NPR; ++k) v = fma(w[k], r[e][k], v);
        }
        double rr = v * sx * 64.0;
        // rint by the 1.5 * 2^52 shift (|rr| <= 64): the FP64 add rounds to nearest
        // even like rint, leaves the integer in the low bits of the sum (two's
        // complement) and needs no conversion instruction
        constexpr double kShift = 6755399441055744.0;
#pragma unroll
        for (int p = 0; p < S; ++p) {
          const double tt = rr + kShift;
          const double d = tt - kShift;
          rr = (rr - d) * 128.0;
          const uint32_t lo = (uint32_t)__double2loint(tt);
          pk[p][e >> 2] |= (lo & 0xFFu) << (8 * (e & 3));
        }
      }
      const int chunk = (lt * QGp + q0) / QC;
#pragma unroll
      for (int p = 0; p < S; ++p)
        *reinterpret_cast<uint4*>(dc + (size_t)(p * nch + chunk) * NR * 32) =
            make_uint4(pk[p][0], pk[p][1], pk[p][2], pk[p][3]);
    }
  }
}

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// 64x64 lower tiles, 4 waves of 32x32 sub-tiles. Per 32-k chunk the digit
// planes of the tile's rows (and, off the diagonal, of its columns) are staged
// through LDS ([plane][row][32 bytes], rows padded to 48 bytes); lane l reads
// row l%32, bytes 16*(l/32) .. +16 of a plane as its MFMA operand.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void gemm_gram_i8(I8GramArgs a) {
  constexpr int S = kI8Digits, RL = 48, PER = S * 128 / 256;  // 16-byte pieces per thread and operand
  int b = xcd_remap(blockIdx.x, gridDim.x);
  const int cell = b / a.ntiles, tile = b % a.ntiles;
  if (cell >= a.ncells) return;
  int ti, tj;
  tile_coords(tile, 1, a.tiles_m, ti, tj);
  const int i0 = ti * 64, j0 = tj * 64;
  const bool dtile = ti == tj;
  __shared__ __attribute__((aligned(16))) int8_t Xs[S][64][RL];
  __shared__ __attribute__((aligned(16))) int8_t Ys[S][64][RL];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  const int8_t* __restrict__ D = a.dig + (size_t)cell * a.sd;
  const size_t plane = (size_t)a.nch * a.NR * 32;
  i32x16 acc[S];
#pragma unroll
  for (int l = 0; l < S; ++l)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[l][v] = 0;
  // piece e = t + 256 r: plane e/128, row (e/2)%64, half e%2
  i32x4 xr[PER], yr[PER];
  auto fetch = [&](int c) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int e = t + 256 * r, p = e >> 7, row = (e >> 1) & 63, h = e & 1;
      const int8_t* src = D + p * plane + ((size_t)c * a.NR + row) * 32 + 16 * h;
      xr[r] = *reinterpret_cast<const i32x4*>(src + (size_t)i0 * 32);
      yr[r] = *reinterpret_cast<const i32x4*>(src + (size_t)(dtile ? i0 : j0) * 32);  // a diagonal tile reads its rows twice (L1 hits)
    }
  };
  auto stage = [&]() {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int e = t + 256 * r, p = e >> 7, row = (e >> 1) & 63, h = e & 1;
      *reinterpret_cast<i32x4*>(&Xs[p][row][16 * h]) = xr[r];
      *reinterpret_cast<i32x4*>(&Ys[p][row][16 * h]) = yr[r];
    }
    __syncthreads();
  };
  const int fr = lane & 31, fk = 16 * (lane >> 5);
  auto multiply = [&]() {
    i32x4 bf[S];
#pragma unroll
    for (int q = 0; q < S; ++q) bf[q] = *reinterpret_cast<const i32x4*>(&Ys[q][wc * 32 + fr][fk]);
#pragma unroll
    for (int p = 0; p < S; ++p) {
      const i32x4 af = *reinterpret_cast<const i32x4*>(&Xs[p][wr * 32 + fr][fk]);
#pragma unroll
      for (int q = 0; q < S - p; ++q) acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af, bf[q], acc[p + q], 0, 0, 0);
    }
  };
  // k-chunks of segments that are zero for this tile are skipped (kseg/32 chunks per segment)
  const uint64_t km = a.kmask ? a.kmask[tile] : ~0ull;
  const int cps = a.kmask ? (int)(a.kseg / 32) : a.nch;
  int seg = 0, segend = cps;
  auto skip_zero_segments = [&](int c) {
    while (c < a.nch && !((km >> seg) & 1ull)) {
      c = segend;
      ++seg;
      segend += cps;
    }
    return c;
  };
  auto next_chunk = [&](int c) {
    if (c >= a.nch) return c;
    ++c;
    if (c >= segend) {
      ++seg;
      segend += cps;
      c = skip_zero_segments(c);
    }
    return c;
  };
  const int si0 = i0 + wr * 32, sj0 = j0 + wc * 32;
  const bool busy = si0 < a.N && sj0 < a.N && si0 + 31 >= sj0;
  int kc = skip_zero_segments(0);
  if (kc < a.nch) fetch(kc);
  while (kc < a.nch) {
    stage();
    const int kn = next_chunk(kc);
    if (kn < a.nch) fetch(kn);
    if (busy) multiply();
    kc = kn;
  }
  // 32x32 i32 result layout: D[a][b] (a the first operand's row, b the second's)
  // in lane b + 32*((a/4)%2), element 4*(a/8) + a%4
  double* __restrict__ C = a.C + (size_t)cell * a.sc;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int i = si0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    const int j = sj0 + (lane & 31);
    double val = 0.0;
#pragma unroll
    for (int l = S - 1; l >= 0; --l) val += ldexp((double)acc[l][v], 2 - 7 * (l + 2));  // exact scaling
    if (i < a.N && j < a.N && i >= j) C[(size_t)i + (size_t)j * a.ldc] = val;
  }
}

// ------------------------------------------------------------ panel kernels
// Columns [k0, k0+kb) (kb <= 64) of each cell's lower-stored matrix:
//   potrf_wave   one wavefront per cell: POTRF of the diagonal block in
//                registers; the factor to Ld (row-major, for the back solve)
//                and to Lt (column-major, with the reciprocals of its diagonal)
//   panel_trsm   one thread per row below the block: X <- X L^-T by
//                substitution, column by column, L staged in LDS per cell.
// The diagonal block of A itself is left untouched (no later stage reads it).
// The solve is a substitution, not a product with an explicit inverse of L: on
// cells whose equilibrated Gram matrix has condition 1e12 the explicit inverse
// makes K up to 20x less accurate than the substitution does.
constexpr int PNB = 64;


// Diagonal block (kb <= NB <= 64 columns of a 64-column panel) of one cell per
// wavefront (4 cells per workgroup), in registers: lane i holds row i of the
// block. Right-looking Cholesky with the pivot and the scaled column broadcast
// across lanes (no LDS, no barriers). Writes the factor by columns (coalesced),
// Ld[(k0+c)*64 + r] = L[k0+r][k0+c] for r, c < kb (r < c: zero), and, for the
// panel solve, Lt[j][m] = L[m][j] (m > j), Lt[j][j] = 1 / L[j][j],
// zero outside the block. NB is the smallest of 8, 16, 32, 64 that holds kb,
// so short last panels cost little.
// A non-positive pivot sets info to its 1-based column (first failure only) and
// turns the rest of the cell's factor into NaN, as LAPACK leaves it unusable.
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void potrf_wave(
    const double* __restrict__ A, int64_t lda, int64_t sA, int k0, int kb, double* __restrict__ Ld, int64_t sLd,
    double* __restrict__ Lt, int32_t* __restrict__ info, int info_base, int ncells) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncells) return;  // whole wavefronts only: no barrier below
  const double* Ar = A + (size_t)c * sA + (size_t)(k0 + lane) + (size_t)k0 * lda;  // row `lane`, column k0
  double a[NB];
  int ll, lf, ls;  // see below
  asm volatile("v_mov_b32 %0, %1" : "=v"(ll) : "v"(lane));
#pragma unroll
  for (int cc = 0; cc < NB; ++cc) {
    double v = (ll == cc) ? 1.0 : 0.0;
    if (ll < kb && cc < kb && cc <= ll) v = *Ar;
    a[cc] = v;
    Ar += lda;
  }
  // per-phase opaque copies of the lane index (ll, lf, ls): keep the compiler
  // from sharing the per-column lane masks across phases, which would hold them
  // all live in scalar registers
  asm volatile("v_mov_b32 %0, %1" : "=v"(lf) : "v"(lane));
  int fail = -1;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    __builtin_amdgcn_sched_barrier(0);
    double d = wave_read(a[j], j);
    const bool bad = !(d > 0.0);
    fail = (fail < 0 && bad) ? j : fail;
    d = bad ? __builtin_nan("") : d;
    const double sd = sqrt(d), rs = 1.0 / sd;
    const double lij = lf > j ? a[j] * rs : 0.0;
    a[j] = lf == j ? sd : (lf > j ? lij : a[j]);
#pragma unroll
    for (int k = j + 1; k < NB; ++k) {
      a[k] = fma(-lij, wave_read(a[j], k), a[k]);
      // materialise the update now: left to itself the compiler sinks it to a[k]'s
      // next use and keeps every broadcast value live in scalar registers meanwhile
      asm volatile("" : "+v"(a[k]));
    }
  }
  double* Lc = Ld + (size_t)c * sLd;
  asm volatile("v_mov_b32 %0, %1" : "=v"(ls) : "v"(lane));
  if (ls < kb) {
#pragma unroll
    for (int cc = 0; cc < NB; ++cc)  // cc is a constant after unrolling: a[] stays in registers
      if (cc < kb) Lc[(size_t)(k0 + cc) * PNB + ls] = cc <= ls ? a[cc] : 0.0;  // column cc: lanes contiguous
  }
  if (lane == 0 && fail >= 0 && fail < kb && info) atomicCAS(&info[c], 0, info_base + k0 + fail + 1);
  // Lt, column j of L contiguous (lane = row): coalesced stores; the diagonal
  // entry, 1 / L[lane][lane], is written last
  double* Ltc = Lt + (size_t)c * PNB * PNB + lane;
  int lt;
  asm volatile("v_mov_b32 %0, %1" : "=v"(lt) : "v"(lane));
  const bool row_in = lt < kb;
  double dl = 1.0;
#pragma unroll
  for (int jj = 0; jj < PNB; ++jj) {
    __builtin_amdgcn_sched_barrier(0);
    double v = 0.0;
    if constexpr (NB == PNB) {
      v = (row_in && jj < kb && lt > jj) ? a[jj] : 0.0;
      dl = lt == jj ? a[jj] : dl;
    } else if (jj < NB) {
      v = (row_in && jj < kb && lt > jj) ? a[jj < NB ? jj : 0] : 0.0;
      dl = lt == jj ? a[jj < NB ? jj : 0] : dl;
    }
    Ltc[(size_t)jj * PNB] = v;
  }
  if (row_in) Ltc[(size_t)lt * PNB] = 1.0 / dl;
}

// one column of the substitution: y_J *= 1/L_JJ, y_m -= L_mJ y_J (m > J)
template <int NB, int J>
__device__ __forceinline__ void trsm_column(double (&y)[NB], const double* L) {
  y[J] *= L[J * PNB + J];
#pragma unroll
  for (int m = J + 1; m < NB; ++m) y[m] = fma(-L[J * PNB + m], y[J], y[m]);
}
// columns J, J+1, ... while below kb
template <int NB, int J>
__device__ __forceinline__ void trsm_columns(double (&y)[NB], const double* L, int kb) {
  if constexpr (J < NB) {
    if (J < kb) {
      trsm_column<NB, J>(y, L);
      trsm_columns<NB, J + 1>(y, L, kb);
    }
  }
}

// X <- X L^-T for rows [k0+kb, D) of columns [k0, k0+kb), one thread per row:
// for j = 0..kb-1, y_j *= 1/L_jj, then y_m -= L_mj y_j (m > j) -- the column
// order of the reference BLAS DTRSM (side R, lower, transposed). The workgroup
// (one cell) stages the cell's L columns in LDS once; every thread then reads
// the same L_mj (LDS broadcast). Read through the scalar cache instead, the
// L blocks of the cells resident on a CU do not fit and every wave waits on L2.
template <int NB>
__global__ __launch_bounds__(256) void panel_trsm(double* __restrict__ A, int64_t lda, int64_t sA, int D, int k0,
                                                  int kb, const double* __restrict__ Lt) {
  __shared__ double Ls[NB * PNB];
  const int c = blockIdx.y;
  const double* __restrict__ Lc = Lt + (size_t)c * PNB * PNB;
  for (int e = threadIdx.x; e < NB * PNB; e += blockDim.x) Ls[e] = Lc[e];
  __syncthreads();
  const int r = k0 + kb + blockIdx.x * 256 + threadIdx.x;
  if (r >= D) return;  // no barrier below
  double* Ar = A + (size_t)c * sA + (size_t)r + (size_t)k0 * lda;
  // per-phase opaque copies of kb: otherwise the compiler keeps the 64 masks
  // j < kb of the loads live in scalar registers until the stores
  int kl, ks, kw;
  asm volatile("s_mov_b32 %0, %1" : "=s"(kl) : "s"(kb));
  double y[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) y[j] = j < kl ? Ar[(size_t)j * lda] : 0.0;
  asm volatile("s_mov_b32 %0, %1" : "=s"(ks) : "s"(kb));
  trsm_columns<NB, 0>(y, Ls, ks);
  asm volatile("s_mov_b32 %0, %1" : "=s"(kw) : "s"(kb));
#pragma unroll
  for (int j = 0; j < NB; ++j)
    if (j < kw) Ar[(size_t)j * lda] = y[j];
}

// ---------------------------------------------------------- small kernels
// zero the trailing block rows/cols [r0, D) (lower part and the diagonal)
__global__ void zero_trailing(double* __restrict__ A, int64_t lda, int64_t sA, int r0, int D) {
  int c = blockIdx.z, col = r0 + blockIdx.y;
  int n = D - r0;
  double* Ac = A + (size_t)c * sA + (size_t)col * lda;
  for (int r = col + blockIdx.x * blockDim.x + threadIdx.x; r < D; r += gridDim.x * blockDim.x) Ac[r] = 0.0;
  (void)n;
}

// POEQU: s_i = 1 / sqrt(A_ii), i < n
__global__ void diag_scale_factors(const double* __restrict__ A, int64_t lda, int64_t sA, int r0, int n,
                                   double* __restrict__ s, int64_t ss, int sign) {
  int c = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = sign * A[(size_t)c * sA + (size_t)(r0 + i) * (lda + 1)];
  s[(size_t)c * ss + i] = 1.0 / sqrt(d);
}

// A[r, j] *= sign * s'_r * s_j for the lower part of the square block [r0, D):
// s'_r = s[r - r0] if r - r0 < n else 1; only columns j - r0 < ncols are scaled by s_j.
// One wavefront per column (rows contiguous), 4 columns per workgroup.
__global__ __launch_bounds__(256) void apply_scaling(double* __restrict__ A, int64_t lda, int64_t sA, int r0, int D,
                                                     int n, int ncol, const double* __restrict__ s, int64_t ss,
                                                     double sign) {
  const int c = blockIdx.y, jc = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (jc >= ncol) return;
  const int col = r0 + jc;
  const double* sc = s + (size_t)c * ss;
  const double sj = jc < n ? sc[jc] : 1.0;
  double* Ac = A + (size_t)c * sA + (size_t)col * lda;
  for (int r = col + lane; r < D; r += 64) {
    const int ir = r - r0;
    const double si = ir < n ? sc[ir] : 1.0;
    Ac[r] = sign * Ac[r] * si * sj;
  }
}

// residual row: A[row_r, i] = A[row_l, i] - sum_j u[c][j] A[r0 + j, i], one wave per column i
__global__ __launch_bounds__(256) void residual_row(double* __restrict__ A, int64_t lda, int64_t sA, int N, int M,
                                                    int r0, int row_l, int row_r, const double* __restrict__ u) {
  int c = blockIdx.y;
  int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (i >= N) return;
  const double* col = A + (size_t)c * sA + (size_t)i * lda;
  double acc = 0.0;
  if (u) {
    const double* uc = u + (size_t)c * M;
    for (int j = lane; j < M; j += 64) acc = fma(uc[j], col[r0 + j], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  }
  // without u the row is cleared (it must not carry stale workspace values)
  if (lane == 0) A[(size_t)c * sA + row_r + (size_t)i * lda] = u ? col[row_l] - acc : 0.0;
}

// out[c][a][b] = sign * sym(A)[r0 + a][r0 + b] (row-major n x n), 32x32 tiles
__global__ __launch_bounds__(256) void extract_sym(const double* __restrict__ A, int64_t lda, int64_t sA, int r0,
                                                   int n, double sign, double* __restrict__ out) {
  __shared__ double tile[32][33];
  int c = blockIdx.z, ta = blockIdx.y, tb = blockIdx.x;
  const double* Ac = A + (size_t)c * sA;
  double* oc = out + (size_t)c * n * n;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  if (ta < tb) {
    // strictly upper tile: value(a,b) = A[r0+b, r0+a] with b > a: read along b, write along b
    for (int yy = ty; yy < 32; yy += 8) {
      int a = ta * 32 + yy, b = tb * 32 + tx;
      if (a < n && b < n) oc[(size_t)a * n + b] = sign * Ac[(size_t)(r0 + b) + (size_t)(r0 + a) * lda];
    }
    return;
  }
  for (int yy = ty; yy < 32; yy += 8) {
    int a = ta * 32 + tx, b = tb * 32 + yy;
    double v = 0.0;
    if (a < n && b < n)
      v = (a >= b) ? Ac[(size_t)(r0 + a) + (size_t)(r0 + b) * lda] : Ac[(size_t)(r0 + b) + (size_t)(r0 + a) * lda];
    tile[tx][yy] = v;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    int a = ta * 32 + yy, b = tb * 32 + tx;
    if (a < n && b < n) oc[(size_t)a * n + b] = sign * tile[yy][tx];
  }
}

// out[c][a] = sign * A[row, r0 + a]
__global__ void extract_row(const double* __restrict__ A, int64_t lda, int64_t sA, int row, int r0, int n,
                            double sign, double* __restrict__ out) {
  int c = blockIdx.y, a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  out[(size_t)c * n + a] = sign * A[(size_t)c * sA + row + (size_t)(r0 + a) * lda];
}

// ------------------------------------------------------------------ adjoint
// e = G^-1 r from the stored factor of S G S = Lt Lt^T: the residual row of A
// holds y = Lt^-1 S r after the partial Cholesky, so e = S Lt^-T y by one
// back substitution (diagonal blocks in Ld, the rest below them in A).
__global__ __launch_bounds__(256) void backsolve_kernel(const double* __restrict__ A, int64_t lda, int64_t sA,
                                                        int N, int row_r, const double* __restrict__ Ld,
                                                        int64_t sLd, const double* __restrict__ s, int64_t ss,
                                                        double* __restrict__ e, int64_t se) {
  extern __shared__ double z[];
  const int c = blockIdx.x, t = threadIdx.x;
  const double* Ac = A + (size_t)c * sA;
  const double* Lc = Ld + (size_t)c * sLd;
  for (int i = t; i < N; i += blockDim.x) z[i] = Ac[(size_t)row_r + (size_t)i * lda];
  for (int j = N - 1; j >= 0; --j) {
    __syncthreads();
    const int kj = j & ~63;
    const double zj = z[j] / Lc[(size_t)j * 64 + (j - kj)];
    __syncthreads();
    if (t == 0) z[j] = zj;
    // y_i -= L[j][i] z_j for i < j
    for (int i = t; i < j; i += blockDim.x) {
      double lji = (i >= kj) ? Lc[(size_t)i * 64 + (j - kj)] : Ac[(size_t)j + (size_t)i * lda];
      z[i] = fma(-lji, zj, z[i]);
    }
  }
  __syncthreads();
  for (int i = t; i < N; i += blockDim.x) e[(size_t)c * se + i] = s[(size_t)c * ss + i] * z[i];
}

// Field recovery after a condensed solve, one workgroup per cell: from K
// (row-major M x M, symmetric) and rhs (M) of the local assembly and the
// trace/flux part u_T of u (entries F..M-1), the lower triangle of S K_FF S
// (s_j = 1/sqrt(K_jj)) into A and below it the row A[F][j] = s_j b_j with
// b = rhs_F - K_FT u_T. K's symmetry lets every read run along rows of K.
__global__ __launch_bounds__(256) void recover_setup(const double* __restrict__ K, const double* __restrict__ rhs,
                                                     const double* __restrict__ u, int M, int F,
                                                     double* __restrict__ A, int64_t lda, int64_t sA,
                                                     double* __restrict__ s, int64_t ss) {
  const int c = blockIdx.x, t = threadIdx.x;
  const double* Kc = K + (size_t)c * M * M;
  const double* uc = u + (size_t)c * M;
  double* Ac = A + (size_t)c * sA;
  double* sc = s + (size_t)c * ss;
  for (int j = t; j < F; j += blockDim.x) sc[j] = 1.0 / sqrt(Kc[(size_t)j * M + j]);
  __syncthreads();
  for (int e = t; e < F * F; e += blockDim.x) {
    const int i = e % F, j = e / F;  // A[i][j], i >= j, read as K[j][i]
    if (i >= j) Ac[(size_t)i + (size_t)j * lda] = sc[i] * sc[j] * Kc[(size_t)j * M + i];
  }
  for (int j = t; j < F; j += blockDim.x) {
    double b = rhs[(size_t)c * M + j];
    for (int q = F; q < M; ++q) b = fma(-Kc[(size_t)q * M + j], uc[q], b);  // K[j][q] = K[q][j]
    Ac[(size_t)F + (size_t)j * lda] = sc[j] * b;
  }
}

__device__ inline double dcoef_value(const DSummand& sm, int k, const double* cc, const double* n) {
  if (sm.slot != k) return 0.0;
  double v = sm.scale;
  double a = cc[k];
  v *= (sm.fn == DPG_FN_SQRT) ? 0.5 / sqrt(a) : (sm.fn == DPG_FN_INV ? -1.0 / (a * a) : 1.0);
  if (sm.normal >= 0) v *= n ? n[sm.normal] : 0.0;
  return v;
}

// The same with the side normal held in registers (n0, n1, n2 selected, never
// addressed: a pointer to a local array would put it in scratch memory)
__device__ inline double normal_comp(bool hasn, double n0, double n1, double n2, int i) {
  return hasn ? (i == 0 ? n0 : (i == 1 ? n1 : n2)) : 0.0;
}
__device__ inline double dcoef_value_n(const DSummand& sm, int k, const double* cc, bool hasn, double n0, double n1,
                                       double n2) {
  if (sm.slot != k) return 0.0;
  double v = sm.scale;
  double a = cc[k];
  v *= (sm.fn == DPG_FN_SQRT) ? 0.5 / sqrt(a) : (sm.fn == DPG_FN_INV ? -1.0 / (a * a) : 1.0);
  if (sm.normal >= 0) v *= normal_comp(hasn, n0, n1, n2, sm.normal);
  return v;
}

// ------------------------------------------------- stored factor, adjoint
// The factored first N columns of a cell's augmented matrix as a row-major
// (N + M + 1) x N block F: rows 0..N-1 the lower factor Lt of S G S = Lt Lt^T
// (entries above the diagonal 0), rows N..N+M-1 the rows of (Lt^-1 S B)^T,
// row N+M (Lt^-1 S l)^T. Entries of Lt inside a diagonal 64-block come from
// Ld, the others from A. 32x32 tiles through LDS: A read along its columns,
// F written along its rows.
__global__ __launch_bounds__(256) void extract_factor(const double* __restrict__ A, int64_t lda, int64_t sA, int N,
                                                      int R, const double* __restrict__ Ld, int64_t sLd,
                                                      double* __restrict__ F, int64_t sF) {
  __shared__ double tile[32][33];
  const int c = blockIdx.z, ti = blockIdx.y, tj = blockIdx.x;  // row tile, column tile of F
  const double* Ac = A + (size_t)c * sA;
  const double* Lc = Ld + (size_t)c * sLd;
  double* Fc = F + (size_t)c * sF;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    const int i = ti * 32 + tx, j = tj * 32 + yy;
    double v = 0.0;
    if (i < R && j < N && i >= j) {
      const int kb = i & ~63;
      v = (i < N && j >= kb) ? Lc[(size_t)j * 64 + (i - kb)] : Ac[(size_t)i + (size_t)j * lda];
    }
    tile[tx][yy] = v;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int i = ti * 32 + yy, j = tj * 32 + tx;
    if (i < R && j < N) Fc[(size_t)i * N + j] = tile[yy][tx];
  }
}

// From a stored factor F: z = y_l - Y u = Lt^-1 S (l - B u) and eta2 = z^T z
// (= r^T G^-1 r). One workgroup per cell, lanes over the test dofs; eta2 by a
// fixed tree, so the same bits on every call. z and eta2 may be NULL.
__global__ __launch_bounds__(256) void factored_residual(const double* __restrict__ F, int64_t sF, int N, int M,
                                                         const double* __restrict__ u, double* __restrict__ z,
                                                         int64_t sz, double* __restrict__ eta2) {
  __shared__ double part[256];
  const int c = blockIdx.x, t = threadIdx.x;
  const double* Fc = F + (size_t)c * sF;
  const double* uc = u + (size_t)c * M;
  double acc = 0.0;
  for (int j = t; j < N; j += 256) {
    double v = Fc[(size_t)(N + M) * N + j];
    for (int m = 0; m < M; ++m) v = fma(-Fc[(size_t)(N + m) * N + j], uc[m], v);
    if (z) z[(size_t)c * sz + j] = v;
    acc = fma(v, v, acc);
  }
  part[t] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) part[t] += part[t + h];
    __syncthreads();
  }
  if (t == 0 && eta2) eta2[c] = part[0];
}

// X = S Lt^-T [Y | y_l | z] for the columns the adjoint needs (row-major
// N x ncol): columns 0..nT-1 from the rows N.. of F (Y = Lt^-1 S B, nT = M or
// 0), then y_l (row N+M of F) when with_y, then z when z is not NULL. The
// columns are G^-1 B, G^-1 l and G^-1 (l - B u). One workgroup per cell, a
// lane per column: substitution from the last row up with Lt's entries as
// wave-uniform loads, then the scaling by s.
__global__ __launch_bounds__(256) void factor_backsolve(const double* __restrict__ F, int64_t sF, int N, int M,
                                                        const double* __restrict__ s, int64_t ss, int nT, int with_y,
                                                        const double* __restrict__ z, int64_t sz,
                                                        double* __restrict__ X, int64_t sX, int ncol) {
  const int c = blockIdx.x;
  const double* Fc = F + (size_t)c * sF;
  const double* sc = s + (size_t)c * ss;
  double* Xc = X + (size_t)c * sX;
  for (int m = threadIdx.x; m < ncol; m += blockDim.x) {
    const double* b = m < nT ? Fc + (size_t)(N + m) * N
                             : ((with_y && m == nT) ? Fc + (size_t)(N + M) * N : z + (size_t)c * sz);
    for (int j = N - 1; j >= 0; --j) {
      double x = b[j];
      for (int i = j + 1; i < N; ++i) x = fma(-Fc[(size_t)i * N + j], Xc[(size_t)i * ncol + m], x);
      Xc[(size_t)j * ncol + m] = x / Fc[(size_t)j * N + j];
    }
    for (int j = 0; j < N; ++j) Xc[(size_t)j * ncol + m] *= sc[j];
  }
}

// Adjoint of the local assembly with respect to the per-cell coefficients:
//   grad[c][k] = d/dcoef_k ( <Kbar, K> + <rbar, rhs> + eta2(u) ),
// every part present only when its cotangent (Kbar, rbar) or u is given. With
// T = G^-1 B, y = G^-1 l, e = G^-1 (l - B u) (the columns of X) and
// Ks = sym(Kbar):
//   Bbar = 2 T Ks + y rbar^T - 2 e u^T,   Gbar = -T Ks T^T - sym(T rbar y^T) - e e^T,
//   d/dk = <Bbar, dB/dk> + <Gbar, dG/dk>   (l does not depend on the coefficients).
// B and G are sums over cubature points of w * coef * op(test) * trial and of
// w * lt(test) lt(test), so at a point only the test operators applied to the
// functions whose test-basis coefficients are the columns of X are needed,
// P[m] = op(sum_i X[i][m] psi_i)(x_q): the reference prims of those functions
// are contracted from the tables (recomputed, not stored), then transformed:
//   Gram point, IP term lt:      2 w dlt_k(X)^T Gp lt(X)
//   volume / side term (coef):   w dcoef_k op(X)^T Bp phi
// with Gp, Bp the cotangents in the coordinates of X's columns and phi the
// trial basis values at the point (flux values times the side parity). One
// workgroup per cell; its waves take the points in turn (npw of them, as many
// as their LDS buffers allow), lanes over the columns, each wave's points in a
// fixed order and one fixed-tree reduction per coefficient at the end: the
// same bits on every call. Column layout: T at [0, nT), y at cy (or -1), e at
// ce (or -1).
// LDS written by some lanes of a wave and read by others: ordered at wavefront
// scope (no workgroup barrier: the other waves work on other points)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int DIM>
__global__ __launch_bounds__(256) void adjoint_kernel(const DevProgram* __restrict__ prog, DevTables tab, int N,
                                                      int M, int QG, int Q, int Qs, int nsides,
                                                      const double* __restrict__ nodes,
                                                      const int8_t* __restrict__ parities,
                                                      const double* __restrict__ coefs,
                                                      const double* __restrict__ X, int64_t sX, int ncol, int nT,
                                                      int cy, int ce, const double* __restrict__ Kbar,
                                                      const double* __restrict__ rbar,
                                                      const double* __restrict__ u, double* __restrict__ grad,
                                                      int npw) {
  extern __shared__ double lds[];
  __shared__ double red[256];
  constexpr int NPR = DIM + 1;
  // the wave index as a scalar: the point loop, its tables and the program are
  // then uniform (scalar loads)
  const int c = blockIdx.x, t = threadIdx.x, w = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  const int V = prog->nverts, nvar = prog->nvar, ncoef = prog->ncoef;
  const double* Xn = nodes + (size_t)c * V * DIM;
  const double* cc = coefs + (size_t)c * ncoef;
  const double* Xc = X + (size_t)c * sX;
  const double* Kc = Kbar ? Kbar + (size_t)c * M * M : nullptr;
  const double* rc = rbar ? rbar + (size_t)c * M : nullptr;
  const double* uc = u ? u + (size_t)c * M : nullptr;
  // wave w's own buffers (the waves take the points in turn: npw of them)
  double* Z = lds + (size_t)w * ((size_t)nvar * NPR * ncol + ncol);  // Z[(v * NPR + r) * ncol + m]
  double* Pl = Z + (size_t)nvar * NPR * ncol;                        // P[m] of one IP term
  auto ks = [&](int i, int j) { return Kc[(size_t)i * M + j] + Kc[(size_t)j * M + i]; };
  double acc[DPG_MAX_COEFS];
  for (int k = 0; k < DPG_MAX_COEFS; ++k) acc[k] = 0.0;
  const int npts = QG + Q + nsides * Qs;
  for (int pp = w; w < npw && pp < npts; pp += npw) {
    const int kind = pp < QG ? 0 : (pp < QG + Q ? 1 : 2);
    const int p = kind == 0 ? pp : (kind == 1 ? pp - QG : pp - QG - Q);
    const double* dN = kind == 0 ? tab.dN_ip : (kind == 1 ? tab.dN_vol : tab.dN_side);
    const double* R = kind == 0 ? tab.R_ip : (kind == 1 ? tab.R_vol : tab.R_side);
    PointGeo<DIM> g;
    point_geometry<DIM>(Xn, dN + (size_t)p * V * DIM, V, g);
    double n[3] = {0.0, 0.0, 0.0};
    double wm;
    int side = 0;
    if (kind == 0) {
      wm = tab.w_ip[p] * g.detJ;
    } else if (kind == 1) {
      wm = tab.w_vol[p] * g.detJ;
    } else {
      side = p / Qs;
      double Tg[2][3];
      for (int a = 0; a < DIM - 1; ++a)
        for (int d = 0; d < DIM; ++d) {
          double s = 0.0;
          for (int e = 0; e < DIM; ++e) s += g.J[d][e] * tab.side_tan[((size_t)side * 2 + a) * DIM + e];
          Tg[a][d] = s;
        }
      if constexpr (DIM == 2) {
        n[0] = Tg[0][1];
        n[1] = -Tg[0][0];
      } else {
        n[0] = Tg[0][1] * Tg[1][2] - Tg[0][2] * Tg[1][1];
        n[1] = Tg[0][2] * Tg[1][0] - Tg[0][0] * Tg[1][2];
        n[2] = Tg[0][0] * Tg[1][1] - Tg[0][1] * Tg[1][0];
      }
      double len = 0.0;
      for (int d = 0; d < DIM; ++d) len += n[d] * n[d];
      len = sqrt(len);
      for (int d = 0; d < DIM; ++d) n[d] /= len;
      wm = tab.w_side[p] * len;
    }
    const bool hasn = kind == 2;
    // reference prims of the column functions at this point
    wave_lds_sync();
    for (int m = lane; m < ncol; m += 64)
      for (int v = 0; v < nvar; ++v) {
        double zr[NPR];
#pragma unroll
        for (int r = 0; r < NPR; ++r) zr[r] = 0.0;
        for (int i = prog->test_off[v]; i < prog->test_off[v + 1]; ++i) {
          const double xi = Xc[(size_t)i * ncol + m];
#pragma unroll
          for (int r = 0; r < NPR; ++r) zr[r] = fma(R[((size_t)p * NPR + r) * N + i], xi, zr[r]);
        }
#pragma unroll
        for (int r = 0; r < NPR; ++r) Z[((size_t)v * NPR + r) * ncol + m] = zr[r];
      }
    wave_lds_sync();
    double tr[NPR];
    auto op_col = [&](const DSummand& sm, int m) {
      op_to_prims<DIM>(sm.op, g, nullptr, tr);
      if (hasn && sm.op == DPG_OP_DOT_NORMAL) {  // (J n)^T / detJ with the normal from registers
        const double id = 1.0 / g.detJ;
#pragma unroll
        for (int e = 0; e < DIM; ++e) {
          double sn = n[0] * g.J[0][e];
#pragma unroll
          for (int d = 1; d < DIM; ++d) sn += n[d] * g.J[d][e];
          tr[e] = sn * id;
        }
      }
      double v = 0.0;
#pragma unroll
      for (int r = 0; r < NPR; ++r) v = fma(tr[r], Z[((size_t)sm.var * NPR + r) * ncol + m], v);
      return v;
    };
    if (kind == 0) {
      for (int lt = 0; lt < prog->n_lt; ++lt) {
        bool dep = false;
        for (int si = prog->lt_beg[lt]; si < prog->lt_beg[lt + 1]; ++si) dep |= prog->s[si].slot >= 0;
        if (!dep) continue;  // the term's weight has no coefficient: no contribution
        for (int m = lane; m < ncol; m += 64) {
          double v = 0.0;
          for (int si = prog->lt_beg[lt]; si < prog->lt_beg[lt + 1]; ++si)
            v = fma(coef_value(prog->s[si], cc, nullptr), op_col(prog->s[si], m), v);
          Pl[m] = v;
        }
        wave_lds_sync();
        for (int m = lane; m < ncol; m += 64) {
          // h = (Gbar in column coordinates) P
          double h = 0.0;
          if (m < nT) {
            if (Kc)
              for (int q = 0; q < M; ++q) h = fma(-0.5 * ks(m, q), Pl[q], h);
            if (rc) h = fma(-0.5 * rc[m], Pl[cy], h);
          } else if (m == cy) {
            for (int q = 0; q < M; ++q) h = fma(-0.5 * rc[q], Pl[q], h);
          } else if (m == ce) {
            h = -Pl[ce];
          }
          // d/dcoef_k of the term's value, every k in one pass over the summands
          // (each k's sum keeps the summand order)
          double dv[DPG_MAX_COEFS];
#pragma unroll
          for (int k = 0; k < DPG_MAX_COEFS; ++k) dv[k] = 0.0;
          for (int si = prog->lt_beg[lt]; si < prog->lt_beg[lt + 1]; ++si) {
            const DSummand& sm = prog->s[si];
            if (sm.slot < 0) continue;
            const double d = dcoef_value(sm, sm.slot, cc, nullptr), o = op_col(sm, m);
#pragma unroll
            for (int k = 0; k < DPG_MAX_COEFS; ++k)
              if (k == sm.slot) dv[k] = fma(d, o, dv[k]);
          }
#pragma unroll
          for (int k = 0; k < DPG_MAX_COEFS; ++k)
            if (k < ncoef) acc[k] = fma(2.0 * wm * dv[k], h, acc[k]);
        }
        wave_lds_sync();
      }
    } else {
      // trial blocks at this point: volume field blocks, or the side's trace/flux blocks
      const int nblk = kind == 1 ? prog->n_vb : prog->n_sb;
      for (int gi = 0; gi < nblk; ++gi) {
        const int sb = kind == 1 ? prog->vb_beg[gi] : prog->sb_beg[gi];
        const int se = kind == 1 ? prog->vb_beg[gi + 1] : prog->sb_beg[gi + 1];
        bool dep = false;
        for (int si = sb; si < se; ++si) dep |= prog->s[si].slot >= 0;
        if (!dep) continue;
        int off, card;
        const double* phi;
        double par = 1.0;
        if (kind == 1) {
          off = prog->vb_off[gi];
          card = prog->nf;
          phi = tab.field_vol + (size_t)p * prog->nf;
        } else {
          card = prog->sb_card[gi];
          off = prog->sb_off[gi] + side * card;
          phi = (prog->sb_flux[gi] ? tab.flux : tab.trace) + (size_t)(p % Qs) * card;
          if (prog->sb_flux[gi]) par = (double)parities[(size_t)c * nsides + side];
        }
        for (int m = lane; m < ncol; m += 64) {
          // gb = (Bbar in column coordinates) phi
          double gb = 0.0;
          if (m < nT) {
            if (Kc)
              for (int j = 0; j < card; ++j)
                gb = fma(ks(m, off + j), phi[j], gb);
          } else if (m == cy) {
            for (int j = 0; j < card; ++j) gb = fma(rc[off + j], phi[j], gb);
          } else if (m == ce) {
            for (int j = 0; j < card; ++j) gb = fma(-2.0 * uc[off + j], phi[j], gb);
          }
          gb *= par * wm;
          for (int si = sb; si < se; ++si) {
            const DSummand& sm = prog->s[si];
            if (sm.slot < 0) continue;
            const double d = dcoef_value_n(sm, sm.slot, cc, hasn, n[0], n[1], DIM == 3 ? n[2] : 0.0) *
                             op_col(sm, m) * gb;
#pragma unroll
            for (int k = 0; k < DPG_MAX_COEFS; ++k)
              if (k == sm.slot) acc[k] += d;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < DPG_MAX_COEFS; ++k) {
    if (k >= ncoef) break;
    __syncthreads();
    red[t] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) grad[(size_t)c * ncoef + k] = red[0];
  }
}

// Per cell c of n doubles each: out[2c] = sum (a - b)^2, out[2c+1] = sum b^2
// (one workgroup per cell, a fixed reduction tree).
__global__ __launch_bounds__(256) void cell_diff_norms(const double* __restrict__ a, const double* __restrict__ b,
                                                       int64_t n, double* __restrict__ out) {
  __shared__ double sd[256], sb[256];
  const int c = blockIdx.x, t = threadIdx.x;
  const double* ac = a + (size_t)c * n;
  const double* bc = b + (size_t)c * n;
  double d2 = 0.0, b2 = 0.0;
  for (int64_t i = t; i < n; i += 256) {
    const double d = ac[i] - bc[i];
    d2 = fma(d, d, d2);
    b2 = fma(bc[i], bc[i], b2);
  }
  sd[t] = d2;
  sb[t] = b2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      sd[t] += sd[t + h];
      sb[t] += sb[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[2 * (size_t)c] = sd[0];
    out[2 * (size_t)c + 1] = sb[0];
  }
}

// ------------------------------------------------------------ host launchers
static void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

void launch_point(int dim, int kind, const DevProgram* prog, const DevTables& tab, int npts, int Qs, int ncells,
                  const double* nodes, const int8_t* par, const double* coefs, int ncoef, int nsides,
                  const dpg_load_desc& load, double* W, double* Wrhs, hipStream_t st) {
  int total = ncells * npts;
  dim3 grid((total + 127) / 128);
  if (dim == 2)
    point_kernel<2><<<grid, 128, 0, st>>>(kind, prog, tab, npts, Qs, ncells, nodes, par, coefs, ncoef, nsides, load,
                                           W, Wrhs);
  else
    point_kernel<3><<<grid, 128, 0, st>>>(kind, prog, tab, npts, Qs, ncells, nodes, par, coefs, ncoef, nsides, load,
                                           W, Wrhs);
  check_launch("point_kernel");
}

void launch_operand(int dim, const double* W, int ng, int nvar, const double* R, const int32_t* vod, int N, int npts,
                    int ncells, double* out, int64_t sc, int64_t sb, int64_t sp, hipStream_t st) {
  if (ng == 0 || npts == 0) return;
  dim3 grid((N + 255) / 256, (npts + kOperandPoints - 1) / kOperandPoints, ncells);
  size_t shm = operand_lds_bytes(ng, nvar, dim);
  if (shm > kMaxLds) throw DeviceError("operand_kernel: LDS request above 64 KiB");
  if (dim == 2)
    operand_kernel<double, 3><<<grid, 256, shm, st>>>(W, ng, nvar, R, vod, N, npts, out, sc, sb, sp);
  else
    operand_kernel<double, 4><<<grid, 256, shm, st>>>(W, ng, nvar, R, vod, N, npts, out, sc, sb, sp);
  check_launch("operand_kernel");
}

void launch_rhs(int dim, const double* Wrhs, int nvar, const double* R, const int32_t* vod, int N, int Q, int ncells,
                double* A, int64_t lda, int64_t sA, int row_l, const double* s, int64_t ss, hipStream_t st) {
  dim3 grid((N + 255) / 256, ncells);
  if (dim == 2)
    rhs_kernel<3><<<grid, 256, 0, st>>>(Wrhs, nvar, R, vod, N, Q, A, lda, sA, row_l, s, ss);
  else
    rhs_kernel<4><<<grid, 256, 0, st>>>(Wrhs, nvar, R, vod, N, Q, A, lda, sA, row_l, s, ss);
  check_launch("rhs_kernel");
}

void launch_ip_operand(int dim, const double* Wip, int n_lt, int nvar, const double* R, const int32_t* vod, int N,
                       int QG, int QGp, int ncells, double* out, int64_t sc, int64_t ld, double* s, int64_t ss,
                       hipStream_t st) {
  size_t shm = ip_operand_lds_bytes(n_lt, nvar, dim);
  if (shm > kMaxLds) throw DeviceError("ip_operand_kernel: LDS request above 64 KiB");
  dim3 grid((N + 255) / 256, ncells);
  if (dim == 2)
    ip_operand_kernel<3><<<grid, 256, shm, st>>>(Wip, n_lt, nvar, R, vod, N, QG, QGp, out, sc, ld, s, ss);
  else
    ip_operand_kernel<4><<<grid, 256, shm, st>>>(Wip, n_lt, nvar, R, vod, N, QG, QGp, out, sc, ld, s, ss);
  check_launch("ip_operand_kernel");
}

void launch_gemm(GemmArgs g, int role, hipStream_t st, int tile) {
  if (g.m <= 0 || g.n <= 0 || g.ncells <= 0) return;
  if (tile != 64 && tile != 32 && tile != 128) throw DeviceError("gemm tile must be 128, 64 or 32");
  // the 128 x 64 tile serves dense (not lower-triangle) updates without k masks;
  // other calls of a 128 choice take the 64 x 64 tile
  if (tile == 128 && (role != kGemmUpdate || g.lower || g.kmask || g.kw)) tile = 64;
  if (tile == 128) {
    g.tiles_m = (g.m + 127) / 128;
    g.ntiles = g.tiles_m * ((g.n + 63) / 64);
    int64_t nb = (int64_t)g.ntiles * g.ncells;
    if (nb > 0x7fffffff) throw DeviceError("gemm grid too large");
    gemm_update_f64_w128<<<dim3((unsigned)nb), 256, 0, st>>>(g);
    check_launch("gemm_update_f64_w128");
    return;
  }
  g.tiles_m = (g.m + tile - 1) / tile;
  int tiles_n = (g.n + tile - 1) / tile;
  g.ntiles = g.lower ? g.tiles_m * (g.tiles_m + 1) / 2 : g.tiles_m * tiles_n;
  int64_t nblocks = (int64_t)g.ntiles * g.ncells;
  if (nblocks > 0x7fffffff) throw DeviceError("gemm grid too large");
  dim3 grid((unsigned)nblocks);
  if (tile == 64) {
    if (role == kGemmGram) gemm_gram_f64<<<grid, 256, 0, st>>>(g);
    else if (role == kGemmBmat) gemm_bmat_f64<<<grid, 256, 0, st>>>(g);
    else gemm_update_f64<<<grid, 256, 0, st>>>(g);
  } else {
    if (role == kGemmGram) gemm_gram_f64_t32<<<grid, 64, 0, st>>>(g);
    else if (role == kGemmBmat) gemm_bmat_f64_t32<<<grid, 64, 0, st>>>(g);
    else gemm_update_f64_t32<<<grid, 64, 0, st>>>(g);
  }
  check_launch("gemm_tn_f64");
}

void launch_ip_digits(int dim, const double* Wip, int n_lt, int nvar, const double* R, const int32_t* vod, int N,
                      int QG, int QGp, int NR, int ncells, int8_t* dig, int64_t sd, double* s, int64_t ss,
                      hipStream_t st) {
  if (QGp % 32 != 0 || NR % 64 != 0 || NR < N) throw DeviceError("ip_digits: QGp must be a multiple of 32, NR of 64");
  size_t shm = ip_digits_lds_bytes(n_lt, nvar, dim);
  if (shm > kMaxLds) throw DeviceError("ip_digits_kernel: LDS request above 64 KiB");
  dim3 grid((NR + 127) / 128, ncells);  // two lanes per row
  if (dim == 2)
    ip_digits_kernel<3><<<grid, 256, shm, st>>>(Wip, n_lt, nvar, R, vod, N, QG, QGp, NR, dig, sd, s, ss);
  else
    ip_digits_kernel<4><<<grid, 256, shm, st>>>(Wip, n_lt, nvar, R, vod, N, QG, QGp, NR, dig, sd, s, ss);
  check_launch("ip_digits_kernel");
}

void launch_gram_i8(I8GramArgs a, hipStream_t st) {
  if (a.N <= 0 || a.ncells <= 0) return;
  if (a.NR % 64 != 0 || a.NR < a.N) throw DeviceError("gram_i8: NR must be a multiple of 64 and >= N");
  if (a.kmask && (a.kseg % 32 != 0 || a.kseg <= 0)) throw DeviceError("gram_i8: kseg must be a multiple of 32");
  // exact int32 sums: a level holds at most kI8Digits products of |d| <= 64 per k
  if ((int64_t)a.nch * 32 * kI8Digits * 64 * 64 >= (int64_t)1 << 31) throw DeviceError("gram_i8: k too large");
  a.tiles_m = (a.N + 63) / 64;
  a.ntiles = a.tiles_m * (a.tiles_m + 1) / 2;
  int64_t nblocks = (int64_t)a.ntiles * a.ncells;
  if (nblocks > 0x7fffffff) throw DeviceError("gemm grid too large");
  gemm_gram_i8<<<dim3((unsigned)nblocks), 256, 0, st>>>(a);
  check_launch("gemm_gram_i8");
}

void launch_gemm_mixed(GemmArgs g, int precision, int role, hipStream_t st) {
  if (g.m <= 0 || g.n <= 0 || g.ncells <= 0) return;
  g.tiles_m = (g.m + GT - 1) / GT;
  int tiles_n = (g.n + GT - 1) / GT;
  g.ntiles = g.lower ? g.tiles_m * (g.tiles_m + 1) / 2 : g.tiles_m * tiles_n;
  int64_t nblocks = (int64_t)g.ntiles * g.ncells;
  if (nblocks > 0x7fffffff) throw DeviceError("gemm grid too large");
  dim3 grid((unsigned)nblocks);
  bool gram = role == kGemmGram;
  if (precision == DPG_PREC_FP32) {
    if (gram) gemm_gram_f32<<<grid, 256, 0, st>>>(g);
    else gemm_bmat_f32<<<grid, 256, 0, st>>>(g);
  } else if (precision == DPG_PREC_BF16X3) {
    if (gram) gemm_gram_bf16x3<<<grid, 256, 0, st>>>(g);
    else gemm_bmat_bf16x3<<<grid, 256, 0, st>>>(g);
  } else if (precision == DPG_PREC_FP16X3) {
    if (gram) gemm_gram_f16x3<<<grid, 256, 0, st>>>(g);
    else gemm_bmat_f16x3<<<grid, 256, 0, st>>>(g);
  } else {
    throw DeviceError("unknown precision");
  }
  check_launch("gemm_mixed");
}

void launch_potrf(double* A, int64_t lda, int64_t sA, int k0, int kb, double* Ld, int64_t sLd, double* Lt, int ncells,
                  int32_t* info, int info_base, hipStream_t st) {
  if (kb <= 0 || kb > PNB) throw DeviceError("potrf_wave: panel width must be 1..64");
  const dim3 grid((ncells + 3) / 4);
  if (kb <= 8) potrf_wave<8><<<grid, 256, 0, st>>>(A, lda, sA, k0, kb, Ld, sLd, Lt, info, info_base, ncells);
  else if (kb <= 16) potrf_wave<16><<<grid, 256, 0, st>>>(A, lda, sA, k0, kb, Ld, sLd, Lt, info, info_base, ncells);
  else if (kb <= 32) potrf_wave<32><<<grid, 256, 0, st>>>(A, lda, sA, k0, kb, Ld, sLd, Lt, info, info_base, ncells);
  else potrf_wave<64><<<grid, 256, 0, st>>>(A, lda, sA, k0, kb, Ld, sLd, Lt, info, info_base, ncells);
  check_launch("potrf_wave");
}

void launch_update_solve(GemmArgs g, hipStream_t st) {
  if (g.m <= 0 || g.ncells <= 0) return;
  if (g.n <= 0 || g.n > 64 || !g.Lt || g.lower || g.skip_upper || g.kmask || g.kw || g.srow || g.scol)
    throw DeviceError("update_solve_f64: a dense panel of at most 64 columns with its diagonal factor");
  g.tiles_m = (g.m + 127) / 128;
  g.ntiles = g.tiles_m;
  int64_t nb = (int64_t)g.ntiles * g.ncells;
  if (nb > 0x7fffffff) throw DeviceError("gemm grid too large");
  if (g.n <= 32) {
    update_solve_f64_n32<<<dim3((unsigned)nb), 256, 0, st>>>(g);
    check_launch("update_solve_f64_n32");
    return;
  }
  update_solve_f64<<<dim3((unsigned)nb), 256, 0, st>>>(g);
  check_launch("update_solve_f64");
}

void launch_panel(double* A, int64_t lda, int64_t sA, int D, int k0, int kb, double* Ld, int64_t sLd, double* Lt,
                  int ncells, int32_t* info, int info_base, hipStream_t st) {
  launch_potrf(A, lda, sA, k0, kb, Ld, sLd, Lt, ncells, info, info_base, st);
  int rows = D - k0 - kb;
  if (rows <= 0) return;
  const dim3 tg((rows + 255) / 256, ncells);
  if (kb <= 8) panel_trsm<8><<<tg, 256, 0, st>>>(A, lda, sA, D, k0, kb, Lt);
  else if (kb <= 16) panel_trsm<16><<<tg, 256, 0, st>>>(A, lda, sA, D, k0, kb, Lt);
  else if (kb <= 32) panel_trsm<32><<<tg, 256, 0, st>>>(A, lda, sA, D, k0, kb, Lt);
  else panel_trsm<64><<<tg, 256, 0, st>>>(A, lda, sA, D, k0, kb, Lt);
  check_launch("panel_trsm");
}

__global__ void fill_i32(int32_t* __restrict__ p, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// One workgroup: copies info[0..n) to out (when out is not NULL) and writes the
// number of non-zero entries to *count (per-thread partial sums, then a fixed
// tree in LDS).
__global__ __launch_bounds__(256) void info_count(const int32_t* __restrict__ info, int n, int32_t* __restrict__ out,
                                                  int32_t* __restrict__ count) {
  __shared__ int32_t part[256];
  int c = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    int32_t v = info[i];
    if (out) out[i] = v;
    c += v != 0;
  }
  part[threadIdx.x] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = part[0];
}

void launch_gram_sf(int dim, const SfBlockPair* bps, int nbp, const SfPair* pairs, const double* F, int nF, int q,
                    const double* Wip, int n_lt, int na, int QG, double* A, int64_t lda, int64_t sA, int ncells,
                    size_t lds_bytes, hipStream_t st, const double* s, int64_t ss, bool diag_only) {
  if (ncells <= 0 || nbp <= 0) return;
  if (lds_bytes > 64 * 1024) throw DeviceError("gram_sf: LDS request above 64 KiB");
  dim3 grid(nbp, ncells);
#define DPG_SF_CASE(D, QQ)                                                                                         \
  case QQ:                                                                                                         \
    if (diag_only)                                                                                                 \
      gram_sf_diag_kernel<D, QQ><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wip, n_lt, na, QG, A, lda, sA); \
    else                                                                                                           \
      gram_sf_kernel<D, QQ><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wip, n_lt, na, QG, A, lda, sA, s, ss); \
    break;
  if (dim == 3) {
    switch (q) {
      DPG_SF_CASE(3, 3) DPG_SF_CASE(3, 4) DPG_SF_CASE(3, 5) DPG_SF_CASE(3, 6) DPG_SF_CASE(3, 7) DPG_SF_CASE(3, 8)
      DPG_SF_CASE(3, 9) DPG_SF_CASE(3, 10)
      default:
        if (diag_only)
          gram_sf_diag_kernel<3, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wip, n_lt, na, QG, A, lda, sA);
        else
          gram_sf_kernel<3, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wip, n_lt, na, QG, A, lda, sA, s, ss);
    }
  } else {
    switch (q) {
      DPG_SF_CASE(2, 3) DPG_SF_CASE(2, 4) DPG_SF_CASE(2, 5) DPG_SF_CASE(2, 6) DPG_SF_CASE(2, 7) DPG_SF_CASE(2, 8)
      DPG_SF_CASE(2, 9) DPG_SF_CASE(2, 10) DPG_SF_CASE(2, 11) DPG_SF_CASE(2, 12)
      default:
        if (diag_only)
          gram_sf_diag_kernel<2, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wip, n_lt, na, QG, A, lda, sA);
        else
          gram_sf_kernel<2, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wip, n_lt, na, QG, A, lda, sA, s, ss);
    }
  }
#undef DPG_SF_CASE
  check_launch("gram_sf_kernel");
}

void launch_bmat_sf(int dim, const SfBlockPair* bps, int nbp, const SfPair* pairs, const double* F, int nF, int q,
                    const double* Wvol, int na, int Q, double* A, int64_t lda, int64_t sA, const double* s,
                    int64_t ss, int ncells, size_t lds_bytes, hipStream_t st) {
  if (ncells <= 0 || nbp <= 0) return;
  if (lds_bytes > 64 * 1024) throw DeviceError("bmat_sf: LDS request above 64 KiB");
  dim3 grid(nbp, ncells);
#define DPG_SFB_CASE(D, QQ) \
  case QQ:                  \
    bmat_sf_kernel<D, QQ><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wvol, na, Q, A, lda, sA, s, ss); \
    break;
  if (dim == 3) {
    switch (q) {
      DPG_SFB_CASE(3, 3) DPG_SFB_CASE(3, 4) DPG_SFB_CASE(3, 5) DPG_SFB_CASE(3, 6) DPG_SFB_CASE(3, 7)
      DPG_SFB_CASE(3, 8) DPG_SFB_CASE(3, 9) DPG_SFB_CASE(3, 10)
      default:
        bmat_sf_kernel<3, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wvol, na, Q, A, lda, sA, s, ss);
    }
  } else {
    switch (q) {
      DPG_SFB_CASE(2, 3) DPG_SFB_CASE(2, 4) DPG_SFB_CASE(2, 5) DPG_SFB_CASE(2, 6) DPG_SFB_CASE(2, 7)
      DPG_SFB_CASE(2, 8) DPG_SFB_CASE(2, 9) DPG_SFB_CASE(2, 10) DPG_SFB_CASE(2, 11) DPG_SFB_CASE(2, 12)
      default:
        bmat_sf_kernel<2, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, q, Wvol, na, Q, A, lda, sA, s, ss);
    }
  }
#undef DPG_SFB_CASE
  check_launch("bmat_sf_kernel");
}

void launch_bmat_side_sf(int dim, const SfBlockPair* bps, int nbp, const SfPair* pairs, const double* F, int nF,
                         int nfam, int q, const double* Wside, int npts, int na, double* A, int64_t lda, int64_t sA,
                         const double* s, int64_t ss, int ncells, size_t lds_bytes, hipStream_t st) {
  if (ncells <= 0 || nbp <= 0) return;
  if (lds_bytes > 64 * 1024) throw DeviceError("bmat_side_sf: LDS request above 64 KiB");
  dim3 grid(nbp, ncells);
#define DPG_SFS_CASE(D, QQ)                                                                                     \
  case QQ:                                                                                                      \
    bmat_side_sf_kernel<D, QQ><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, nfam, q, Wside, npts, na, A, lda, \
                                                              sA, s, ss);                                       \
    break;
  if (dim == 3) {
    switch (q) {
      DPG_SFS_CASE(3, 3) DPG_SFS_CASE(3, 4) DPG_SFS_CASE(3, 5) DPG_SFS_CASE(3, 6) DPG_SFS_CASE(3, 7)
      DPG_SFS_CASE(3, 8)
      default:
        bmat_side_sf_kernel<3, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, nfam, q, Wside, npts, na, A, lda,
                                                                 sA, s, ss);
    }
  } else {
    switch (q) {
      DPG_SFS_CASE(2, 3) DPG_SFS_CASE(2, 4) DPG_SFS_CASE(2, 5) DPG_SFS_CASE(2, 6) DPG_SFS_CASE(2, 7)
      DPG_SFS_CASE(2, 8) DPG_SFS_CASE(2, 9) DPG_SFS_CASE(2, 10)
      default:
        bmat_side_sf_kernel<2, 0><<<grid, 256, lds_bytes, st>>>(bps, pairs, F, nF, nfam, q, Wside, npts, na, A, lda,
                                                                 sA, s, ss);
    }
  }
#undef DPG_SFS_CASE
  check_launch("bmat_side_sf_kernel");
}

// a kernel rather than hipMemsetAsync, so that captured graphs replay it reliably
void launch_fill_i32(int32_t* p, int n, int32_t v, hipStream_t st) {
  if (n <= 0) return;
  fill_i32<<<(n + 255) / 256, 256, 0, st>>>(p, n, v);
  check_launch("fill_i32");
}

void launch_info_count(const int32_t* info, int n, int32_t* out, int32_t* count, hipStream_t st) {
  info_count<<<1, 256, 0, st>>>(info, n, out, count);
  check_launch("info_count");
}

void launch_zero_trailing(double* A, int64_t lda, int64_t sA, int r0, int D, int ncells, hipStream_t st) {
  int n = D - r0;
  if (n <= 0) return;
  zero_trailing<<<dim3((n + 255) / 256, n, ncells), 256, 0, st>>>(A, lda, sA, r0, D);
  check_launch("zero_trailing");
}

void launch_scaling(double* A, int64_t lda, int64_t sA, int r0, int D, int n, int ncol, double* s, int64_t ss,
                    int ncells, double sign, hipStream_t st) {
  diag_scale_factors<<<dim3((n + 255) / 256, ncells), 256, 0, st>>>(A, lda, sA, r0, n, s, ss, sign > 0 ? 1 : -1);
  check_launch("diag_scale_factors");
  apply_scaling<<<dim3((ncol + 3) / 4, ncells), 256, 0, st>>>(A, lda, sA, r0, D, n, ncol, s, ss, sign);
  check_launch("apply_scaling");
}

void launch_scale_factors(const double* A, int64_t lda, int64_t sA, int r0, int n, double* s, int64_t ss, int ncells,
                          hipStream_t st) {
  diag_scale_factors<<<dim3((n + 255) / 256, ncells), 256, 0, st>>>(A, lda, sA, r0, n, s, ss, 1);
  check_launch("diag_scale_factors");
}

void launch_residual_row(double* A, int64_t lda, int64_t sA, int N, int M, int r0, int row_l, int row_r,
                         const double* u, int ncells, hipStream_t st) {
  residual_row<<<dim3((N + 3) / 4, ncells), 256, 0, st>>>(A, lda, sA, N, M, r0, row_l, row_r, u);
  check_launch("residual_row");
}

void launch_extract_sym(const double* A, int64_t lda, int64_t sA, int r0, int n, double sign, double* out, int ncells,
                        hipStream_t st) {
  int nt = (n + 31) / 32;
  extract_sym<<<dim3(nt, nt, ncells), 256, 0, st>>>(A, lda, sA, r0, n, sign, out);
  check_launch("extract_sym");
}

void launch_recover_setup(const double* K, const double* rhs, const double* u, int M, int F, double* A, int64_t lda,
                          int64_t sA, double* s, int64_t ss, int ncells, hipStream_t st) {
  recover_setup<<<dim3(ncells), 256, 0, st>>>(K, rhs, u, M, F, A, lda, sA, s, ss);
  check_launch("recover_setup");
}

void launch_backsolve(const double* A, int64_t lda, int64_t sA, int N, int row_r, const double* Ld, int64_t sLd,
                      const double* s, int64_t ss, double* e, int64_t se, int ncells, hipStream_t st) {
  backsolve_kernel<<<dim3(ncells), 256, sizeof(double) * N, st>>>(A, lda, sA, N, row_r, Ld, sLd, s, ss, e, se);
  check_launch("backsolve_kernel");
}

void launch_extract_factor(const double* A, int64_t lda, int64_t sA, int N, int M, const double* Ld, int64_t sLd,
                           double* F, int64_t sF, int ncells, hipStream_t st) {
  if (ncells <= 0 || N <= 0) return;
  const int R = N + M + 1;
  extract_factor<<<dim3((N + 31) / 32, (R + 31) / 32, ncells), 256, 0, st>>>(A, lda, sA, N, R, Ld, sLd, F, sF);
  check_launch("extract_factor");
}

void launch_factored_residual(const double* F, int64_t sF, int N, int M, const double* u, double* z, int64_t sz,
                              double* eta2, int ncells, hipStream_t st) {
  if (ncells <= 0) return;
  factored_residual<<<ncells, 256, 0, st>>>(F, sF, N, M, u, z, sz, eta2);
  check_launch("factored_residual");
}

void launch_factor_backsolve(const double* F, int64_t sF, int N, int M, const double* s, int64_t ss, int nT,
                             int with_y, const double* z, int64_t sz, double* X, int64_t sX, int ncol, int ncells,
                             hipStream_t st) {
  if (ncells <= 0 || ncol <= 0) return;
  factor_backsolve<<<ncells, 256, 0, st>>>(F, sF, N, M, s, ss, nT, with_y, z, sz, X, sX, ncol);
  check_launch("factor_backsolve");
}

size_t adjoint_lds_bytes(int nvar, int dim, int ncol) {
  return sizeof(double) * ((size_t)nvar * (dim + 1) * ncol + ncol);
}

void launch_adjoint(int dim, const DevProgram* prog, const DevTables& tab, int N, int M, int QG, int Q, int Qs,
                    int nsides, int nvar, const double* nodes, const int8_t* par, const double* coefs,
                    const double* X, int64_t sX, int ncol, int nT, int cy, int ce, const double* Kbar,
                    const double* rbar, const double* u, double* grad, int ncells, hipStream_t st) {
  if (ncells <= 0) return;
  const size_t per_wave = adjoint_lds_bytes(nvar, dim, ncol), cap = 64 * 1024;
  if (per_wave > cap) throw DeviceError("adjoint: LDS request above 64 KiB");
  const int npw = (int)std::min<size_t>(4, cap / per_wave);  // waves sharing the points
  const size_t lds = per_wave * npw;
  if (dim == 3)
    adjoint_kernel<3><<<ncells, 256, lds, st>>>(prog, tab, N, M, QG, Q, Qs, nsides, nodes, par, coefs, X, sX, ncol,
                                                nT, cy, ce, Kbar, rbar, u, grad, npw);
  else
    adjoint_kernel<2><<<ncells, 256, lds, st>>>(prog, tab, N, M, QG, Q, Qs, nsides, nodes, par, coefs, X, sX, ncol,
                                                nT, cy, ce, Kbar, rbar, u, grad, npw);
  check_launch("adjoint_kernel");
}

void launch_cell_diff_norms(const double* a, const double* b, int64_t n, int ncells, double* out, hipStream_t st) {
  if (ncells <= 0) return;
  cell_diff_norms<<<ncells, 256, 0, st>>>(a, b, n, out);
  check_launch("cell_diff_norms");
}

void launch_extract_row(const double* A, int64_t lda, int64_t sA, int row, int r0, int n, double sign, double* out,
                        int ncells, hipStream_t st) {
  extract_row<<<dim3((n + 255) / 256, ncells), 256, 0, st>>>(A, lda, sA, row, r0, n, sign, out);
  check_launch("extract_row");
}

}  // namespace dpg

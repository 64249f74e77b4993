// kernels.hpp -- host launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/dpg_capi.h"
#include "device_common.hpp"

namespace dpg {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// C(i,j) = alpha * sum_k X(k,i) Y(k,j) + beta * C(i,j) for every cell of a batch
struct GemmArgs {
  const double* X;
  int64_t ldx, sx;  // X(k,i) = X[k*ldx + i], per-cell stride sx (0: shared)
  const double* Y;
  int64_t ldy, sy;
  double* C;
  int64_t ldc, sc;  // C(i,j) = C[i + j*ldc]
  int32_t m, n, k, lower;
  double alpha, beta;
  int32_t tiles_m, ntiles, ncells;
  int32_t skip_upper;  // 1: C(i,j) with i < j is not needed (not computed for whole 32x32 sub-tiles, not written)
  // optional: scale C(i,j) by srow[i] * scol[j] (per-cell vectors, stride ss)
  const double* srow;
  const double* scol;
  int64_t ss;
  // optional (mixed-precision GEMMs): scale X(k,i) by px[i] and Y(k,j) by py[j]
  // before the conversion to the low-precision operand format
  const double* px;
  const double* py;
  // optional (FP64 GEMMs): k runs over segments of kseg rows (kseg a multiple of
  // the 16-row k-chunk); bit s of kmask[tile] clear means segment s is zero for
  // this tile's rows or columns and is skipped (adds exact zeros otherwise)
  const uint64_t* kmask;
  int64_t kseg;
  // optional (with kmask): kw[tile*nseg + s] scales the staged X rows of segment s
  // (a segment standing for several equal terms counts once per term)
  const double* kw;
  int32_t nseg;
  // update_solve_f64 only: the panel's diagonal factor as potrf_wave leaves it
  // (per cell 64 x 64, Lt[j*64 + m] = L[m][j] below the diagonal, 1 / L[j][j] on it)
  const double* Lt;
};

void launch_point(int dim, int kind, const DevProgram* prog, const DevTables& tab, int npts, int Qs, int ncells,
                  const double* nodes, const int8_t* par, const double* coefs, int ncoef, int nsides,
                  const dpg_load_desc& load, double* W, double* Wrhs, hipStream_t st);
void launch_operand(int dim, const double* W, int ng, int nvar, const double* R, const int32_t* vod, int N, int npts,
                    int ncells, double* out, int64_t sc, int64_t sb, int64_t sp, hipStream_t st);
void launch_rhs(int dim, const double* Wrhs, int nvar, const double* R, const int32_t* vod, int N, int Q, int ncells,
                double* A, int64_t lda, int64_t sA, int row_l, const double* s, int64_t ss, hipStream_t st);
// Gram operand: out[c][q*n_lt + lt][j] = weighted value of IP term lt at point q
// for test dof j (G = out^T out); s[c][j] = 1/sqrt(G_jj)
void launch_ip_operand(int dim, const double* Wip, int n_lt, int nvar, const double* R, const int32_t* vod, int N,
                       int QG, int QGp, int ncells, double* out, int64_t sc, int64_t ld, double* s, int64_t ss,
                       hipStream_t st);
enum { kGemmGram = 0, kGemmBmat = 1, kGemmUpdate = 2 };
// points whose weights one operand workgroup stages in LDS, and the LDS bytes
// of the weight-staging kernels (at most kMaxLds per workgroup)
constexpr int kOperandPoints = 8;
constexpr size_t kMaxLds = 64 * 1024;
inline size_t operand_lds_bytes(int ng, int nvar, int dim) {
  return sizeof(double) * kOperandPoints * (size_t)ng * nvar * (dim + 1);
}
inline size_t ip_operand_lds_bytes(int n_lt, int nvar, int dim) { return sizeof(double) * (size_t)n_lt * nvar * (dim + 1); }
inline size_t ip_digits_lds_bytes(int n_lt, int nvar, int dim) {
  return sizeof(double) * 32 * (size_t)n_lt * nvar * (dim + 1);
}
void launch_gemm(GemmArgs g, int role, hipStream_t st, int tile = 64);
void launch_fill_i32(int32_t* p, int n, int32_t v, hipStream_t st);
// copies info (n) to out (optional) and writes the count of non-zero entries to *count
void launch_info_count(const int32_t* info, int n, int32_t* out, int32_t* count, hipStream_t st);
// sum-factorised Gram matrix, lower triangle of A's first N rows/columns: unscaled, or S G S
// with the DPOEQU factors s; diag_only: G's diagonal alone (what s is computed from)
void launch_gram_sf(int dim, const SfBlockPair* bps, int nbp, const SfPair* pairs, const double* F, int nF, int q,
                    const double* Wip, int n_lt, int na, int QG, double* A, int64_t lda, int64_t sA, int ncells,
                    size_t lds_bytes, hipStream_t st, const double* s = nullptr, int64_t ss = 0,
                    bool diag_only = false);
// s[i] = 1 / sqrt(A[r0 + i][r0 + i]), i < n (the DPOEQU factors alone)
void launch_scale_factors(const double* A, int64_t lda, int64_t sA, int r0, int n, double* s, int64_t ss, int ncells,
                          hipStream_t st);
// the volume part of B, sum-factorised: block pairs (trial field block, test
// block) and their pairs, 1D tables of 4 families on the stiffness points
// (q per direction), point weights Wvol (na per point), stores scaled by s
// (the DPOEQU factors of the test columns; none when s is NULL)
void launch_bmat_sf(int dim, const SfBlockPair* bps, int nbp, const SfPair* pairs, const double* F, int nF, int q,
                    const double* Wvol, int na, int Q, double* A, int64_t lda, int64_t sA, const double* s,
                    int64_t ss, int ncells, size_t lds_bytes, hipStream_t st);
// the side part of B, sum-factorised per side: side A a trace or flux block on
// one side (rows N + its offset), side B a test block restricted to the side;
// nfam 1D families per block pair from F + foff (tangential directions on the
// side's points, the normal one at the side's coordinate), point weights Wside
// (npts side points per cell, na per point, the pair's side from woff)
void launch_bmat_side_sf(int dim, const SfBlockPair* bps, int nbp, const SfPair* pairs, const double* F, int nF,
                         int nfam, int q, const double* Wside, int npts, int na, double* A, int64_t lda, int64_t sA,
                         const double* s, int64_t ss, int ncells, size_t lds_bytes, hipStream_t st);
// the same contraction with FP64 operands converted on load to FP32 (precision
// DPG_PREC_FP32) or to hi+lo BF16/FP16 pairs (DPG_PREC_BF16X3/FP16X3, products
// hi*hi + hi*lo + lo*hi), FP32 MFMA accumulation, FP64 epilogue
void launch_gemm_mixed(GemmArgs g, int precision, int role, hipStream_t st);
// Int8 digit split of the Gram product (precision DPG_PREC_INT8SPLIT).
// The equilibrated operand x(k,j) = s_j out(k,j), with s_j = 1/||out(:,j)|| so
// that |x| <= 1, is written as kI8Digits signed digits of 7 bits:
//   x = sum_p d_p 2^(1-7p) + r,   d_p in [-64, 64],   |r| <= 2^(-7 kI8Digits).
// Digit planes, bytes per cell: kI8Digits * (KG/32) * NR * 32, plane p, k-chunk c
// (32 consecutive k), row j at ((p*(KG/32) + c)*NR + j)*32 + k%32; NR = N rounded
// up to 64, rows N..NR-1 zero. s[c][j] as launch_ip_operand.
constexpr int kI8Digits = 8;
void launch_ip_digits(int dim, const double* Wip, int n_lt, int nvar, const double* R, const int32_t* vod, int N,
                      int QG, int QGp, int NR, int ncells, int8_t* dig, int64_t sd, double* s, int64_t ss,
                      hipStream_t st);
// G(i,j) (lower, i >= j, rows and columns < N) = sum_k x(k,i) x(k,j) from the
// digit planes: the digit products d_p(i) d_q(j) with p + q <= kI8Digits + 1
// on v_mfma_i32_32x32x32_i8, exact int32 sums per level p + q, the levels
// scaled by 2^(2-7(p+q)) and added in FP64 from the smallest. kmask/kseg as
// GemmArgs, kseg in k (a multiple of 32).
struct I8GramArgs {
  const int8_t* dig;
  int64_t sd;  // bytes per cell
  int32_t N, NR, nch, ncells;
  double* C;
  int64_t ldc, sc;
  const uint64_t* kmask;
  int64_t kseg;
  int32_t tiles_m, ntiles;
};
void launch_gram_i8(I8GramArgs a, hipStream_t st);
// Ld: per-cell (rows, 64) store of the factored diagonal blocks (row k0 + r, column r');
// Lt: per-cell 64x64 scratch for the current diagonal block, column-major, with
// the reciprocals of its diagonal
void launch_panel(double* A, int64_t lda, int64_t sA, int D, int k0, int kb, double* Ld, int64_t sLd, double* Lt,
                  int ncells, int32_t* info, int info_base, hipStream_t st);
// the diagonal block's factorisation alone (potrf_wave), for launch_update_solve
void launch_potrf(double* A, int64_t lda, int64_t sA, int k0, int kb, double* Ld, int64_t sLd, double* Lt, int ncells,
                  int32_t* info, int info_base, hipStream_t st);
// rows below a panel's diagonal block: C = beta C + alpha X^T Y (the panel
// update, g.k may be 0) and then C <- C L^-T with the diagonal factor g.Lt, in
// one kernel (128 x 64 tiles, g.n <= 64)
void launch_update_solve(GemmArgs g, hipStream_t st);
void launch_zero_trailing(double* A, int64_t lda, int64_t sA, int r0, int D, int ncells, hipStream_t st);
void launch_scaling(double* A, int64_t lda, int64_t sA, int r0, int D, int n, int ncol, double* s, int64_t ss,
                    int ncells, double sign, hipStream_t st);
void launch_residual_row(double* A, int64_t lda, int64_t sA, int N, int M, int r0, int row_l, int row_r,
                         const double* u, int ncells, hipStream_t st);
// per cell of n doubles: out[2c] = |a_c - b_c|^2, out[2c+1] = |b_c|^2
void launch_cell_diff_norms(const double* a, const double* b, int64_t n, int ncells, double* out, hipStream_t st);
void launch_extract_sym(const double* A, int64_t lda, int64_t sA, int r0, int n, double sign, double* out, int ncells,
                        hipStream_t st);
// field recovery after a condensed solve: S K_FF S and the row s (rhs_F - K_FT u_T) into A
void launch_recover_setup(const double* K, const double* rhs, const double* u, int M, int F, double* A, int64_t lda,
                          int64_t sA, double* s, int64_t ss, int ncells, hipStream_t st);
// e = S Lt^-T y (y = residual row of A): the Riesz representation and the field back solve
void launch_backsolve(const double* A, int64_t lda, int64_t sA, int N, int row_r, const double* Ld, int64_t sLd,
                      const double* s, int64_t ss, double* e, int64_t se, int ncells, hipStream_t st);
// the factored first N columns of the augmented matrices as row-major (N+M+1, N)
// blocks (Lt, (Lt^-1 S B)^T, (Lt^-1 S l)^T), per-cell stride sF
void launch_extract_factor(const double* A, int64_t lda, int64_t sA, int N, int M, const double* Ld, int64_t sLd,
                           double* F, int64_t sF, int ncells, hipStream_t st);
// z = Lt^-1 S (l - B u) and eta2 = z^T z from stored factors (z, eta2 optional)
void launch_factored_residual(const double* F, int64_t sF, int N, int M, const double* u, double* z, int64_t sz,
                              double* eta2, int ncells, hipStream_t st);
// X (N x ncol per cell) = S Lt^-T [Y (nT columns) | y_l (with_y) | z (z != NULL)]
void launch_factor_backsolve(const double* F, int64_t sF, int N, int M, const double* s, int64_t ss, int nT,
                             int with_y, const double* z, int64_t sz, double* X, int64_t sX, int ncol, int ncells,
                             hipStream_t st);
size_t adjoint_lds_bytes(int nvar, int dim, int ncol);
// grad (C, ncoef) of <Kbar, K> + <rbar, rhs> + eta2(u) from X = [G^-1 B | G^-1 l | G^-1 r]
void launch_adjoint(int dim, const DevProgram* prog, const DevTables& tab, int N, int M, int QG, int Q, int Qs,
                    int nsides, int nvar, const double* nodes, const int8_t* par, const double* coefs,
                    const double* X, int64_t sX, int ncol, int nT, int cy, int ce, const double* Kbar,
                    const double* rbar, const double* u, double* grad, int ncells, hipStream_t st);
void launch_extract_row(const double* A, int64_t lda, int64_t sA, int row, int r0, int n, double sign, double* out,
                        int ncells, hipStream_t st);

}  // namespace dpg

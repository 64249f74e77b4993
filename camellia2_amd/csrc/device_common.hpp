// device_common.hpp -- types shared by the HIP kernels and the host orchestration.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dpg {

// One summand of a compiled term: coef(scale, slot, fn, normal) * op(test var).
struct DSummand {
  int32_t var, op, slot, fn, normal, aux;  // aux: load component (rhs) / flux flag (side blocks)
  double scale;
};

constexpr int kMaxSummands = 256;
constexpr int kMaxGroups = 40;
constexpr int kMaxTestVars = 16;

// Compiled per-form "term program" evaluated at every cubature point.
// Groups: IP scalar terms (lt), volume trial blocks (field components),
// side trial variables (traces/fluxes), rhs (one group).
struct DevProgram {
  int32_t dim, nvar, npr;  // npr = dim + 1 reference prims per test dof
  int32_t n_lt, n_vb, n_sb, n_rhs;
  int32_t nverts;  // vertices of the cell topology (geometry map)
  int32_t lt_beg[kMaxGroups + 1];
  int32_t vb_beg[kMaxGroups + 1];
  int32_t sb_beg[kMaxGroups + 1];
  int32_t sb_flux[kMaxGroups];
  int32_t rhs_beg[2];
  // trial-dof offsets of the groups (adjoint): volume group -> block offset,
  // side group -> offset of its side-0 block (sides follow with stride card)
  int32_t vb_off[kMaxGroups];
  int32_t sb_off[kMaxGroups];
  int32_t sb_card[kMaxGroups];
  int32_t nf;
  int32_t ncoef;
  int32_t test_off[kMaxTestVars + 1];  // test dofs of variable v: [test_off[v], test_off[v + 1])
  DSummand s[kMaxSummands];
};

// Sum-factorised Gram matrix (tensor-product cells): the test functions of a
// block (HGRAD var, or one component of an HDIV var) are products of 1D
// families per direction: L (closed Lagrange), L' (its derivative), O (open
// Lagrange). A pair couples prim a of block A with prim b of block B:
//   G_AB += sum_q M_ab(q) R_a(i, q) R_b(j, q),  M_ab(q) = sum_lt W[q][lt][a] W[q][lt][b].
constexpr int kSfFamL = 0, kSfFamD = 1, kSfFamO = 2, kSfFamFld = 3;  // Fld: trial field basis (B only)
constexpr int kSfMaxCls = 9;
struct SfPair {
  int16_t a, b;       // (var * npr + prim) of the two sides
  int8_t fa[3], fb[3];  // 1D family per direction
  int8_t cls, pad;    // class of the direction-1 family pair
};
struct SfBlockPair {
  int32_t offA, offB;  // first row (side A) and column (side B) of the block pair
  int32_t nA[3], nB[3];  // 1D sizes per direction
  int32_t sB[3];       // column stride of each direction's index (side B)
  int32_t foff;        // first entry of the block pair's 1D tables
  int32_t woff;        // first point weight of the block pair (per cell)
  int32_t diag;        // A == B: lower triangle only
  int32_t npairs, pair0, ncls;
  int8_t clsfa[kSfMaxCls], clsfb[kSfMaxCls];  // direction-1 families of each class
  int16_t pad;
};

// per-bucket device tables (reference data shared by all cells)
struct DevTables {
  const double* R_vol;    // (Q, npr, N)
  const double* R_ip;     // (QG, npr, N)
  const double* R_side;   // (S*Qs, npr, N)
  const double* dN_vol;   // (Q, V, dim)
  const double* dN_ip;    // (QG, V, dim)
  const double* dN_side;  // (S*Qs, V, dim)
  const double* Nv_vol;   // (Q, V)
  const double* w_vol;    // (Q)
  const double* w_ip;     // (QG)
  const double* w_side;   // (S*Qs)
  const double* field_vol;  // (Q, nf)
  const double* trace;    // (Qs, ntr)
  const double* flux;     // (Qs, nfl)
  const double* side_tan;  // (S, 2, dim) reference side tangents
  const int32_t* var_of_dof;  // (N)
};

}  // namespace dpg

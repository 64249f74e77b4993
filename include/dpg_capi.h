/*
 * dpg_capi.h -- C ABI of the MI355X-native DPG local-assembly library
 * (libcamellia2_dpg.so).
 *
 * Plain C: fixed-size structs, plain pointers and sizes, no C++ or torch types.
 * Every entry point returns an int status:
 *     0   success
 *    <0   argument / shape / runtime error (message in dpg_last_error())
 *    >0   number of cells whose Gram matrix was not positive definite
 *         (their per-cell info[] holds the LAPACK POTRF convention: the
 *         1-based column at which the factorisation stopped).
 * No exception crosses this boundary.
 *
 * Layouts follow Camellia's Intrepid::FieldContainer (row-major):
 *   nodes      (C, V, d)       V = 4 (quad), 8 (hex, shards vertex order)
 *   parities   (C, S)          int8, +1 / -1 (cellID < neighbourID -> +1, boundary +1)
 *   K          (C, M, M)       optimal-test stiffness  B^T G^-1 B
 *   rhs        (C, M)          B^T G^-1 l
 *   G          (C, N, N)       test Gram matrix
 *   schur      (C, T, T)       K_TT - K_TF K_FF^-1 K_FT   (trace dofs after the field dofs)
 *   schur_rhs  (C, T)          rhs_T - K_TF K_FF^-1 rhs_F
 * M = trial dofs (fields first, then traces/fluxes side by side), N = test dofs.
 */
#ifndef DPG_CAPI_H
#define DPG_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPG_MAX_VARS 16
#define DPG_MAX_BTERMS 64
#define DPG_MAX_IPTERMS 32
#define DPG_MAX_LT_SUMMANDS 8
#define DPG_MAX_RHS_TERMS 8
#define DPG_MAX_LOAD_MODES 32
#define DPG_MAX_COEFS 8

/* cell topologies */
enum { DPG_QUAD = 0, DPG_HEX = 1, DPG_TRI = 2 };
/* trial variable kinds (VarFactory::fieldVar / traceVar / fluxVar).
 * With H1 order k: FIELD and FIELD_VECTOR use HGRAD Cn(k-1) (HVOL, identity
 * transform), TRACE uses Cn(k) on every side (HGRAD trace), FLUX uses
 * Cn(k-1) on every side (L2 flux, multiplied by the side parity). */
enum { DPG_FIELD = 0, DPG_FIELD_VECTOR = 1, DPG_TRACE = 2, DPG_FLUX = 3 };
/* test spaces: HGRAD Cn(kt), HDIV In(kt) */
enum { DPG_HGRAD = 0, DPG_HDIV = 1 };
/* scalar test operators */
enum {
  DPG_OP_VALUE = 0,      /* v (HGRAD)                         */
  DPG_OP_DX = 1,         /* dv/dx                             */
  DPG_OP_DY = 2,         /* dv/dy                             */
  DPG_OP_DZ = 3,         /* dv/dz                             */
  DPG_OP_X = 4,          /* tau_x (HDIV)                      */
  DPG_OP_Y = 5,          /* tau_y                             */
  DPG_OP_Z = 6,          /* tau_z                             */
  DPG_OP_DIV = 7,        /* div tau                           */
  DPG_OP_DOT_NORMAL = 8  /* tau . n  (boundary-only)          */
};
enum { DPG_FN_VALUE = 0, DPG_FN_SQRT = 1, DPG_FN_INV = 2 };
/* DPG_PREC_INT8SPLIT: the Gram product G = A^T A on int8 MFMA with the
   equilibrated operand split into 8 signed 7-bit digits (error-free splitting;
   digit products exact in int32, levels summed in FP64); B, l and everything
   after them FP64. G is not available as an output in this mode. */
enum { DPG_PREC_FP64 = 0, DPG_PREC_FP32 = 1, DPG_PREC_BF16X3 = 2, DPG_PREC_FP16X3 = 3, DPG_PREC_INT8SPLIT = 4 };

/* coefficient of a scalar product: scale * [fn(coef[c][slot])] * [n_comp] */
typedef struct {
  double scale;
  int32_t coef_slot;   /* -1: none, else per-cell coefficient slot        */
  int32_t coef_fn;     /* DPG_FN_VALUE, DPG_FN_SQRT or DPG_FN_INV (1/a)      */
  int32_t normal_comp; /* -1: none, else outward unit normal component    */
  int32_t pad;
} dpg_coef;

/* one scalar bilinear product  coef * (trial component) * (test operator).
 * It is integrated over the sides when the trial variable is a trace/flux,
 * the test operator is DOT_NORMAL or the coefficient holds a normal
 * (LinearTerm boundary-only summands), otherwise over the cell volume. */
typedef struct {
  int32_t trial_var;
  int32_t trial_comp; /* 0 for scalar variables, d for FIELD_VECTOR component d */
  int32_t test_var;
  int32_t test_op;
  dpg_coef coef;
} dpg_bterm;

typedef struct {
  int32_t test_var;
  int32_t test_op;
  dpg_coef coef;
} dpg_lt_summand;

/* scalar linear term of the test inner product: G += (lt, lt) */
typedef struct {
  int32_t n;
  int32_t pad;
  dpg_lt_summand s[DPG_MAX_LT_SUMMANDS];
} dpg_ipterm;

/* load functional term:  l_i += scale * (f_{load_comp}, op(psi_i)) */
typedef struct {
  int32_t test_var;
  int32_t test_op;
  int32_t load_comp;
  int32_t pad;
  double scale;
} dpg_rhsterm;

/* compiled BF + IP + RHS (the term tables of a Camellia formulation) */
typedef struct {
  int32_t spatial_dim;
  int32_t n_trial;
  int32_t n_test;
  int32_t n_coefs; /* per-cell coefficients consumed by coef_slot */
  int32_t trial_kind[DPG_MAX_VARS];
  int32_t test_space[DPG_MAX_VARS];
  int32_t n_bterms;
  int32_t n_ipterms;
  int32_t n_rhs;
  int32_t pad;
  dpg_bterm bterms[DPG_MAX_BTERMS];
  dpg_ipterm ipterms[DPG_MAX_IPTERMS];
  dpg_rhsterm rhs[DPG_MAX_RHS_TERMS];
} dpg_form_desc;

/* load f_c(x) = constant[c] + sum_m amp[c][m] prod_d sin(k[c][m][d] pi x_d) */
typedef struct {
  int32_t ncomp;
  int32_t nmodes[3];
  double constant[3];
  double amp[3][DPG_MAX_LOAD_MODES];
  int32_t k[3][DPG_MAX_LOAD_MODES][3];
} dpg_load_desc;

/* element type = bucket key (cell topology, H1 order, test degree) */
typedef struct {
  int32_t topology;
  int32_t h1_order;    /* k: fields Cn(k-1), traces Cn(k), fluxes Cn(k-1) */
  int32_t test_degree; /* kt: HGRAD Cn(kt), HDIV In(kt)                 */
  int32_t pad;
} dpg_bucket_desc;

/* derived sizes of one bucket */
typedef struct {
  int32_t N;        /* test dofs                          */
  int32_t M;        /* trial dofs                         */
  int32_t F;        /* field dofs (first F trial dofs)    */
  int32_t T;        /* trace/flux dofs  (M - F)           */
  int32_t Q;        /* volume points, stiffness/rhs cache */
  int32_t Qs;       /* points per side                    */
  int32_t QG;       /* volume points, Gram (IP) cache     */
  int32_t nsides;
  int32_t nverts;
  int32_t pad;
  double flops_per_cell; /* algorithmic flop count used for GFLOP/s */
} dpg_bucket_info;

typedef struct {
  int32_t precision;       /* DPG_PREC_*: arithmetic of the G/B/l contractions */
  int32_t refine;          /* 1: one FP64 refinement step on G^-1 B            */
  int32_t max_batch_cells; /* 0: derived from workspace_bytes                  */
  int32_t profile;         /* 1: time every kernel launch (dpg_plan_timings)   */
  double workspace_bytes;  /* device workspace budget, 0: 16 GiB               */
} dpg_options;

/* pipeline stages timed when dpg_options.profile is set */
enum {
  DPG_STAGE_POINT = 0,   /* per-point geometry and weights            */
  DPG_STAGE_OPERAND = 1, /* weighted operator values                  */
  DPG_STAGE_GRAM = 2,    /* G contraction (MFMA)                      */
  DPG_STAGE_BMAT = 3,    /* B and l contractions                      */
  DPG_STAGE_SCALE = 4,   /* DPOEQU scaling                            */
  DPG_STAGE_PANEL = 5,   /* diagonal-block Cholesky + panel solve     */
  DPG_STAGE_UPDATE = 6,  /* trailing updates (MFMA)                   */
  DPG_STAGE_EXTRACT = 7, /* K, rhs, G, Schur copies to the outputs    */
  DPG_STAGE_ADJOINT = 8, /* adjoint: back solves and point contractions */
  DPG_STAGE_SOLVE = 9,   /* panel update fused with the panel solve (MFMA) */
  DPG_NSTAGES = 10
};

/* accumulated per-stage kernel times (HIP events on the call's stream) with
 * the algorithmic flops and HBM bytes of the timed launches; the analogue of
 * the reference's timing callbacks (BF.h:40-41,190-191) */
typedef struct {
  double ms[DPG_NSTAGES];
  double flops[DPG_NSTAGES];
  double bytes[DPG_NSTAGES];
  int64_t launches[DPG_NSTAGES];
  int64_t cells;
} dpg_timings;

/* one call's inputs; all pointers are device pointers */
typedef struct {
  int32_t ncells;
  int32_t pad;
  const double* nodes;     /* (C, V, d)                */
  const int8_t* parities;  /* (C, S)                   */
  const double* coefs;     /* (C, n_coefs) or NULL     */
  dpg_load_desc load;
} dpg_batch;

/* one call's outputs; device pointers, NULL = not requested */
typedef struct {
  double* K;         /* (C, M, M) */
  double* rhs;       /* (C, M)    */
  double* G;         /* (C, N, N) */
  double* schur;     /* (C, T, T) */
  double* schur_rhs; /* (C, T)    */
  int32_t* info;     /* (C)       */
  /* The factorisation kept for dpg_energy_error_factored and dpg_adjoint
   * (both or neither): with S = diag(factor_scale) the DPOEQU factors and
   * S G S = L L^T the Cholesky factor of the equilibrated Gram matrix,
   *   factor (C, N + M + 1, N) row-major: rows 0..N-1 L (lower; 0 above the
   *          diagonal), rows N..N+M-1 (L^-1 S B)^T, row N+M (L^-1 S l)^T
   *   factor_scale (C, N): the diagonal of S. */
  double* factor;
  double* factor_scale;
} dpg_outputs;

typedef struct dpg_plan dpg_plan;

/* ---- library ----------------------------------------------------------- */
const char* dpg_version(void);
/* message of the last failing call on this thread */
const char* dpg_last_error(void);

/* ---- formulations (host only, no GPU) ------------------------------------
 * Term tables of Camellia's ultraweak formulations with the graph-norm inner
 * product (TBF::graphNorm) and the load functional. */
/* PoissonFormulation(spaceDim, conforming, ULTRAWEAK)  PoissonFormulation.cpp:30-88 */
int dpg_form_poisson_uw(int spatial_dim, dpg_form_desc* out);
/* ConvectionDiffusionReactionFormulation(ULTRAWEAK) with per-cell
 * coefficients [eps, beta_1..beta_d, alpha]; sqrt(eps) enters as in
 * ConvectionDiffusionReactionFormulation.cpp:71-111 */
int dpg_form_cdr_uw(int spatial_dim, dpg_form_desc* out);
/* ConvectionDiffusionFormulation (ultraweak, graph norm), ConvectionDiffusionFormulation.cpp:23-96,
 * with per-cell coefficients [eps, beta_1..beta_d]; (sigma/eps, tau) uses DPG_FN_INV */
int dpg_form_convection_diffusion_uw(int spatial_dim, dpg_form_desc* out);
/* StokesVGPFormulation steady, conforming traces, StokesVGPFormulation.cpp:372-408, rhs :1015-1028 */
int dpg_form_stokes_vgp_uw(int spatial_dim, double mu, dpg_form_desc* out);
/* TBF::graphNorm (BF.cpp:1039-1141): replace out->ipterms from the bilinear terms */
int dpg_form_graph_norm(dpg_form_desc* form);

/* ---- seeded synthetic inputs (host only) -------------------------------- */
/* nx*ny quads on [0,1]^2, interior vertices jittered by U(-jitter, jitter)*h,
 * cells numbered x fastest; nodes (C,4,2) CCW; parities (C,4) */
int dpg_mesh_quad(int nx, int ny, double jitter, uint64_t seed, double* nodes, int8_t* parities);
/* nx*ny*nz hexes on [0,1]^3, shards Hexahedron<8> vertex order; nodes (C,8,3); parities (C,6) */
int dpg_mesh_hex(int nx, int ny, int nz, double jitter, uint64_t seed, double* nodes, int8_t* parities);
/* hp mesh graded toward the outflow edge x = 1: an n0*n0 grid whose column of
 * cells touching x = 1 is refined isotropically `levels` times (each level
 * halves h; 2:1 balanced, hanging nodes on the refinement fronts). Cells are
 * numbered coarse levels first, then y, then x; regular interior vertices are
 * jittered by U(-jitter, jitter) times the smallest adjacent h, hanging ones
 * stay on their edges. Parity rule of GlobalDofAssignment.cpp:165-185 with the
 * neighbour across the side (the coarse cell, or the first fine cell across a
 * refined side): +1 on the boundary or toward a higher cell number.
 * Returns the number of cells n0^2 + 3 n0 (2^levels - 1); with nodes = NULL
 * only the count. nodes (C,4,2) CCW, parities (C,4), level (C) or NULL. */
int dpg_mesh_quad_graded(int n0, int levels, double jitter, uint64_t seed, double* nodes, int8_t* parities,
                         int32_t* level);
/* The same two meshes generated on the device (device arrays nodes, parities,
 * stream-ordered; synchronises the stream): bit for bit the host generators'
 * output. Each vertex reaches its jitter draws in the generator's stream by a
 * jump ahead of the linear xoshiro256** state. */
int dpg_mesh_quad_device(int nx, int ny, double jitter, uint64_t seed, double* nodes, int8_t* parities,
                         void* stream);
int dpg_mesh_hex_device(int nx, int ny, int nz, double jitter, uint64_t seed, double* nodes, int8_t* parities,
                        void* stream);
/* sine-mode load: nmodes modes per component, amp ~ N(0,1), k ~ U{1..kmax} */
int dpg_load_sine_modes(int spatial_dim, int ncomp, int nmodes, int kmax, uint64_t seed, dpg_load_desc* out);
/* xoshiro256** seeded by splitmix64: n uniforms on [lo,hi) / n normals (Box-Muller) */
int dpg_rng_uniform(uint64_t seed, double lo, double hi, double* out, size_t n);
int dpg_rng_normal(uint64_t seed, double* out, size_t n);

/* sizes / algorithmic flop count of a bucket (host only) */
int dpg_bucket_sizes(const dpg_form_desc* form, const dpg_bucket_desc* bucket, dpg_bucket_info* out);

/* ---- GPU plan and assembly ---------------------------------------------- */
/* Builds the device basis/cubature tables of every bucket (the BasisCache
 * role, BasisCache.cpp:236-267) on the current HIP device. */
int dpg_plan_create(const dpg_form_desc* form, const dpg_bucket_desc* buckets, int nbuckets,
                    const dpg_options* options, dpg_plan** out);
int dpg_plan_destroy(dpg_plan* plan);
/* A second handle on a plan for concurrent calls: the device tables (bases,
 * cubature, term programs, tile choices as they are at the call) are shared
 * and read-only; the clone has its own workspace, INFO buffer and profiler, so
 * calls on the two handles may run at the same time on different streams
 * (calls on one handle are stream-ordered and must not overlap).
 * workspace_bytes > 0 replaces the original's workspace budget. The tables
 * are freed with the last handle (dpg_plan_destroy on each). The reference has
 * no counterpart: TBF::localStiffnessMatrixAndRHS (BF.h:122-125) is single
 * threaded and its BasisCaches memoise mutable state, so one call at a time. */
int dpg_plan_clone_workspace(const dpg_plan* plan, double workspace_bytes, dpg_plan** out);
/* bytes of device workspace the plan holds */
int dpg_plan_workspace(const dpg_plan* plan, size_t* bytes);
/* timings accumulated since the last reset (needs options.profile = 1) */
int dpg_plan_timings(dpg_plan* plan, dpg_timings* out, int reset);

/* Local assembly of one element type: replaces TBF::localStiffnessMatrixAndRHS
 * (BF.cpp:1308-1496, FACTORED_CHOLESKY branch) for all cells of the batch,
 * plus the static condensation of CondensedDofInterpreter::interpretLocalData
 * (CondensedDofInterpreter.cpp:1063-1225) when outputs->schur is set.
 * Stream-ordered on `stream` (a hipStream_t, NULL = default stream).
 * Returns after the work is enqueued unless outputs->info is non-NULL, in
 * which case it synchronises the stream and returns the non-SPD cell count. */
int dpg_assemble(dpg_plan* plan, int bucket, const dpg_batch* batch, const dpg_outputs* outputs,
                 void* stream);

/* dpg_assemble without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same inputs,
 * outputs, results and return value as dpg_assemble (outputs->G excepted:
 * supported as there), but the workspace is a static device array of
 * DPG_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on until
 * the process ends, instead of the plan's workspace: the cells are assembled in
 * batches that fit it. A call whose element type needs more than the scratch
 * for a single cell fails with a negative status and a message (dpg_assemble
 * has no such limit). No per-launch timings are recorded. All calls on one GPU
 * share the scratch, whatever the plan: calls that overlap in time on one GPU
 * (other host threads, other streams) race and give wrong results. The current
 * device must be the plan's device. */
#define DPG_NOALLOC_SCRATCH_BYTES (256u << 20)
int dpg_assemble_noalloc(dpg_plan* plan, int bucket, const dpg_batch* batch, const dpg_outputs* outputs,
                         void* stream);

/* Energy-error indicator of TSolution::rankLocalEnergyError (Solution.cpp:2653-2798):
 * r = l - B u,  G e = r,  eta2[c] = r^T e  (= e^T G e; Camellia reports sqrt). */
int dpg_energy_error(dpg_plan* plan, int bucket, const dpg_batch* batch, const double* u,
                     double* eta2, void* stream);

/* The energy error from a stored factorisation (outputs->factor of a
 * dpg_assemble of the same cells): eta2[c] = |y_l - Y u|^2 with Y = L^-1 S B
 * and y_l = L^-1 S l the factor's rows, i.e. r^T G^-1 r with r = l - B u as in
 * TSolution::computeErrorRepresentation (Solution.cpp:2688-2722) -- without
 * re-assembling G and B or refactoring. u (C, M), eta2 (C). */
int dpg_energy_error_factored(dpg_plan* plan, int bucket, int ncells, const double* factor, const double* u,
                              double* eta2, void* stream);

/* Adjoint of the local assembly with respect to the per-cell PDE coefficients
 * (no counterpart in the reference, which has no derivatives). For every cell
 *   L_c = <Kbar_c, K_c> + <rhsbar_c, rhs_c> + eta2_c(u_c),
 * each part present only when its input (Kbar (C, M, M), rhsbar (C, M), u
 * (C, M)) is not NULL, and grad[c][k] = d L_c / d coefs[c][k] (C, n_coefs):
 *   with T = G^-1 B, y = G^-1 l, e = G^-1 (l - B u), Ks = sym(Kbar),
 *   Bbar = 2 T Ks + y rhsbar^T - 2 e u^T,  Gbar = -T Ks T^T - sym(T rhsbar y^T) - e e^T,
 *   grad = <Bbar, dB/dcoef> + <Gbar, dG/dcoef>,
 * the basis values recomputed at the quadrature points. eta2 (C, needs u) is
 * r^T G^-1 r as dpg_energy_error. factor / factor_scale: the factorisation a
 * dpg_assemble of the same cells kept (outputs->factor): then nothing is
 * re-assembled or refactored; NULL: the call assembles and factors G, B, l
 * itself. grad, eta2 may be NULL. Returns the count of non-SPD cells (0 with a
 * stored factor). */
int dpg_adjoint(dpg_plan* plan, int bucket, const dpg_batch* batch, const double* u, const double* Kbar,
                const double* rhsbar, const double* factor, const double* factor_scale, double* eta2, double* grad,
                void* stream);

/* ---- hp meshes: tile selection and the captured all-bucket launch --------
 * No counterpart in the reference, which assembles element types one after the
 * other on the CPU (Solution.cpp:1162-1298); these serve the same loop. */

/* FP64 contraction choices of one bucket: GEMM tile edge (64 or 32) per role,
 * and whether G and B come from the sum-factorised kernels (quads and hexes
 * only). sumfact = 1 contracts G = sum_q M(q) R R direction by direction
 * (about 60x fewer flops than A^T A on the 32^3 hex workload), and B's volume
 * and side parts the same way from the point weights (no B operand). Its
 * rounding differs from the dense path's: on the 32^3 hexes K is 9e-11 from an
 * 80-bit evaluation (the dense G: 6.6e-10; far fewer roundings per entry), on
 * the 2D Stokes mesh 1.5e-9 (dense: 2.4e-9). dpg_plan_autotune switches it on
 * only where it is faster and passes a gate on K (see there). */
typedef struct {
  int32_t gram;    /* G = A^T A (dense path)                        */
  int32_t bmat;    /* B blocks                                      */
  int32_t update;  /* Cholesky panel/trailing updates; also 128: the
                      rows below each diagonal block are updated
                      and solved against it in one kernel on
                      128 x 64 tiles (128 x 32 for a panel of at
                      most 32 columns; the diagonal blocks and the
                      trailing update then take 64)                 */
  int32_t sumfact; /* 1: sum-factorised G and B (see above)         */
} dpg_tiles;
int dpg_plan_get_tiles(const dpg_plan* plan, int bucket, dpg_tiles* out);
int dpg_plan_set_tiles(dpg_plan* plan, int bucket, const dpg_tiles* tiles);
/* Autotuner: times both tile edges for every role (and the 128 x 64 tile for
 * the updates) on the first `sample` cells of `batch` (HIP events, one warm-up
 * pass each) and keeps the fastest per role. Tile edges never change results
 * (every output accumulates in the same k order).
 * Then, for quads and hexes in FP64, the G path: the sum-factorised G is taken
 * when the whole pipeline runs faster with it and, on the first (at most 128)
 * sample cells, it moves K by no more than 4 times what a second rounding of
 * the dense G does (the dense G with its equal IP terms summed separately
 * instead of merged: the same products in another FP64 order); otherwise G is
 * dense. Blocks until done; `stream` runs the timings. */
int dpg_plan_autotune(dpg_plan* plan, int bucket, const dpg_batch* batch, int sample, void* stream);

/* one element type's share of a mesh: its cells, coefficients and outputs */
typedef struct {
  int32_t bucket;
  int32_t pad;
  dpg_batch batch;
  dpg_outputs out; /* info (C) optional: LAPACK INFO per cell, written on device */
} dpg_work_item;

typedef struct dpg_graph dpg_graph;
/* Captures the assembly of every item (dpg_assemble's pipeline, FP64 or the
 * plan's precision) into one hipGraph: items are ordered by decreasing flops
 * and dealt to `lanes` streams forked inside the graph (longest processing
 * time first), each lane with its own workspace of at most
 * options.workspace_bytes / lanes. The items' device pointers are baked into
 * the graph; replays recompute from their current contents. `stream` is only
 * used for capture and is not launched on. */
int dpg_graph_create(dpg_plan* plan, const dpg_work_item* items, int nitems, int lanes, void* stream,
                     dpg_graph** out);
/* one hipGraphLaunch of the captured assembly on `stream` (asynchronous) */
int dpg_graph_launch(dpg_graph* graph, void* stream);
/* device bytes held by the graph's lane workspaces */
int dpg_graph_workspace(const dpg_graph* graph, size_t* bytes);
int dpg_graph_destroy(dpg_graph* graph);

/* The error representation function of each cell (RieszRep, RieszRep.cpp:86,
 * 244; the e of TSolution::computeErrorRepresentation, Solution.cpp:2688-2722):
 * e (C, N) solves G e = l - B u in the cell's test basis (test dof order),
 * from the same factorisation dpg_energy_error uses; eta2 = r^T e. Returns
 * the number of non-SPD cells. */
int dpg_riesz_representation(dpg_plan* plan, int bucket, const dpg_batch* batch, const double* u, double* e,
                             void* stream);

/* Field dofs back from a condensed solve: for every cell of the batch,
 * u_F = K_FF^-1 (rhs_F - K_FT u_T), with K (C, M, M) and rhs (C, M) the
 * dpg_assemble outputs of these cells and u (C, M) the cells' trial
 * coefficients, whose trailing T entries (traces and fluxes, e.g. gathered
 * from a global solution) are read and whose leading F entries are written.
 * Replaces the field back-substitution of
 * CondensedDofInterpreter::interpretGlobalCoefficients
 * (CondensedDofInterpreter.cpp:1229-1330); K_FF is equilibrated and factored
 * by the same partial Cholesky as the local assembly (the reference uses an
 * LU). Returns the number of cells whose K_FF was not SPD. */
int dpg_recover_fields(dpg_plan* plan, int bucket, int ncells, const double* K, const double* rhs, double* u,
                       void* stream);

/* ------------------------------------------------------------------------
 * Preparation of the inputs on the device (device arrays, stream-ordered
 * unless stated).
 * ---------------------------------------------------------------------- */
/* Side parities (C, S) of a conforming mesh of one topology from its cell ->
 * vertex ids (C, V): +1 on the boundary, +1 for the lower cell ID of two peer
 * cells, -1 for the other -- the rule of the reference's cell side parities
 * (GlobalDofAssignment.cpp:160-185) for meshes without hanging sides. Sides
 * are matched by their vertex ids. Synchronises the stream; an error when a
 * side has more than two cells. */
int dpg_side_parities(int topology, int ncells, const int32_t* cell_vertices, int8_t* parities, void* stream);
/* Element-type bucketing (Solution.cpp:1162-1215 groups cells by ElementType):
 * perm (C) lists the cells sorted by type_key (>= 0), stably; bucket b holds
 * perm[offsets[b] .. offsets[b+1]) and has key bucket_keys[b]. offsets needs
 * room for C + 1 entries, bucket_keys for C. Synchronises the stream; returns
 * the number of buckets. */
int dpg_sort_cells(int ncells, const int32_t* type_key, int32_t* perm, int32_t* offsets, int32_t* bucket_keys,
                   void* stream);
/* Gathers cells into bucket order: out[i] = in[perm[i]] for the nodes (C,
 * nodes_per_cell doubles), parities (C, nsides) and coefficients (C, ncoefs);
 * any size may be 0 (its arrays unused). */
int dpg_gather_cells(int ncells, const int32_t* perm, int nodes_per_cell, const double* nodes, double* nodes_out,
                     int nsides, const int8_t* parities, int8_t* parities_out, int ncoefs, const double* coefs,
                     double* coefs_out, void* stream);

/* ------------------------------------------------------------------------
 * Condensed (trace and flux) system of a conforming mesh: local-to-global
 * interpretation of the cells' Schur complements and their sparse assembly.
 * Replaces GDAMinimumRule::interpretLocalData (GDAMinimumRule.cpp:714-780,
 * through LocalDofMapper) followed by the global insertion of
 * TSolution::populateStiffnessAndLoad (Solution.cpp:1263-1293) for the
 * condensed solve (CondensedDofInterpreter.cpp:1084,1194-1224), on meshes
 * without hanging sides and with one polynomial order (one bucket): H^1 traces
 * share vertex, edge and face nodes between cells, L^2 fluxes belong to one
 * side, shared entities are seen in one canonical orientation.
 * ---------------------------------------------------------------------- */
typedef struct dpg_skeleton dpg_skeleton;

/* Host. cell_vertices (C, V): global vertex ids, in the vertex order of the
 * cell's nodes in dpg_batch. gdof (C, T) receives the global index of each of
 * the cell's trailing trial dofs (trace and flux blocks in the cell's dof
 * order: the rows of dpg_outputs.schur). Returns the number of global dofs
 * (>= 0) or < 0 on error. */
int dpg_skeleton_numbering(const dpg_form_desc* form, const dpg_bucket_desc* bucket, int ncells,
                           const int32_t* cell_vertices, int32_t* gdof);
/* Host. Reference coordinates (T, dim) of the nodes of those T dofs (their
 * nodal side basis functions are 1 there). */
int dpg_skeleton_nodes(const dpg_form_desc* form, const dpg_bucket_desc* bucket, double* xi);
/* Device symbolic phase from a host gdof (C, T): the CSR pattern (rows of
 * int64 pointers, int32 columns, both triangles) of sum_c P_c^T S_c P_c and
 * the order in which each nonzero sums its contributions. Needs
 * C * T * T < 2^32. Synchronises `stream`. */
int dpg_skeleton_create(int ncells, int T, const int32_t* gdof, int nglobal, void* stream, dpg_skeleton** out);
/* The same for cells with different numbers of global dofs (hp meshes after
 * their constraints, dpg_hp_system): cell c holds the global dofs
 * dofs[offsets[c] .. offsets[c+1]) (host arrays); dpg_skeleton_assemble then
 * reads each cell's n_c x n_c block, row-major, one after the other (and n_c
 * load entries per cell). */
int dpg_skeleton_create_ragged(int ncells, const int32_t* offsets, const int32_t* dofs, int nglobal, void* stream,
                               dpg_skeleton** out);
/* nnz, global size and the device CSR pattern (owned by the skeleton; any
 * argument may be NULL) */
int dpg_skeleton_info(const dpg_skeleton* sk, int64_t* nnz, int64_t* nglobal, const int64_t** row_ptr,
                      const int32_t** col);
/* copies the CSR pattern into caller-owned device buffers row_ptr (nglobal + 1)
 * and col (nnz), stream-ordered */
int dpg_skeleton_pattern(const dpg_skeleton* sk, int64_t* row_ptr, int32_t* col, void* stream);
/* Device numeric phase, stream-ordered: values (nnz) from schur (C, T, T) and
 * rhs (nglobal) from schur_rhs (C, T), every sum in increasing (cell, local
 * row, local column) order, so the same bits on every call. values or rhs may
 * be NULL to skip that part. */
int dpg_skeleton_assemble(const dpg_skeleton* sk, const double* schur, const double* schur_rhs, double* values,
                          double* rhs, void* stream);
/* Host. entity (nglobal) receives the mesh entity of every global dof of the
 * numbering dpg_skeleton_numbering gives for the same cells: vertices first,
 * then edges, then (3D) faces; a side's flux dofs share its entity with the
 * trace dofs interior to it. Returns the number of entities or < 0 on error.
 * The blocks of dpg_skeleton_set_blocks. */
int dpg_skeleton_entities(const dpg_form_desc* form, const dpg_bucket_desc* bucket, int ncells,
                          const int32_t* cell_vertices, int32_t* entity);
/* Block-Jacobi preconditioning for dpg_skeleton_solve: the dofs of one entity
 * (host array entity (nglobal) with ids in [0, nentities); at most 64 dofs per
 * entity) form a block whose dense diagonal block of the matrix is inverted
 * (Cholesky) at the start of every solve. nentities = 0 returns to point
 * Jacobi. The reference's multigrid smoothers offer point and block Jacobi
 * (GMGOperator.cpp:2241-2262, Ifpack_PointRelaxation / Ifpack_BlockRelaxation
 * over a graph partition); here the blocks are the mesh entities. */
int dpg_skeleton_set_blocks(dpg_skeleton* sk, const int32_t* entity, int nentities, void* stream);
/* Solves the assembled system for its free dofs with the others prescribed:
 * rows of free dofs of A x = rhs, with x[d] given where fixed[d] = 1 (device
 * arrays of nglobal; x in/out, its free entries the initial guess). Jacobi-
 * (or, after dpg_skeleton_set_blocks, block-Jacobi-) preconditioned conjugate
 * gradients (the free block is symmetric positive definite) until
 * ||r|| <= tol ||r_0|| or maxit iterations, every reduction in a fixed order. Stands in for the global solve of TSolution::solve
 * (Solution.cpp:3480, Solver.h:111-118: Amesos direct or an iterative solver).
 * iters and relres (final ||r|| / ||r_0||) may be NULL. */
int dpg_skeleton_solve(const dpg_skeleton* sk, const double* values, const double* rhs, const int32_t* fixed,
                       double* x, double tol, int maxit, int* iters, double* relres, void* stream);
/* flags[c] = err[c] > rel_threshold * max err (device arrays of ncells): the
 * marking of TRefinementStrategy::getCellsAboveErrorThreshhold
 * (RefinementStrategy.cpp:249-257; ErrorIndicator.cpp:24-33), e.g. with
 * err = sqrt(eta2) from dpg_energy_error. Returns the number of marked cells. */
int dpg_mark_cells(const double* err, int ncells, double rel_threshold, int32_t* flags, void* stream);
int dpg_skeleton_destroy(dpg_skeleton* sk);

/* ------------------------------------------------------------------------
 * Condensed system of a 2D hp mesh: quads and triangles, hanging sides of a
 * 2:1 refinement, one element type (bucket) per cell. The minimum rule of
 * GDAMinimumRule::interpretLocalData (GDAMinimumRule.cpp:714-780) with the
 * side constraints of BasisReconciliation (BasisReconciliation.cpp:259):
 * every side of the mesh (a side of two cells, a boundary side, or the side of
 * a coarse cell whose neighbours are two fine cells) carries the smallest
 * order of its cells; H1 traces have a dof per regular vertex and order - 1 per
 * side, a hanging vertex takes the coarse side's value there; L2 fluxes have
 * `order` dofs per side. Each cell's trace and flux dofs are the global
 * functions at their nodes (local = M_c global); the cells' Schur blocks
 * enter as M_c^T S_c M_c.
 * ---------------------------------------------------------------------- */
typedef struct dpg_hp_system dpg_hp_system;
/* Host + device set-up once per mesh: cell_bucket (C) indexes buckets,
 * cell_vertices (C, 4) the global vertex ids in [0, nvertices) in each cell's
 * vertex order (triangles: 3, the fourth ignored), vertex_coords (nvertices, 2)
 * their coordinates (a hanging vertex is the midpoint of its coarse side). */
int dpg_hp_system_create(const dpg_form_desc* form, const dpg_bucket_desc* buckets, int nbuckets, int ncells,
                         const int32_t* cell_bucket, const int32_t* cell_vertices, int nvertices,
                         const double* vertex_coords, void* stream, dpg_hp_system** out);
/* Host only: the constrained numbering itself. sizes[4] receives the global
 * dof count, the total of the cells' global dof lists, the number of local
 * trace/flux rows (sum of T over the cells) and of (row, global dof, weight)
 * entries; with the arrays non-NULL (sizes from a first call): offsets (C + 1)
 * and dofs (the global dofs each cell touches), row_ptr (rows + 1), entry_dof
 * and entry_weight (entries) -- local = M global, rows in cell order then the
 * cell's trace/flux dof order. */
int dpg_hp_numbering(const dpg_form_desc* form, const dpg_bucket_desc* buckets, int nbuckets, int ncells,
                     const int32_t* cell_bucket, const int32_t* cell_vertices, int nvertices,
                     const double* vertex_coords, int64_t* sizes, int32_t* offsets, int32_t* dofs, int32_t* row_ptr,
                     int32_t* entry_dof, double* entry_weight);
/* global dofs, mesh entities (vertices and sides, the blocks for
 * dpg_skeleton_set_blocks) and the skeleton of the assembled matrix (owned by
 * the hp system; for dpg_skeleton_solve, _set_blocks, _pattern, _info). */
int dpg_hp_system_info(const dpg_hp_system* hs, int64_t* nglobal, int64_t* nentities, dpg_skeleton** sk);
/* Host arrays (nglobal each, any may be NULL): fixed = 1 on the trace dofs of
 * boundary sides; entity of every dof; a (cell, local trace/flux index) whose
 * local dof equals the global dof (-1 when none does). */
int dpg_hp_system_dofs(const dpg_hp_system* hs, int32_t* fixed, int32_t* entity, int32_t* rep_cell,
                       int32_t* rep_local);
/* values (nnz) and rhs (nglobal) from the buckets' Schur blocks: schur[b]
 * (C_b, T_b, T_b) and schur_rhs[b] (C_b, T_b) of the cells of bucket b in
 * increasing cell order (host arrays of device pointers). Stream-ordered. */
int dpg_hp_system_assemble(dpg_hp_system* hs, const double* const* schur, const double* const* schur_rhs,
                           double* values, double* rhs, void* stream);
/* the cells' local trace and flux coefficients from a global vector x: the
 * trailing T_b entries of every row of u[b] (C_b, M_b) are written (for
 * dpg_recover_fields). */
int dpg_hp_system_local(const dpg_hp_system* hs, const double* x, double* const* u, void* stream);
int dpg_hp_system_destroy(dpg_hp_system* hs);

/* Two-level (p-multigrid) preconditioning of the condensed system, the
 * multilevel solve the reference offers through GMGSolver (GMGSolver.h:23,
 * GMGOperator: a coarse space of lower polynomial order, block-Jacobi
 * smoothing, GMGOperator.cpp:2241-2262). The coarse level is the lowest-order
 * trace and flux space of the same mesh (H1 traces linear on every side: one
 * dof per vertex; L2 fluxes constant per side), numbered as
 * dpg_skeleton_numbering numbers H1 order 1; the prolongation P interpolates
 * the coarse functions at the fine nodes. Quads and hexes. Host and device
 * set-up, once per mesh; cell_vertices (C, V) on the host as for
 * dpg_skeleton_numbering. */
typedef struct dpg_coarse dpg_coarse;
int dpg_coarse_create(const dpg_form_desc* form, const dpg_bucket_desc* bucket, int ncells,
                      const int32_t* cell_vertices, void* stream, dpg_coarse** out);
/* coarse global dofs, nonzeros of the coarse matrix, coarse dofs per cell (any may be NULL) */
int dpg_coarse_info(const dpg_coarse* co, int64_t* nglobal_coarse, int64_t* nnz_coarse, int* T_coarse);
int dpg_coarse_destroy(dpg_coarse* co);
/* dpg_skeleton_solve with the two-level preconditioner: the coarse matrix
 * P^T A P of the free dofs is assembled from the cells' Schur blocks schur
 * (C, T, T), the matrix of values; flexible conjugate gradients, each
 * preconditioner application = damped block-Jacobi smoothing (the entity
 * blocks of dpg_skeleton_set_blocks, else point Jacobi; damping 1 / lambda_max
 * of the smoothed operator, estimated by power iteration at the start of the
 * call, which keeps the preconditioner positive definite), a coarse
 * correction by coarse_iters Jacobi-CG steps on the coarse matrix, and a
 * second smoothing step. Same arguments and stopping rule as
 * dpg_skeleton_solve. */
int dpg_skeleton_solve_twolevel(const dpg_skeleton* sk, const dpg_coarse* co, const double* schur,
                                const double* values, const double* rhs, const int32_t* fixed, double* x, double tol,
                                int maxit, int coarse_iters, int* iters, double* relres, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DPG_CAPI_H */

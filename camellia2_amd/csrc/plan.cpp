// plan.cpp -- C-ABI entry points of the device path: plan creation (device
// basis/cubature tables per element type, the compiled term program) and the
// batched local-assembly pipeline that replaces TBF::localStiffnessMatrixAndRHS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dpg_capi.h"
#include "element.hpp"
#include "errors.hpp"
#include "kernels.hpp"

namespace dpg {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

#define HIP_CHECK(expr)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) throw ::dpg::DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct DeviceBuffer {
  void* p = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t n) {
    if (n <= bytes) return;
    release();
    HIP_CHECK(hipMalloc(&p, n));
    bytes = n;
  }
};

template <class T>
T* upload(std::vector<std::unique_ptr<DeviceBuffer>>& owner, const std::vector<T>& host) {
  owner.emplace_back(new DeviceBuffer());
  size_t n = std::max<size_t>(1, host.size()) * sizeof(T);
  owner.back()->ensure(n);
  if (!host.empty()) HIP_CHECK(hipMemcpy(owner.back()->p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  return (T*)owner.back()->p;
}

DSummand summand(int var, int op, const dpg_coef& c, int aux) {
  DSummand s;
  s.var = var;
  s.op = op;
  s.slot = c.coef_slot;
  s.fn = c.coef_fn;
  s.normal = c.normal_comp;
  s.aux = aux;
  s.scale = c.scale;
  return s;
}

bool bterm_on_side(const dpg_form_desc& f, const dpg_bterm& t) {
  int kind = f.trial_kind[t.trial_var];
  return kind == DPG_TRACE || kind == DPG_FLUX || t.test_op == DPG_OP_DOT_NORMAL || t.coef.normal_comp >= 0;
}

void validate_form(const dpg_form_desc& f) {
  if (f.spatial_dim != 2 && f.spatial_dim != 3) throw std::invalid_argument("spatial_dim must be 2 or 3");
  if (f.n_bterms < 0 || f.n_bterms > DPG_MAX_BTERMS || f.n_ipterms <= 0 || f.n_ipterms > DPG_MAX_IPTERMS ||
      f.n_rhs < 0 || f.n_rhs > DPG_MAX_RHS_TERMS || f.n_coefs < 0 || f.n_coefs > DPG_MAX_COEFS)
    throw std::invalid_argument("term counts out of range");
  auto check_op = [&](int var, int op, bool side_ok) {
    if (var < 0 || var >= f.n_test) throw std::invalid_argument("test variable index out of range");
    bool hgrad = f.test_space[var] == DPG_HGRAD;
    bool ok = hgrad ? (op == DPG_OP_VALUE || (op >= DPG_OP_DX && op < DPG_OP_DX + f.spatial_dim))
                    : ((op >= DPG_OP_X && op < DPG_OP_X + f.spatial_dim) || op == DPG_OP_DIV ||
                       (side_ok && op == DPG_OP_DOT_NORMAL));
    if (!ok) throw std::invalid_argument("test operator does not apply to the test space");
  };
  auto check_coef = [&](const dpg_coef& c, bool side_ok) {
    if (c.coef_slot >= f.n_coefs || c.coef_slot < -1) throw std::invalid_argument("coefficient slot out of range");
    if (c.coef_fn < DPG_FN_VALUE || c.coef_fn > DPG_FN_INV) throw std::invalid_argument("unknown coefficient function");
    if (c.normal_comp >= f.spatial_dim || (!side_ok && c.normal_comp >= 0))
      throw std::invalid_argument("normal component out of range");
  };
  for (int i = 0; i < f.n_bterms; ++i) {
    const dpg_bterm& t = f.bterms[i];
    if (t.trial_var < 0 || t.trial_var >= f.n_trial) throw std::invalid_argument("trial variable index out of range");
    int kind = f.trial_kind[t.trial_var];
    int ncomp = kind == DPG_FIELD_VECTOR ? f.spatial_dim : 1;
    if (t.trial_comp < 0 || t.trial_comp >= ncomp) throw std::invalid_argument("trial component out of range");
    bool side = bterm_on_side(f, t);
    if (side && (kind == DPG_FIELD || kind == DPG_FIELD_VECTOR))
      throw std::invalid_argument("field variables on the mesh skeleton are not supported");
    check_op(t.test_var, t.test_op, side);
    check_coef(t.coef, side);
  }
  for (int i = 0; i < f.n_ipterms; ++i) {
    if (f.ipterms[i].n <= 0 || f.ipterms[i].n > DPG_MAX_LT_SUMMANDS) throw std::invalid_argument("bad IP term");
    for (int s = 0; s < f.ipterms[i].n; ++s) {
      check_op(f.ipterms[i].s[s].test_var, f.ipterms[i].s[s].test_op, false);
      check_coef(f.ipterms[i].s[s].coef, false);
    }
  }
  for (int i = 0; i < f.n_rhs; ++i) {
    check_op(f.rhs[i].test_var, f.rhs[i].test_op, false);
    if (f.rhs[i].load_comp < 0 || f.rhs[i].load_comp > 2) throw std::invalid_argument("bad load component");
  }
  // LDS of the kernels that stage per-point weights: volume groups (field
  // components), side groups (trace and flux variables), IP terms
  int n_vb = 0, n_sb = 0;
  for (int v = 0; v < f.n_trial; ++v) {
    const int kind = f.trial_kind[v];
    if (kind == DPG_FIELD) n_vb += 1;
    else if (kind == DPG_FIELD_VECTOR) n_vb += f.spatial_dim;
    else n_sb += 1;
  }
  const int dim = f.spatial_dim, nvar = f.n_test;
  if (operand_lds_bytes(std::max(n_vb, n_sb), nvar, dim) > kMaxLds || ip_operand_lds_bytes(f.n_ipterms, nvar, dim) > kMaxLds)
    throw std::invalid_argument("the form has too many test variables x terms for the operand kernels' LDS (64 KiB)");
}

}  // namespace

// Per-launch kernel timing with HIP events on the call's stream (only when
// the plan was created with options.profile = 1).
class Profiler {
 public:
  bool on = false;
  ~Profiler() {
    for (auto e : pool_) (void)hipEventDestroy(e);
    for (auto& r : pending_) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
  }
  hipEvent_t start(hipStream_t st) {
    if (!on) return nullptr;
    hipEvent_t e = get();
    HIP_CHECK(hipEventRecord(e, st));
    return e;
  }
  void stop(hipEvent_t a, hipStream_t st, int stage, double flops, double bytes) {
    if (!on) return;
    hipEvent_t b = get();
    HIP_CHECK(hipEventRecord(b, st));
    pending_.push_back({stage, a, b, flops, bytes});
  }
  // waits for the stream and folds the pending launches into the totals
  void collect(hipStream_t st, int64_t cells) {
    if (!on) return;
    HIP_CHECK(hipStreamSynchronize(st));
    for (auto& r : pending_) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      acc.ms[r.stage] += ms;
      acc.flops[r.stage] += r.flops;
      acc.bytes[r.stage] += r.bytes;
      acc.launches[r.stage] += 1;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    pending_.clear();
    acc.cells += cells;
  }
  dpg_timings acc{};

 private:
  struct Rec {
    int stage;
    hipEvent_t a, b;
    double flops, bytes;
  };
  hipEvent_t get() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  std::vector<hipEvent_t> pool_;
  std::vector<Rec> pending_;
};

struct BucketPlan {
  dpg_bucket_desc desc{};
  ElementLayout L;
  DevProgram prog{};
  DevProgram* prog_d = nullptr;
  DevTables tab{};
  std::vector<int> vb_block;  // volume group -> index in L.blocks
  std::vector<int> sb_var;    // side group -> trial variable
  int npr = 0, D = 0;
  int64_t lda = 0, Npad = 0;
  // per-cell element counts (doubles) of the workspace regions
  int64_t nA = 0, nAt = 0, nTv = 0, nTs = 0, nWip = 0, nWvol = 0, nWrhs = 0, nWside = 0, nS = 0, nLd = 0, nLi = 0;
  int64_t per_cell_bytes = 0;
  // adjoint only, after the regions above: stored factor, X = [G^-1 B | G^-1 l | G^-1 r], z
  int64_t nFac = 0, nX = 0, nZ = 0;
  int64_t per_cell_bytes_adj = 0;
  double flops = 0.0;
  double gram_flops = 0.0;  // sum over IP terms of QG * n(n+1), n = dofs the term touches
  int tile[3] = {64, 64, 64};  // FP64 GEMM tile edge per role: Gram, B blocks, Cholesky updates
  int QGp = 0;  // Gram points per IP term in the operand, padded to the k-chunk (16 rows; 32 for int8 digits)
  int NR = 0;         // int8 digit planes: rows (N rounded up to 64)
  int64_t nDig = 0;   // int8 digit planes: bytes per cell
  // per lower tile of G (64- and 32-wide tiles): bit lt set when IP term lt
  // reaches both the tile's rows and its columns
  const uint64_t* gmask[2] = {nullptr, nullptr};         // Gram k-segments per (64- or 32-) tile
  const uint64_t* gmask_merged[2] = {nullptr, nullptr};  // the same with equal terms merged ...
  const double* gweight[2] = {nullptr, nullptr};         // ... and the merged segments' weights
  // Exact zeros left of the B rows: a trial variable's row of B is +0 in every
  // cell at the columns of the test variables no bilinear term couples it with
  // (their point weights are written as zeros), and its solved row stays +0 up
  // to its first coupled column. zmask: per lower tile of the trailing block
  // (64- and 32-wide) the k-segments (zseg columns each) some product in the
  // tile needs.
  const uint64_t* zmask[2] = {nullptr, nullptr};
  double zkeep[2] = {1.0, 1.0};  // share of the trailing GEMM's tile x k work the masks keep
  int zseg = 16;
  // sum-factorised Gram (tensor-product cells, FP64): block pairs, pairs, 1D tables
  bool sf_ok = false;   // available for this bucket
  bool sf_on = false;   // used (dpg_plan_set_tiles, or the autotuner's accuracy-gated choice)
  bool merge_off = false;  // dense G without merged equal terms (the autotuner's second rounding of G)
  int sf_nbp = 0, sf_nF = 0, sf_q = 0;
  size_t sf_lds = 0;
  const SfBlockPair* sf_bps = nullptr;
  const SfPair* sf_pairs = nullptr;
  const double* sf_F = nullptr;
  // the volume part of B sum-factorised along with G (sf_on): block pairs
  // (trial field block, test block), 1D tables on the stiffness points
  bool sfb_ok = false;
  int sfb_nbp = 0, sfb_nF = 0, sfb_q = 0;
  size_t sfb_lds = 0;
  const SfBlockPair* sfb_bps = nullptr;
  const SfPair* sfb_pairs = nullptr;
  const double* sfb_F = nullptr;
  // ... and its side part (with sfb_ok): block pairs (trace or flux block on a
  // side, test block), per-side 1D tables of 5 * dim families
  bool sfs_ok = false;
  int sfs_nbp = 0, sfs_nF = 0, sfs_nfam = 0, sfs_q = 0;
  size_t sfs_lds = 0;
  const SfBlockPair* sfs_bps = nullptr;
  const SfPair* sfs_pairs = nullptr;
  const double* sfs_F = nullptr;
};

// ---- static scratch of dpg_assemble_noalloc ---------------------------------
// Workspace of the entry point that allocates nothing: a file-scope device
// array, one copy per GPU for the life of the process, shared by every call on
// that GPU. Host code never passes `scratch` itself to a kernel (on the host the
// name is only a stand-in); it uses the device address resolved per device by
// static_scratch().
static __device__ __align__(256) unsigned char scratch[DPG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, which must be the current device.
// Resolved once per device (hipGetSymbolAddress) and kept.
unsigned char* static_scratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw DeviceError("hipGetSymbolAddress returned NULL for the static scratch");
    g_scratch_ptr[device] = (unsigned char*)p;
  }
  return g_scratch_ptr[device];
}

}  // namespace dpg

struct dpg_plan {
  dpg_form_desc form{};
  dpg_options opt{};
  int device = 0;
  std::vector<dpg::BucketPlan> buckets;
  // device tables (bases, cubature, term programs, masks): read-only once built,
  // shared by the handles dpg_plan_clone_workspace makes
  std::shared_ptr<std::vector<std::unique_ptr<dpg::DeviceBuffer>>> tables =
      std::make_shared<std::vector<std::unique_ptr<dpg::DeviceBuffer>>>();
  dpg::DeviceBuffer ws;
  dpg::DeviceBuffer info;
  dpg::Profiler prof;
};

namespace dpg {
namespace {

void compile_program(const dpg_form_desc& f, BucketPlan& B) {
  DevProgram& P = B.prog;
  std::memset(&P, 0, sizeof(P));
  P.dim = f.spatial_dim;
  P.nverts = B.L.rc.nverts;
  P.nvar = f.n_test;
  P.npr = f.spatial_dim + 1;
  int ns = 0;
  auto push = [&](const DSummand& s) {
    if (ns >= kMaxSummands) throw std::invalid_argument("form has too many summands");
    P.s[ns++] = s;
  };
  // IP terms
  P.n_lt = f.n_ipterms;
  if (P.n_lt > kMaxGroups) throw std::invalid_argument("too many IP terms");
  for (int i = 0; i < f.n_ipterms; ++i) {
    P.lt_beg[i] = ns;
    for (int s = 0; s < f.ipterms[i].n; ++s)
      push(summand(f.ipterms[i].s[s].test_var, f.ipterms[i].s[s].test_op, f.ipterms[i].s[s].coef, 0));
  }
  P.lt_beg[P.n_lt] = ns;
  // volume trial blocks: field components
  B.vb_block.clear();
  for (int bi = 0; bi < (int)B.L.blocks.size(); ++bi)
    if (B.L.blocks[bi].side < 0) B.vb_block.push_back(bi);
  P.n_vb = (int)B.vb_block.size();
  if (P.n_vb > kMaxGroups) throw std::invalid_argument("too many field blocks");
  for (int g = 0; g < P.n_vb; ++g) {
    const TrialBlock& blk = B.L.blocks[B.vb_block[g]];
    P.vb_beg[g] = ns;
    for (int i = 0; i < f.n_bterms; ++i) {
      const dpg_bterm& t = f.bterms[i];
      if (t.trial_var != blk.var || t.trial_comp != blk.comp || bterm_on_side(f, t)) continue;
      push(summand(t.test_var, t.test_op, t.coef, 0));
    }
  }
  P.vb_beg[P.n_vb] = ns;
  P.nf = B.L.nf;
  P.ncoef = f.n_coefs;
  static_assert(kMaxTestVars == DPG_MAX_VARS, "test variable tables");
  for (int v = 0; v < f.n_test; ++v) P.test_off[v] = B.L.test_off[v];
  P.test_off[f.n_test] = B.L.N;
  for (int g = 0; g < P.n_vb; ++g) P.vb_off[g] = B.L.blocks[B.vb_block[g]].offset;
  // side groups: trace / flux variables
  B.sb_var.clear();
  for (int v = 0; v < f.n_trial; ++v)
    if (f.trial_kind[v] == DPG_TRACE || f.trial_kind[v] == DPG_FLUX) B.sb_var.push_back(v);
  P.n_sb = (int)B.sb_var.size();
  if (P.n_sb > kMaxGroups) throw std::invalid_argument("too many side variables");
  for (int g = 0; g < P.n_sb; ++g) {
    int v = B.sb_var[g];
    P.sb_flux[g] = f.trial_kind[v] == DPG_FLUX ? 1 : 0;
    P.sb_off[g] = -1;
    for (const TrialBlock& blk : B.L.blocks)
      if (blk.var == v && blk.side == 0) {
        P.sb_off[g] = blk.offset;
        P.sb_card[g] = blk.card;
      }
    P.sb_beg[g] = ns;
    for (int i = 0; i < f.n_bterms; ++i) {
      const dpg_bterm& t = f.bterms[i];
      if (t.trial_var != v) continue;
      push(summand(t.test_var, t.test_op, t.coef, 0));
    }
  }
  P.sb_beg[P.n_sb] = ns;
  P.rhs_beg[0] = ns;
  for (int i = 0; i < f.n_rhs; ++i) {
    dpg_coef c{};
    c.scale = f.rhs[i].scale;
    c.coef_slot = -1;
    c.normal_comp = -1;
    push(summand(f.rhs[i].test_var, f.rhs[i].test_op, c, f.rhs[i].load_comp));
  }
  P.rhs_beg[1] = ns;
}

// Tensor-product test blocks of a quad or hex bucket (HGRAD var; HDIV var
// component): 1D sizes per direction and, per prim, the 1D family per direction.
struct SfTestBlock {
  int var, off, n[3], nprim, prim[4], fam[4][3];
};
std::vector<SfTestBlock> sf_test_blocks(const dpg_form_desc& f, const ElementLayout& L) {
  const int dim = L.dim, npr = dim + 1, kt = L.kt;
  using Blk = SfTestBlock;
  std::vector<Blk> blocks;
  for (int v = 0; v < f.n_test; ++v) {
    if (f.test_space[v] == DPG_HGRAD) {
      Blk b{};
      b.var = v;
      b.off = L.test_off[v];
      for (int d = 0; d < 3; ++d) b.n[d] = d < dim ? kt + 1 : 1;
      b.nprim = npr;
      for (int r = 0; r < npr; ++r) {
        b.prim[r] = r;
        for (int d = 0; d < 3; ++d) b.fam[r][d] = (r >= 1 && d == r - 1) ? kSfFamD : kSfFamL;
      }
      blocks.push_back(b);
    } else {
      int off = L.test_off[v];
      for (int cpt = 0; cpt < dim; ++cpt) {
        Blk b{};
        b.var = v;
        b.off = off;
        int card = 1;
        for (int d = 0; d < 3; ++d) {
          b.n[d] = d < dim ? (d == cpt ? kt + 1 : kt) : 1;
          card *= b.n[d];
        }
        b.nprim = 2;
        b.prim[0] = cpt;  // reference component cpt
        b.prim[1] = dim;  // reference divergence
        for (int d = 0; d < 3; ++d) {
          b.fam[0][d] = d == cpt ? kSfFamL : kSfFamO;
          b.fam[1][d] = d == cpt ? kSfFamD : kSfFamO;
        }
        blocks.push_back(b);
        off += card;
      }
    }
  }
  return blocks;
}

// Tensor-product structure of the test functions for the sum-factorised Gram
// matrix: blocks (HGRAD var; HDIV var component), their prims' 1D families,
// and the (block, block) pairs that some IP term couples.
void build_sum_factorisation(dpg_plan* plan, BucketPlan& B) {
  const dpg_form_desc& f = plan->form;
  const ElementLayout& L = B.L;
  if (L.topo != DPG_QUAD && L.topo != DPG_HEX) return;
  const int dim = L.dim, npr = dim + 1, kt = L.kt;
  using Blk = SfTestBlock;
  const std::vector<Blk> blocks = sf_test_blocks(f, L);
  // prims each IP term reaches (op_to_prims: grad -> all derivative prims, HDIV value -> all components)
  const int na = f.n_test * npr;
  std::vector<std::vector<char>> touch(f.n_ipterms, std::vector<char>(na, 0));
  for (int i = 0; i < f.n_ipterms; ++i)
    for (int k = 0; k < f.ipterms[i].n; ++k) {
      int v = f.ipterms[i].s[k].test_var, op = f.ipterms[i].s[k].test_op;
      if (op == DPG_OP_VALUE) touch[i][v * npr] = 1;
      else if (op >= DPG_OP_DX && op < DPG_OP_DX + dim)
        for (int e = 1; e <= dim; ++e) touch[i][v * npr + e] = 1;
      else if (op >= DPG_OP_X && op < DPG_OP_X + dim)
        for (int e = 0; e < dim; ++e) touch[i][v * npr + e] = 1;
      else if (op == DPG_OP_DIV) touch[i][v * npr + dim] = 1;
    }
  auto coupled = [&](int a, int b) {
    for (int i = 0; i < f.n_ipterms; ++i)
      if (touch[i][a] && touch[i][b]) return true;
    return false;
  };
  std::vector<SfBlockPair> bps;
  std::vector<SfPair> pairs;
  for (size_t ia = 0; ia < blocks.size(); ++ia)
    for (size_t ib = 0; ib <= ia; ++ib) {
      const Blk &A = blocks[ia], &Bb = blocks[ib];
      SfBlockPair bp{};
      bp.offA = A.off;
      bp.offB = Bb.off;
      for (int d = 0; d < 3; ++d) {
        bp.nA[d] = A.n[d];
        bp.nB[d] = Bb.n[d];
      }
      bp.sB[0] = 1;
      bp.sB[1] = Bb.n[0];
      bp.sB[2] = Bb.n[0] * Bb.n[1];
      bp.diag = ia == ib;
      bp.pair0 = (int)pairs.size();
      for (int ra = 0; ra < A.nprim; ++ra)
        for (int rb = 0; rb < Bb.nprim; ++rb) {
          int a = A.var * npr + A.prim[ra], b = Bb.var * npr + Bb.prim[rb];
          if (!coupled(a, b)) continue;
          SfPair p{};
          p.a = (int16_t)a;
          p.b = (int16_t)b;
          for (int d = 0; d < 3; ++d) {
            p.fa[d] = (int8_t)A.fam[ra][d];
            p.fb[d] = (int8_t)Bb.fam[rb][d];
          }
          int cl = -1;
          for (int k = 0; k < bp.ncls; ++k)
            if (bp.clsfa[k] == p.fa[0] && bp.clsfb[k] == p.fb[0]) cl = k;
          if (cl < 0) {
            if (bp.ncls >= kSfMaxCls) throw std::logic_error("sum factorisation: too many classes");
            cl = bp.ncls++;
            bp.clsfa[cl] = p.fa[0];
            bp.clsfb[cl] = p.fb[0];
          }
          p.cls = (int8_t)cl;
          pairs.push_back(p);
        }
      bp.npairs = (int)pairs.size() - bp.pair0;
      bps.push_back(bp);  // without pairs too: it writes the block's zeros over the last call's factor
    }
  // 1D tables on the Gram points: L, L', O (O padded with a zero row)
  const int q = L.QG1, nF = kt + 1;
  Rule1D r = gauss_legendre(q);
  LagrangeLine Cl(gll_points(kt)), Op(gauss_legendre(kt).x);
  std::vector<double> F((size_t)3 * nF * q, 0.0);
  std::vector<double> v(nF), dv(nF), o(nF);
  for (int k = 0; k < q; ++k) {
    Cl.eval(r.x[k], v.data(), dv.data());
    Op.eval(r.x[k], o.data(), nullptr);
    for (int i = 0; i < nF; ++i) {
      F[((size_t)kSfFamL * nF + i) * q + k] = v[i];
      F[((size_t)kSfFamD * nF + i) * q + k] = dv[i];
      F[((size_t)kSfFamO * nF + i) * q + k] = i < kt ? o[i] : 0.0;
    }
  }
  size_t lds = 0;
  for (const SfBlockPair& bp : bps) {
    size_t t1 = dim == 3 ? (size_t)bp.npairs * q * q : (size_t)bp.ncls * q;
    size_t t2 = dim == 3 ? (size_t)bp.ncls * bp.nA[1] * bp.nB[1] * q : 0;
    lds = std::max(lds, sizeof(double) * ((size_t)bp.npairs * L.QG + 3 * (size_t)nF * q + t1 + t2));
  }
  if (bps.empty() || lds > 64 * 1024) return;  // the dense contraction stays in use
  auto& own = *plan->tables;
  B.sf_bps = upload(own, bps);
  B.sf_pairs = upload(own, pairs);
  B.sf_F = upload(own, F);
  B.sf_nbp = (int)bps.size();
  B.sf_nF = nF;
  B.sf_q = q;
  B.sf_lds = lds;
  B.sf_ok = true;
  B.sf_on = false;  // off until dpg_plan_set_tiles or dpg_plan_autotune switches it on (see dpg_tiles.sumfact)
}

// Reference prims (value, d/dxi_e; or components, divergence) a test operator
// reaches through the point's geometry (op_to_prims): bit r set for prim r.
unsigned prims_of_op(int op, int dim) {
  if (op == DPG_OP_VALUE) return 1u;
  if (op >= DPG_OP_DX && op < DPG_OP_DX + dim) return ((1u << dim) - 1u) << 1;
  if (op >= DPG_OP_X && op < DPG_OP_X + dim) return (1u << dim) - 1u;
  if (op == DPG_OP_DIV) return 1u << dim;
  if (op == DPG_OP_DOT_NORMAL) return (1u << dim) - 1u;
  return 0u;
}

// The volume part of B the same way: B[f, j] = sum_q sum_r W[q][g][var j][r]
// phi_f(q) R_r(j, q) for the trial field block g (a tensor-product scalar
// basis, one 1D family) and the test block of j, its prims r the group's
// summands reach. Taken together with the sum-factorised G (sf_on).
void build_sum_factorisation_b(dpg_plan* plan, BucketPlan& B) {
  const dpg_form_desc& f = plan->form;
  const ElementLayout& L = B.L;
  if (L.topo != DPG_QUAD && L.topo != DPG_HEX) return;
  const int dim = L.dim, npr = dim + 1, kt = L.kt, nvar = f.n_test;
  LagrangeLine Fld(gll_points(std::max(L.field_deg, 1)));
  const int nfl = Fld.size();
  if ((dim == 3 ? nfl * nfl * nfl : nfl * nfl) != L.nf) return;
  const std::vector<SfTestBlock> blocks = sf_test_blocks(f, L);
  const DevProgram& P = B.prog;
  std::vector<SfBlockPair> bps;
  std::vector<SfPair> pairs;
  for (int g = 0; g < P.n_vb; ++g) {
    const TrialBlock& blk = L.blocks[B.vb_block[g]];
    if (blk.card != L.nf) return;
    std::vector<unsigned> reach(nvar, 0u);  // prims of each test var the group's summands reach
    for (int si = P.vb_beg[g]; si < P.vb_beg[g + 1]; ++si) reach[P.s[si].var] |= prims_of_op(P.s[si].op, dim);
    for (const SfTestBlock& tb : blocks) {
      SfBlockPair bp{};
      bp.offA = L.N + blk.offset;
      bp.offB = tb.off;
      for (int d = 0; d < 3; ++d) {
        bp.nA[d] = d < dim ? nfl : 1;
        bp.nB[d] = tb.n[d];
      }
      bp.sB[0] = 1;
      bp.sB[1] = tb.n[0];
      bp.sB[2] = tb.n[0] * tb.n[1];
      bp.diag = 0;
      bp.pair0 = (int)pairs.size();
      for (int ri = 0; ri < tb.nprim; ++ri) {
        const int r = tb.prim[ri];
        if (!((reach[tb.var] >> r) & 1u)) continue;
        SfPair p{};
        p.a = p.b = (int16_t)((g * nvar + tb.var) * npr + r);
        for (int d = 0; d < 3; ++d) {
          p.fa[d] = (int8_t)kSfFamFld;
          p.fb[d] = (int8_t)tb.fam[ri][d];
        }
        int cl = -1;
        for (int k = 0; k < bp.ncls; ++k)
          if (bp.clsfa[k] == p.fa[0] && bp.clsfb[k] == p.fb[0]) cl = k;
        if (cl < 0) {
          if (bp.ncls >= kSfMaxCls) throw std::logic_error("sum factorisation: too many classes");
          cl = bp.ncls++;
          bp.clsfa[cl] = p.fa[0];
          bp.clsfb[cl] = p.fb[0];
        }
        p.cls = (int8_t)cl;
        pairs.push_back(p);
      }
      bp.npairs = (int)pairs.size() - bp.pair0;
      bps.push_back(bp);  // without pairs too: it writes the block's zeros (the dense GEMM writes every column)
    }
  }
  // 1D tables on the stiffness points: L, L', O (padded with a zero row), Fld
  const int q = L.Q1, nF = std::max(kt + 1, nfl);
  Rule1D r = gauss_legendre(q);
  LagrangeLine Cl(gll_points(kt)), Op(gauss_legendre(kt).x);
  std::vector<double> F((size_t)4 * nF * q, 0.0);
  std::vector<double> v(nF), dv(nF), o(nF), fl(nF);
  for (int k = 0; k < q; ++k) {
    Cl.eval(r.x[k], v.data(), dv.data());
    Op.eval(r.x[k], o.data(), nullptr);
    Fld.eval(r.x[k], fl.data(), nullptr);
    for (int i = 0; i < nF; ++i) {
      F[((size_t)kSfFamL * nF + i) * q + k] = i <= kt ? v[i] : 0.0;
      F[((size_t)kSfFamD * nF + i) * q + k] = i <= kt ? dv[i] : 0.0;
      F[((size_t)kSfFamO * nF + i) * q + k] = i < kt ? o[i] : 0.0;
      F[((size_t)kSfFamFld * nF + i) * q + k] = i < nfl ? fl[i] : 0.0;
    }
  }
  size_t lds = 0;
  for (const SfBlockPair& bp : bps) {
    size_t t1 = dim == 3 ? (size_t)bp.npairs * q * q : (size_t)bp.ncls * q;
    size_t t2 = dim == 3 ? (size_t)bp.ncls * bp.nA[1] * bp.nB[1] * q : 0;
    lds = std::max(lds, sizeof(double) * ((size_t)bp.npairs * L.Q + 4 * (size_t)nF * q + t1 + t2));
  }
  if (bps.empty() || lds > 64 * 1024) return;  // the dense contraction stays in use
  auto& own = *plan->tables;
  B.sfb_bps = upload(own, bps);
  B.sfb_pairs = upload(own, pairs);
  B.sfb_F = upload(own, F);
  B.sfb_nbp = (int)bps.size();
  B.sfb_nF = nF;
  B.sfb_q = q;
  B.sfb_lds = lds;
  B.sfb_ok = true;
}

// The side part of B the same way, one side at a time: B[t, j] = sum_q sum_r
// W_s[q][g][var j][r] phi_t(q) R_r(j, x_s(q)) for the trace or flux block of
// group g on side s (a tensor-product basis in the side's coordinates) and the
// test block of j restricted to the side. The kernel's directions are the
// side's coordinates (the reference directions they run along, with their
// signs, from the affine side map) and last the normal direction, whose one
// point is the side's reference coordinate; the test block's columns keep
// their reference order through per-direction strides.
void build_sum_factorisation_side(dpg_plan* plan, BucketPlan& B) {
  const dpg_form_desc& f = plan->form;
  const ElementLayout& L = B.L;
  if (!B.sfb_ok) return;
  const int dim = L.dim, npr = dim + 1, kt = L.kt, nvar = f.n_test, S = L.rc.nsides, q = L.Qs1;
  const DevProgram& P = B.prog;
  const int nfam = 5 * dim, na = P.n_sb * nvar * npr;
  // family of base b along kernel direction k; b: 0 L, 1 L', 2 O, 3 the trace
  // basis, 4 the flux basis (3 and 4 along the normal direction: the constant 1)
  auto id = [](int k, int base) { return k * 5 + base; };
  const std::vector<SfTestBlock> blocks = sf_test_blocks(f, L);
  LagrangeLine Cl(gll_points(kt)), Op(gauss_legendre(kt).x);
  std::unique_ptr<LagrangeLine> Tr, Fl;
  if (L.trace_deg >= 1) Tr.reset(new LagrangeLine(gll_points(L.trace_deg)));
  if (L.flux_deg >= 1) Fl.reset(new LagrangeLine(gll_points(L.flux_deg)));
  int nF = kt + 1;
  if (Tr) nF = std::max(nF, Tr->size());
  if (Fl) nF = std::max(nF, Fl->size());
  // side geometry: reference direction and sign of each side coordinate, the normal direction and its coordinate
  std::vector<std::array<int, 3>> dirs(S);
  std::vector<std::array<double, 3>> sign(S);
  for (int sd = 0; sd < S; ++sd) {
    const double* p = &L.rc.side_param[(size_t)sd * 3 * dim];
    int used = 0;
    for (int a = 0; a < dim - 1; ++a) {
      int dd = -1;
      for (int d = 0; d < dim; ++d)
        if (p[(1 + a) * dim + d] != 0.0) {
          if (dd >= 0 || std::fabs(p[(1 + a) * dim + d]) != 1.0 || p[d] != 0.0) return;  // not axis-aligned
          dd = d;
        }
      if (dd < 0) return;
      dirs[sd][a] = dd;
      sign[sd][a] = p[(1 + a) * dim + dd];
      used |= 1 << dd;
    }
    for (int d = 0; d < dim; ++d)
      if (!((used >> d) & 1)) {
        if (std::fabs(p[d]) != 1.0) return;
        dirs[sd][dim - 1] = d;
        sign[sd][dim - 1] = p[d];  // the normal direction's coordinate
      }
  }
  // per side: nfam * nF * q table entries
  Rule1D r = gauss_legendre(q);
  std::vector<double> F((size_t)S * nfam * nF * q, 0.0);
  std::vector<double> v(nF), dv(nF), o(nF), tb1(nF);
  for (int sd = 0; sd < S; ++sd) {
    double* Fs = &F[(size_t)sd * nfam * nF * q];
    for (int k = 0; k < dim; ++k) {
      const bool normal = k == dim - 1;
      for (int j = 0; j < (normal ? 1 : q); ++j) {
        const double xi = normal ? sign[sd][k] : sign[sd][k] * r.x[j];
        Cl.eval(xi, v.data(), dv.data());
        Op.eval(xi, o.data(), nullptr);
        for (int i = 0; i < nF; ++i) {
          Fs[((size_t)id(k, 0) * nF + i) * q + j] = i <= kt ? v[i] : 0.0;
          Fs[((size_t)id(k, 1) * nF + i) * q + j] = i <= kt ? dv[i] : 0.0;
          Fs[((size_t)id(k, 2) * nF + i) * q + j] = i < kt ? o[i] : 0.0;
        }
        for (int b = 3; b < 5; ++b) {
          const LagrangeLine* T1d = b == 3 ? Tr.get() : Fl.get();
          if (normal) {
            Fs[((size_t)id(k, b) * nF + 0) * q + j] = 1.0;
          } else if (T1d) {  // the trial bases along the side coordinate itself
            T1d->eval(r.x[j], tb1.data(), nullptr);
            for (int i = 0; i < T1d->size(); ++i) Fs[((size_t)id(k, b) * nF + i) * q + j] = tb1[i];
          }
        }
      }
    }
  }
  std::vector<SfBlockPair> bps;
  std::vector<SfPair> pairs;
  for (int g = 0; g < P.n_sb; ++g) {
    std::vector<unsigned> reach(nvar, 0u);
    for (int si = P.sb_beg[g]; si < P.sb_beg[g + 1]; ++si) reach[P.s[si].var] |= prims_of_op(P.s[si].op, dim);
    for (const TrialBlock& blk : L.blocks) {
      if (blk.var != B.sb_var[g] || blk.side < 0) continue;
      const LagrangeLine* T1d = blk.kind == DPG_TRACE ? Tr.get() : Fl.get();
      if (!T1d) return;
      const int n1 = T1d->size();
      if ((dim == 3 ? n1 * n1 : n1) != blk.card) return;
      const int sd = blk.side, tbase = blk.kind == DPG_TRACE ? 3 : 4;
      for (const SfTestBlock& tb : blocks) {
        SfBlockPair bp{};
        bp.offA = L.N + blk.offset;
        bp.offB = tb.off;
        int stride[3] = {1, tb.n[0], tb.n[0] * tb.n[1]};
        for (int k = 0; k < 3; ++k) {
          bp.nA[k] = k < dim - 1 ? n1 : 1;
          bp.nB[k] = k < dim ? tb.n[dirs[sd][k]] : 1;
          bp.sB[k] = k < dim ? stride[dirs[sd][k]] : 0;
        }
        bp.foff = sd * nfam * nF * q;
        bp.woff = sd * L.Qs * na;
        bp.diag = 0;
        bp.pair0 = (int)pairs.size();
        for (int ri = 0; ri < tb.nprim; ++ri) {
          const int pr = tb.prim[ri];
          if (!((reach[tb.var] >> pr) & 1u)) continue;
          SfPair p{};
          p.a = p.b = (int16_t)((g * nvar + tb.var) * npr + pr);
          for (int k = 0; k < 3; ++k) {
            p.fa[k] = (int8_t)(k < dim ? id(k, tbase) : 0);
            p.fb[k] = (int8_t)(k < dim ? id(k, tb.fam[ri][dirs[sd][k]]) : 0);
          }
          int cl = -1;
          for (int c2 = 0; c2 < bp.ncls; ++c2)
            if (bp.clsfa[c2] == p.fa[0] && bp.clsfb[c2] == p.fb[0]) cl = c2;
          if (cl < 0) {
            if (bp.ncls >= kSfMaxCls) throw std::logic_error("sum factorisation: too many classes");
            cl = bp.ncls++;
            bp.clsfa[cl] = p.fa[0];
            bp.clsfb[cl] = p.fb[0];
          }
          p.cls = (int8_t)cl;
          pairs.push_back(p);
        }
        bp.npairs = (int)pairs.size() - bp.pair0;
        bps.push_back(bp);  // without pairs too: it writes the block's zeros
      }
    }
  }
  const int QP = dim == 3 ? q * q * q : q * q;  // the kernel's point slots per pair
  size_t lds = 0;
  for (const SfBlockPair& bp : bps) {
    size_t t1 = dim == 3 ? (size_t)bp.npairs * q * q : (size_t)bp.ncls * q;
    size_t t2 = dim == 3 ? (size_t)bp.ncls * bp.nA[1] * bp.nB[1] * q : 0;
    lds = std::max(lds, sizeof(double) * ((size_t)bp.npairs * QP + (size_t)nfam * nF * q + t1 + t2));
  }
  if (bps.empty() || lds > 64 * 1024) return;
  auto& own = *plan->tables;
  B.sfs_bps = upload(own, bps);
  B.sfs_pairs = upload(own, pairs);
  B.sfs_F = upload(own, F);
  B.sfs_nbp = (int)bps.size();
  B.sfs_nF = nF;
  B.sfs_nfam = nfam;
  B.sfs_q = q;
  B.sfs_lds = lds;
  B.sfs_ok = true;
}

void build_bucket(dpg_plan* plan, const dpg_bucket_desc& desc, BucketPlan& B) {
  const dpg_form_desc& f = plan->form;
  B.desc = desc;
  B.L = make_layout(f, desc);
  const ElementLayout& L = B.L;
  B.npr = L.dim + 1;
  B.flops = flops_per_cell(f, L);
  B.gram_flops = 0.0;
  for (int i = 0; i < f.n_ipterms; ++i) {
    bool used[DPG_MAX_VARS] = {false};
    for (int k = 0; k < f.ipterms[i].n; ++k) used[f.ipterms[i].s[k].test_var] = true;
    double n = 0;
    for (int v = 0; v < f.n_test; ++v)
      if (used[v]) n += L.test_card[v];
    B.gram_flops += (double)L.QG * n * (n + 1.0);
  }
  compile_program(f, B);
  build_sum_factorisation(plan, B);
  build_sum_factorisation_b(plan, B);
  build_sum_factorisation_side(plan, B);
  RefTables T = make_tables(f, L);
  auto& own = *plan->tables;
  B.tab.R_vol = upload(own, T.vol.R);
  B.tab.R_ip = upload(own, T.ip.R);
  B.tab.R_side = upload(own, T.side.R);
  B.tab.dN_vol = upload(own, T.vol.dN);
  B.tab.dN_ip = upload(own, T.ip.dN);
  B.tab.dN_side = upload(own, T.side.dN);
  B.tab.Nv_vol = upload(own, T.vol.Nv);
  B.tab.w_vol = upload(own, T.vol.w);
  B.tab.w_ip = upload(own, T.ip.w);
  B.tab.w_side = upload(own, T.side.w);
  B.tab.field_vol = upload(own, T.vol.field);
  B.tab.trace = upload(own, T.trace);
  B.tab.flux = upload(own, T.flux);
  std::vector<double> tan((size_t)L.rc.nsides * 2 * L.dim, 0.0);
  for (int s = 0; s < L.rc.nsides; ++s)
    for (int a = 0; a < L.dim - 1; ++a)
      for (int d = 0; d < L.dim; ++d)
        tan[((size_t)s * 2 + a) * L.dim + d] = L.rc.side_param[((size_t)s * 3 + 1 + a) * L.dim + d];
  B.tab.side_tan = upload(own, tan);
  std::vector<int32_t> vod(L.var_of_test_dof.begin(), L.var_of_test_dof.end());
  B.tab.var_of_dof = upload(own, vod);
  std::vector<DevProgram> pv(1, B.prog);
  B.prog_d = upload(own, pv);

  // workspace layout (doubles per cell)
  B.D = L.N + L.M + 2;
  B.lda = round_up(B.D, 8);
  B.Npad = round_up(L.N, 8);
  int S = L.rc.nsides, nvar = f.n_test;
  B.nA = B.lda * B.D;
  const bool i8 = plan->opt.precision == DPG_PREC_INT8SPLIT;
  B.QGp = (int)round_up(L.QG, i8 ? 32 : 16);
  B.nAt = (int64_t)B.prog.n_lt * B.QGp * B.Npad;  // weighted IP-term values, term-major rows
  if (i8) {
    // the same region holds the operand's int8 digit planes instead
    B.NR = (int)round_up(L.N, 64);
    B.nDig = (int64_t)kI8Digits * B.prog.n_lt * B.QGp * B.NR;
    B.nAt = std::max(B.nAt, (B.nDig + 7) / 8);
  }
  {
    // test variables each IP term reaches
    std::vector<uint64_t> tv(f.n_test, 0);
    for (int i = 0; i < f.n_ipterms; ++i)
      for (int k = 0; k < f.ipterms[i].n; ++k) tv[f.ipterms[i].s[k].test_var] |= 1ull << i;
    // IP terms a and b give the same operand values on the dofs of the test
    // variables in `vars` when their summands on those variables are the same
    // list (operator and coefficient, in order)
    auto same_on = [&](int a, int b, uint64_t vars) {
      auto restricted = [&](int t) {
        std::vector<const dpg_lt_summand*> r;
        for (int k = 0; k < f.ipterms[t].n; ++k)
          if ((vars >> f.ipterms[t].s[k].test_var) & 1ull) r.push_back(&f.ipterms[t].s[k]);
        return r;
      };
      const auto ra = restricted(a), rb = restricted(b);
      if (ra.size() != rb.size()) return false;
      for (size_t k = 0; k < ra.size(); ++k) {
        const dpg_lt_summand &x = *ra[k], &y = *rb[k];
        if (x.test_var != y.test_var || x.test_op != y.test_op || x.coef.scale != y.coef.scale ||
            x.coef.coef_slot != y.coef.coef_slot || x.coef.coef_fn != y.coef.coef_fn ||
            x.coef.normal_comp != y.coef.normal_comp)
          return false;
      }
      return true;
    };
    const int nlt = B.prog.n_lt;
    // the merged segments below are whole IP terms: one term per segment, every
    // segment a whole number of the GEMM's 16-row k-chunks
    if (nlt != f.n_ipterms || B.QGp % 16 != 0)
      throw std::logic_error("Gram operand segments must be whole IP terms of whole k-chunks");
    for (int ts = 0; ts < 2; ++ts) {
      const int T = ts == 0 ? 64 : 32, nt = (L.N + T - 1) / T;
      std::vector<uint64_t> reach(nt, 0), vars(nt, 0), mask, mmask;  // terms reaching some dof of a T-block
      std::vector<double> weight;
      for (int j = 0; j < L.N; ++j) {
        reach[j / T] |= tv[L.var_of_test_dof[j]];
        vars[j / T] |= 1ull << L.var_of_test_dof[j];
      }
      for (int ti = 0; ti < nt; ++ti)
        for (int tj = 0; tj <= ti; ++tj) {
          const uint64_t m = reach[ti] & reach[tj], u = vars[ti] | vars[tj];
          mask.push_back(m);
          // merged: terms equal on the variables of both blocks share one segment,
          // whose staged rows are weighted by their count (the graph norm's
          // (tau - grad v)_d and tau_d terms on tau-tau blocks, for instance)
          uint64_t mm = m;
          std::vector<double> w(nlt, 1.0);
          for (int a = 0; a < nlt; ++a) {
            if (!((mm >> a) & 1ull)) continue;
            for (int b = a + 1; b < nlt; ++b)
              if (((mm >> b) & 1ull) && same_on(a, b, u)) {
                mm &= ~(1ull << b);
                w[a] += 1.0;
              }
          }
          mmask.push_back(mm);
          weight.insert(weight.end(), w.begin(), w.end());
        }
      B.gmask[ts] = upload(*plan->tables, mask);
      B.gmask_merged[ts] = upload(*plan->tables, mmask);
      B.gweight[ts] = upload(*plan->tables, weight);
    }
  }
  {
    // first column each augmented row can be non-zero in (rounded down to the
    // 16-column k-chunk); the load and residual rows are taken as dense
    const int N = L.N, D = B.D;
    std::vector<int> z(D, 0);
    for (const TrialBlock& blk : L.blocks) {
      int zc = N;
      for (int i = 0; i < f.n_bterms; ++i)
        if (f.bterms[i].trial_var == blk.var) zc = std::min(zc, L.test_off[f.bterms[i].test_var]);
      zc = zc / 16 * 16;
      for (int r = 0; r < blk.card; ++r) z[N + blk.offset + r] = zc;
    }
    B.zseg = 16 * (int)((N + 16 * 64 - 1) / (16 * 64));  // at most 64 segments
    const int n = D - N, nseg = (N + B.zseg - 1) / B.zseg;
    for (int ts = 0; ts < 2; ++ts) {
      const int T = ts == 0 ? 64 : 32, nt = (n + T - 1) / T;
      std::vector<int> zt(nt, N);  // smallest z over each T-block of the trailing rows
      for (int r = 0; r < n; ++r) zt[r / T] = std::min(zt[r / T], z[N + r]);
      std::vector<uint64_t> mask;
      double kept = 0.0;
      for (int ti = 0; ti < nt; ++ti)
        for (int tj = 0; tj <= ti; ++tj) {
          const int k0 = std::max(zt[ti], zt[tj]);  // products left of it are 0 x (something)
          uint64_t m = 0;
          for (int sg = 0; sg < nseg; ++sg)
            if ((sg + 1) * B.zseg > k0) {
              m |= 1ull << sg;
              kept += std::min(N, (sg + 1) * B.zseg) - sg * B.zseg;
            }
          mask.push_back(m);
        }
      B.zmask[ts] = upload(*plan->tables, mask);
      B.zkeep[ts] = kept / ((double)mask.size() * N);
    }
  }
  B.nTv = (int64_t)B.prog.n_vb * L.Q * B.Npad;
  B.nTs = (int64_t)B.prog.n_sb * S * L.Qs * B.Npad;
  B.nWip = (int64_t)L.QG * B.prog.n_lt * nvar * B.npr;
  B.nWvol = (int64_t)L.Q * B.prog.n_vb * nvar * B.npr;
  B.nWrhs = (int64_t)L.Q * nvar * B.npr;
  B.nWside = (int64_t)S * L.Qs * B.prog.n_sb * nvar * B.npr;
  B.nS = round_up(std::max(L.N, L.M), 8);
  B.nLd = (int64_t)round_up(L.N + L.F, 64) * 64;  // factored diagonal blocks of both stages
  B.nLi = 64 * 64;                                 // current diagonal block, column-major
  int64_t tot = B.nA + B.nAt + B.nTv + B.nTs + B.nWip + B.nWvol + B.nWrhs + B.nWside + B.nS + B.nLd + B.nLi;
  B.per_cell_bytes = tot * (int64_t)sizeof(double);
  B.nFac = round_up((int64_t)(L.N + L.M + 1) * L.N, 8);
  B.nX = round_up((int64_t)L.N * (L.M + 2), 8);
  B.nZ = round_up(L.N, 8);
  B.per_cell_bytes_adj = B.per_cell_bytes + (B.nFac + B.nX + B.nZ) * (int64_t)sizeof(double);
}

struct Work {
  double *A, *At, *Tv, *Ts, *Wip, *Wvol, *Wrhs, *Wside, *s, *Ld, *Li;
  double *Fac = nullptr, *X = nullptr, *Z = nullptr;  // adjoint regions (carve with adj)
  int32_t* info;  // per-cell LAPACK INFO of the batch
};

// per-cell regions of a batch of nb cells, carved from `base` (with adj, also
// the adjoint's regions: B.per_cell_bytes_adj per cell)
Work carve(void* base, int32_t* info, const BucketPlan& B, int nb, bool adj = false) {
  Work w;
  w.info = info;
  double* p = (double*)base;
  auto take = [&](int64_t per) {
    double* r = p;
    p += per * nb;
    return r;
  };
  w.A = take(B.nA);
  w.At = take(B.nAt);
  w.Tv = take(B.nTv);
  w.Ts = take(B.nTs);
  w.Wip = take(B.nWip);
  w.Wvol = take(B.nWvol);
  w.Wrhs = take(B.nWrhs);
  w.Wside = take(B.nWside);
  w.s = take(B.nS);
  w.Ld = take(B.nLd);
  w.Li = take(B.nLi);
  if (adj) {
    w.Fac = take(B.nFac);
    w.X = take(B.nX);
    w.Z = take(B.nZ);
  }
  return w;
}

int batch_cells(dpg_plan* plan, const BucketPlan& B, int ncells, int64_t per_cell_bytes = 0) {
  if (plan->opt.max_batch_cells > 0) return std::min(std::min(ncells, plan->opt.max_batch_cells), 65535);
  double budget = plan->opt.workspace_bytes > 0 ? plan->opt.workspace_bytes : 16.0 * (1ull << 30);
  int64_t nb = (int64_t)(budget / (double)(per_cell_bytes > 0 ? per_cell_bytes : B.per_cell_bytes));
  nb = std::max<int64_t>(1, std::min<int64_t>(nb, ncells));
  return (int)std::min<int64_t>(nb, 65535);  // cells are a grid y/z dimension
}

// Left-looking partial Cholesky of columns [kbeg, kend) of the augmented
// matrices of a batch (columns before kbeg already hold their final values,
// columns [kbeg, kend) have not been touched by the factorisation yet):
//   panel update  A[k0:D, k0:k1] -= A[k0:D, kbeg:k0] A[k0:k1, kbeg:k0]^T   (MFMA)
//   panel         POTRF of the diagonal block, X <- X L^-T below it
// then the trailing block A[kend:D, kend:D] -= A[kend:D, kbeg:kend] (...)^T.
// zeros: the matrix is a bucket's augmented [G; B; l; r] over the columns of G
// (kbeg = 0, kend = N): the closing product of the B rows skips the k-segments
// where the B rows' leading exact zeros leave every product of a tile zero
// (zmask). (Splitting the panel updates at the zero runs as well was measured
// slower: the extra, partly empty row tiles cost more than the skipped work.)
void factor_columns(dpg_plan* plan, const BucketPlan& B, int nb, const Work& w, int kbeg, int kend, double beta_trailing,
                    hipStream_t st, bool zeros = false) {
  Profiler& P = plan->prof;
  const int64_t sA = B.nA, lda = B.lda;
  int32_t* info = w.info;
  // update tile 128: the rows below each diagonal block are updated and solved
  // in one kernel (update_solve_f64) after the block's own update and POTRF
  const bool fused = B.tile[2] == 128;
  for (int k0 = kbeg; k0 < kend; k0 += 64) {
    int kb = std::min(64, kend - k0);
    double rows = B.D - k0;
    hipEvent_t e;
    if (fused) {
      const double kk = k0 - kbeg, below = rows - kb;
      if (k0 > kbeg) {  // the diagonal block's update (its lower part)
        GemmArgs t{};
        t.X = w.A + k0 + (size_t)kbeg * lda;
        t.ldx = lda;
        t.sx = sA;
        t.Y = t.X;
        t.ldy = lda;
        t.sy = sA;
        t.C = w.A + k0 + (size_t)k0 * lda;
        t.ldc = lda;
        t.sc = sA;
        t.m = kb;
        t.n = kb;
        t.k = k0 - kbeg;
        t.alpha = -1.0;
        t.beta = 1.0;
        t.ncells = nb;
        t.skip_upper = 1;
        e = P.start(st);
        launch_gemm(t, kGemmUpdate, st, 64);
        P.stop(e, st, DPG_STAGE_UPDATE, nb * 2.0 * (kb * (kb + 1) / 2.0) * kk, nb * (kb * kk + kb * kb * 2.0) * 8.0);
      }
      e = P.start(st);
      launch_potrf(w.A, lda, sA, k0, kb, w.Ld, B.nLd, w.Li, nb, info, 0, st);
      P.stop(e, st, DPG_STAGE_PANEL, nb * (double)kb * kb * kb / 3.0, nb * (double)kb * kb * 3.0 * 8.0);
      if (below > 0) {
        GemmArgs t{};
        t.X = w.A + k0 + kb + (size_t)kbeg * lda;
        t.ldx = lda;
        t.sx = sA;
        t.Y = w.A + k0 + (size_t)kbeg * lda;
        t.ldy = lda;
        t.sy = sA;
        t.C = w.A + k0 + kb + (size_t)k0 * lda;
        t.ldc = lda;
        t.sc = sA;
        t.m = (int)below;
        t.n = kb;
        t.k = k0 - kbeg;
        t.alpha = -1.0;
        t.beta = 1.0;
        t.ncells = nb;
        t.Lt = w.Li;
        e = P.start(st);
        launch_update_solve(t, st);
        P.stop(e, st, DPG_STAGE_SOLVE, nb * (2.0 * below * kb * kk + below * (double)kb * kb),
               nb * (below * kk + kb * kk + below * kb * 2.0 + kb * 64.0) * 8.0);
      }
      continue;
    }
    if (k0 > kbeg) {
      GemmArgs t{};
      t.X = w.A + k0 + (size_t)kbeg * lda;
      t.ldx = lda;
      t.sx = sA;
      t.Y = t.X;
      t.ldy = lda;
      t.sy = sA;
      t.C = w.A + k0 + (size_t)k0 * lda;
      t.ldc = lda;
      t.sc = sA;
      t.m = B.D - k0;
      t.n = kb;
      t.k = k0 - kbeg;
      t.alpha = -1.0;
      t.beta = 1.0;
      t.ncells = nb;
      t.skip_upper = 1;  // the panel's diagonal block: its factorisation reads the lower part only
      double kk = k0 - kbeg;
      e = P.start(st);
      launch_gemm(t, kGemmUpdate, st, B.tile[2]);
      P.stop(e, st, DPG_STAGE_UPDATE, nb * 2.0 * (rows * kb - kb * (kb - 1) / 2.0) * kk,
             nb * (rows * kk + rows * kb * 2.0) * 8.0);
    }
    e = P.start(st);
    launch_panel(w.A, lda, sA, B.D, k0, kb, w.Ld, B.nLd, w.Li, nb, info, 0, st);
    P.stop(e, st, DPG_STAGE_PANEL, nb * ((double)kb * kb * kb / 3.0 + (rows - kb) * kb * kb),
           nb * rows * kb * 16.0);
  }
  int n = B.D - kend;
  if (n > 0) {
    GemmArgs t{};
    t.X = w.A + kend + (size_t)kbeg * lda;
    t.ldx = lda;
    t.sx = sA;
    t.Y = t.X;
    t.ldy = lda;
    t.sy = sA;
    t.C = w.A + kend + (size_t)kend * lda;
    t.ldc = lda;
    t.sc = sA;
    t.m = t.n = n;
    t.k = kend - kbeg;
    t.lower = 1;
    t.alpha = -1.0;
    t.beta = beta_trailing;
    t.ncells = nb;
    if (zeros) {  // the B rows' leading zeros: per tile, the k-segments some product needs
      t.kmask = B.zmask[B.tile[2] == 32 ? 1 : 0];
      t.kseg = B.zseg;
    }
    hipEvent_t e = P.start(st);
    launch_gemm(t, kGemmUpdate, st, B.tile[2]);
    P.stop(e, st, DPG_STAGE_UPDATE, nb * (double)n * (n + 1.0) * t.k * (zeros ? B.zkeep[B.tile[2] == 32 ? 1 : 0] : 1.0),
           nb * ((double)n * t.k + n * (n + 1.0) / 2.0 * (beta_trailing != 0.0 ? 2.0 : 1.0)) * 8.0);
  }
}

// Stage 1 for cells [c0, c0+nb): contractions into A, DPOEQU scaling and the
// partial Cholesky over the first N columns.
void stage_factor(dpg_plan* plan, const BucketPlan& B, const dpg_batch& bt, int c0, int nb, const Work& w,
                  const double* u, double* Gout, hipStream_t st) {
  Profiler& P = plan->prof;
  const ElementLayout& L = B.L;
  const int dim = L.dim, V = L.rc.nverts, S = L.rc.nsides, N = L.N, M = L.M, nvar = plan->form.n_test;
  const double* nodes = bt.nodes + (size_t)c0 * V * dim;
  const int8_t* par = bt.parities + (size_t)c0 * S;
  const double* coefs = bt.coefs ? bt.coefs + (size_t)c0 * plan->form.n_coefs : nullptr;
  const int64_t sA = B.nA, lda = B.lda;
  const double D8 = 8.0;
  // per-point weights
  hipEvent_t e = P.start(st);
  launch_point(dim, 0, B.prog_d, B.tab, L.QG, L.Qs, nb, nodes, par, coefs, plan->form.n_coefs, S, bt.load, w.Wip,
               nullptr, st);
  launch_point(dim, 1, B.prog_d, B.tab, L.Q, L.Qs, nb, nodes, par, coefs, plan->form.n_coefs, S, bt.load, w.Wvol,
               w.Wrhs, st);
  if (B.prog.n_sb > 0)
    launch_point(dim, 2, B.prog_d, B.tab, S * L.Qs, L.Qs, nb, nodes, par, coefs, plan->form.n_coefs, S, bt.load,
                 w.Wside, nullptr, st);
  P.stop(e, st, DPG_STAGE_POINT, 0.0, nb * D8 * (B.nWip + B.nWvol + B.nWrhs + B.nWside + V * dim));
  const int prec = plan->opt.precision;
  const bool sf = B.sf_on && prec == DPG_PREC_FP64;
  const bool sfb = sf && B.sfb_ok;  // the volume part of B sum-factorised too: no volume B operand
  const bool sfs = sfb && B.sfs_ok;  // and its side part: no side B operand
  // operands: weighted IP-term values (with the DPOEQU factors), B operands
  e = P.start(st);
  const int KG = B.prog.n_lt * B.QGp;
  const bool i8 = prec == DPG_PREC_INT8SPLIT;
  if (i8 && Gout) throw std::invalid_argument("G is an output of the FP64 and FP32-class precisions only");
  if (i8)
    launch_ip_digits(dim, w.Wip, B.prog.n_lt, nvar, B.tab.R_ip, B.tab.var_of_dof, N, L.QG, B.QGp, B.NR, nb,
                     (int8_t*)w.At, B.nAt * (int64_t)sizeof(double), w.s, B.nS, st);
  else if (!sf)
    launch_ip_operand(dim, w.Wip, B.prog.n_lt, nvar, B.tab.R_ip, B.tab.var_of_dof, N, L.QG, B.QGp, nb, w.At, B.nAt,
                      B.Npad, w.s, B.nS, st);
  if (!sfb)
    launch_operand(dim, w.Wvol, B.prog.n_vb, nvar, B.tab.R_vol, B.tab.var_of_dof, N, L.Q, nb, w.Tv, B.nTv,
                   (int64_t)L.Q * B.Npad, B.Npad, st);
  if (B.prog.n_sb > 0 && !sfs)
    launch_operand(dim, w.Wside, B.prog.n_sb, nvar, B.tab.R_side, B.tab.var_of_dof, N, S * L.Qs, nb, w.Ts, B.nTs,
                   (int64_t)S * L.Qs * B.Npad, B.Npad, st);
  double opnd = (sf ? 0.0 : (double)KG) + (sfb ? 0.0 : (double)B.prog.n_vb * L.Q) +
                (sfs ? 0.0 : (double)B.prog.n_sb * S * L.Qs);
  P.stop(e, st, DPG_STAGE_OPERAND, nb * opnd * N * 2.0 * B.npr, nb * opnd * N * D8);
  if (sf && !Gout) {
    // G's diagonal first (the same sums as the full pass), the DPOEQU factors
    // from it, then the full sum-factorised G written as S G S (the products of
    // the scaling pass, in its order); B and l are scaled in their epilogues
    e = P.start(st);
    launch_gram_sf(dim, B.sf_bps, B.sf_nbp, B.sf_pairs, B.sf_F, B.sf_nF, B.sf_q, w.Wip, B.prog.n_lt,
                   nvar * B.npr, L.QG, w.A, lda, sA, nb, B.sf_lds, st, nullptr, 0, true);
    P.stop(e, st, DPG_STAGE_GRAM, 0.0, nb * (double)B.nWip * D8);
    e = P.start(st);
    launch_scale_factors(w.A, lda, sA, 0, N, w.s, B.nS, nb, st);
    P.stop(e, st, DPG_STAGE_SCALE, 0.0, nb * (double)N * 2.0 * D8);
    e = P.start(st);
    launch_gram_sf(dim, B.sf_bps, B.sf_nbp, B.sf_pairs, B.sf_F, B.sf_nF, B.sf_q, w.Wip, B.prog.n_lt,
                   nvar * B.npr, L.QG, w.A, lda, sA, nb, B.sf_lds, st, w.s, B.nS);
    P.stop(e, st, DPG_STAGE_GRAM, nb * B.gram_flops, nb * ((double)B.nWip + N * (N + 1.0) / 2.0) * D8);
  } else if (sf) {
    // G an output too: sum-factorised G (unscaled), its copy-out, then the
    // DPOEQU factors from its diagonal and S G S in place
    e = P.start(st);
    launch_gram_sf(dim, B.sf_bps, B.sf_nbp, B.sf_pairs, B.sf_F, B.sf_nF, B.sf_q, w.Wip, B.prog.n_lt,
                   nvar * B.npr, L.QG, w.A, lda, sA, nb, B.sf_lds, st);
    P.stop(e, st, DPG_STAGE_GRAM, nb * B.gram_flops, nb * ((double)B.nWip + N * (N + 1.0) / 2.0) * D8);
    e = P.start(st);
    launch_extract_sym(w.A, lda, sA, 0, N, 1.0, Gout + (size_t)c0 * N * N, nb, st);
    P.stop(e, st, DPG_STAGE_EXTRACT, 0.0, nb * (double)N * N * 2.0 * D8);
    e = P.start(st);
    launch_scaling(w.A, lda, sA, 0, N, N, N, w.s, B.nS, nb, 1.0, st);
    P.stop(e, st, DPG_STAGE_SCALE, 0.0, nb * (double)N * (N + 1) / 2.0 * 16.0);
  }
  // DPOEQU scaling is applied in the GEMM epilogues, except when G itself is an
  // output of the dense path: then G is written unscaled, copied out, and scaled in place
  const double* s_ep = (Gout && !sf) ? nullptr : w.s;
  // G = At^T At (lower), scaled by s_i s_j in the epilogue
  GemmArgs g{};
  g.X = w.At;
  g.ldx = B.Npad;
  g.sx = B.nAt;
  g.Y = w.At;
  g.ldy = B.Npad;
  g.sy = B.nAt;
  g.C = w.A;
  g.ldc = lda;
  g.sc = sA;
  g.m = g.n = N;
  g.k = KG;
  g.lower = 1;
  g.alpha = 1.0;
  g.beta = 0.0;
  g.ncells = nb;
  g.srow = s_ep;
  g.scol = s_ep;
  g.ss = B.nS;
  const int ts = B.tile[0] == 64 ? 0 : 1;
  g.kmask = B.merge_off ? B.gmask[ts] : B.gmask_merged[ts];
  g.kw = B.merge_off ? nullptr : B.gweight[ts];
  g.nseg = B.prog.n_lt;
  g.kseg = B.QGp;
  if (prec != DPG_PREC_FP64 && !i8) {
    // low-precision operands are pre-scaled (keeps FP16 in range): X(k,i) s_i, Y(k,j) s_j
    g.srow = g.scol = nullptr;
    g.px = g.py = s_ep;
    g.kmask = nullptr;
    g.kw = nullptr;
  }
  if (i8) {
    // G = x^T x from the digit planes (x already carries the DPOEQU factors)
    I8GramArgs a{};
    a.dig = (const int8_t*)w.At;
    a.sd = B.nAt * (int64_t)sizeof(double);
    a.N = N;
    a.NR = B.NR;
    a.nch = KG / 32;
    a.ncells = nb;
    a.C = w.A;
    a.ldc = lda;
    a.sc = sA;
    a.kmask = B.gmask[0];
    a.kseg = B.QGp;
    e = P.start(st);
    launch_gram_i8(a, st);
    P.stop(e, st, DPG_STAGE_GRAM, nb * B.gram_flops, nb * ((double)KG * B.NR * kI8Digits + N * (N + 1.0) / 2.0 * D8));
  } else if (!sf) {
    e = P.start(st);
    if (prec == DPG_PREC_FP64) launch_gemm(g, kGemmGram, st, B.tile[0]);
    else launch_gemm_mixed(g, prec, kGemmGram, st);
    P.stop(e, st, DPG_STAGE_GRAM, nb * B.gram_flops, nb * ((double)KG * N + N * (N + 1.0) / 2.0) * D8);
  }
  if (Gout && !sf) {
    e = P.start(st);
    launch_extract_sym(w.A, lda, sA, 0, N, 1.0, Gout + (size_t)c0 * N * N, nb, st);
    P.stop(e, st, DPG_STAGE_EXTRACT, 0.0, nb * (double)N * N * 2.0 * D8);
  }
  // B^T blocks: rows N + trial offset, columns = test dofs
  e = P.start(st);
  double bfl = 0.0, bby = 0.0;
  if (sfb) {
    launch_bmat_sf(dim, B.sfb_bps, B.sfb_nbp, B.sfb_pairs, B.sfb_F, B.sfb_nF, B.sfb_q, w.Wvol,
                   B.prog.n_vb * nvar * B.npr, L.Q, w.A, lda, sA, s_ep, B.nS, nb, B.sfb_lds, st);
    // counted as the dense contraction it replaces; bytes: weights in, B out
    bfl += 2.0 * L.F * N * L.Q;
    bby += ((double)B.nWvol + (double)L.F * N) * D8;
  }
  for (int vb = 0; vb < (sfb ? 0 : B.prog.n_vb); ++vb) {
    const TrialBlock& blk = L.blocks[B.vb_block[vb]];
    GemmArgs b{};
    b.X = B.tab.field_vol;
    b.ldx = L.nf;
    b.sx = 0;
    b.Y = w.Tv + (size_t)vb * L.Q * B.Npad;
    b.ldy = B.Npad;
    b.sy = B.nTv;
    b.C = w.A + N + blk.offset;
    b.ldc = lda;
    b.sc = sA;
    b.m = blk.card;
    b.n = N;
    b.k = L.Q;
    b.alpha = 1.0;
    b.ncells = nb;
    b.scol = s_ep;
    b.ss = B.nS;
    if (prec == DPG_PREC_FP64 || i8) {
      launch_gemm(b, kGemmBmat, st, B.tile[1]);
    } else {
      b.scol = nullptr;
      b.py = s_ep;
      launch_gemm_mixed(b, prec, kGemmBmat, st);
    }
    bfl += 2.0 * b.m * b.n * b.k;
    bby += ((double)b.k * b.n + (double)b.m * b.n) * D8;
  }
  if (sfs) {
    launch_bmat_side_sf(dim, B.sfs_bps, B.sfs_nbp, B.sfs_pairs, B.sfs_F, B.sfs_nF, B.sfs_nfam, B.sfs_q, w.Wside,
                        S * L.Qs, B.prog.n_sb * nvar * B.npr, w.A, lda, sA, s_ep, B.nS, nb, B.sfs_lds, st);
    bfl += 2.0 * L.T * N * L.Qs;  // counted as the side contractions it replaces
    bby += ((double)B.nWside + (double)L.T * N) * D8;
  }
  for (int sb = 0; sb < (sfs ? 0 : B.prog.n_sb); ++sb) {
    int v = B.sb_var[sb];
    for (const TrialBlock& blk : L.blocks) {
      if (blk.var != v || blk.side < 0) continue;
      GemmArgs b{};
      b.X = blk.kind == DPG_TRACE ? B.tab.trace : B.tab.flux;
      b.ldx = blk.card;
      b.sx = 0;
      b.Y = w.Ts + ((size_t)sb * S + blk.side) * L.Qs * B.Npad;
      b.ldy = B.Npad;
      b.sy = B.nTs;
      b.C = w.A + N + blk.offset;
      b.ldc = lda;
      b.sc = sA;
      b.m = blk.card;
      b.n = N;
      b.k = L.Qs;
      b.alpha = 1.0;
      b.ncells = nb;
      b.scol = s_ep;
      b.ss = B.nS;
      if (prec == DPG_PREC_FP64 || i8) {
        launch_gemm(b, kGemmBmat, st, B.tile[1]);
      } else {
        b.scol = nullptr;
        b.py = s_ep;
        launch_gemm_mixed(b, prec, kGemmBmat, st);
      }
      bfl += 2.0 * b.m * b.n * b.k;
      bby += ((double)b.k * b.n + (double)b.m * b.n) * D8;
    }
  }
  // load vector row, residual row (from the already scaled rows)
  launch_rhs(dim, w.Wrhs, nvar, B.tab.R_vol, B.tab.var_of_dof, N, L.Q, nb, w.A, lda, sA, N + M, s_ep, B.nS, st);
  launch_residual_row(w.A, lda, sA, N, M, N, N + M, N + M + 1, u ? u + (size_t)c0 * M : nullptr, nb, st);
  bfl += 2.0 * L.Q * N * B.npr + (u ? 2.0 * M * N : 0.0);
  P.stop(e, st, DPG_STAGE_BMAT, nb * bfl, nb * bby);
  if (Gout && !sf) {
    e = P.start(st);
    launch_scaling(w.A, lda, sA, 0, B.D, N, N, w.s, B.nS, nb, 1.0, st);
    P.stop(e, st, DPG_STAGE_SCALE, 0.0, nb * ((double)N * (N + 1) / 2.0 + (double)(B.D - N) * N) * 16.0);
  }
  launch_fill_i32(w.info, nb, 0, st);
  // the trailing block starts as 0 and is written (beta = 0) by the closing SYRK
  factor_columns(plan, B, nb, w, 0, N, 0.0, st, true);
}

// Stage 2: static condensation of the field dofs of K = -A[N:, N:]
void stage_schur(dpg_plan* plan, const BucketPlan& B, int nb, const Work& w, hipStream_t st) {
  const ElementLayout& L = B.L;
  const int N = L.N, F = L.F;
  Profiler& P = plan->prof;
  // negate the trailing block and equilibrate its field rows/columns
  hipEvent_t e = P.start(st);
  launch_scaling(w.A, B.lda, B.nA, N, B.D, F, B.D - N, w.s, B.nS, nb, -1.0, st);
  P.stop(e, st, DPG_STAGE_SCALE, 0.0, nb * (double)(B.D - N) * (B.D - N + 1) / 2.0 * 16.0);
  factor_columns(plan, B, nb, w, N, N + F, 1.0, st);
}

// dpg_assemble's pipeline for cells [c0, c0+nb) of one batch
void assemble_batch(dpg_plan* plan, const BucketPlan& B, const dpg_batch& bt, int c0, int nb, const Work& w,
                    const dpg_outputs& o, hipStream_t st) {
  const int N = B.L.N, M = B.L.M, F = B.L.F, T = B.L.T;
  stage_factor(plan, B, bt, c0, nb, w, nullptr, o.G, st);
  hipEvent_t e = plan->prof.start(st);
  if (o.factor) {
    // before the Schur stage, which re-uses the scale vector
    launch_extract_factor(w.A, B.lda, B.nA, N, M, w.Ld, B.nLd, o.factor + (size_t)c0 * (N + M + 1) * N,
                          (int64_t)(N + M + 1) * N, nb, st);
    launch_extract_row(w.s, 1, B.nS, 0, 0, N, 1.0, o.factor_scale + (size_t)c0 * N, nb, st);
  }
  if (o.K) launch_extract_sym(w.A, B.lda, B.nA, N, M, -1.0, o.K + (size_t)c0 * M * M, nb, st);
  if (o.rhs) launch_extract_row(w.A, B.lda, B.nA, N + M, N, M, -1.0, o.rhs + (size_t)c0 * M, nb, st);
  plan->prof.stop(e, st, DPG_STAGE_EXTRACT, 0.0,
                  nb * ((double)M * M * 2.0 + M * 2.0 + (o.factor ? 2.0 * (N + M + 1.0) * N + 2.0 * N : 0.0)) * 8.0);
  if (o.schur || o.schur_rhs) {
    stage_schur(plan, B, nb, w, st);
    e = plan->prof.start(st);
    if (o.schur) launch_extract_sym(w.A, B.lda, B.nA, N + F, T, 1.0, o.schur + (size_t)c0 * T * T, nb, st);
    if (o.schur_rhs)
      launch_extract_row(w.A, B.lda, B.nA, N + M, N + F, T, 1.0, o.schur_rhs + (size_t)c0 * T, nb, st);
    plan->prof.stop(e, st, DPG_STAGE_EXTRACT, 0.0, nb * ((double)T * T * 2.0 + T * 2.0) * 8.0);
  }
}

// The sum-factorised G (tensor-product cells) against the dense one, on the
// first cells of a batch, with the tiles already chosen: it is switched on when
// it is faster over the whole pipeline and moves K no more than 4 times what a
// second rounding of the dense G does (the same G with its equal IP terms summed
// separately instead of merged, i.e. another FP64 summation order of the same
// products). A form without mergeable terms has no second rounding (the two
// dense K agree bit for bit), and then the sum-factorised G is not taken. On
// the 32^3-hex workload it passes (K 9e-11 from an 80-bit evaluation against
// 6.6e-10 for the dense G: far fewer roundings per entry). The timing runs over
// the whole sample: on the 2D Stokes mesh the sum-factorised path wins on a
// few dozen cells and loses on thousands, and G stays dense there.
void choose_gram_path(dpg_plan* plan, BucketPlan& B, const dpg_batch& bt, int ns, const Work& w, hipStream_t st) {
  const int N = B.L.N, M = B.L.M;
  const int64_t mm = (int64_t)M * M;
  // up to 128 cells, and at most 512 MiB per copy of K
  const int ng = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)ns, 128, (512ll << 20) / (8 * mm)}));
  DeviceBuffer kd, kx, nrm;
  kd.ensure(sizeof(double) * ng * mm);
  kx.ensure(sizeof(double) * ng * mm);
  nrm.ensure(sizeof(double) * 2 * ng);
  std::vector<double> h(2 * (size_t)ng);
  // timed on the whole sample (a few dozen cells would favour the sum-factorised
  // kernels, whose small workgroups fill the GPU sooner than the GEMM tiles);
  // K of its first ng cells compared
  auto run = [&](double* K) {  // warm pass, timed pass; K of the timed pass
    double t = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      plan->prof.acc = dpg_timings{};
      stage_factor(plan, B, bt, 0, ns, w, nullptr, nullptr, st);
      launch_extract_sym(w.A, B.lda, B.nA, N, M, -1.0, K, ng, st);
      stage_schur(plan, B, ns, w, st);
      plan->prof.collect(st, ns);
      t = 0.0;
      for (int k = 0; k < DPG_NSTAGES; ++k) t += plan->prof.acc.ms[k];
    }
    return t;
  };
  auto worst = [&]() {  // max over cells of |K_x - K_d| / |K_d|
    launch_cell_diff_norms((const double*)kx.p, (const double*)kd.p, mm, ng, (double*)nrm.p, st);
    HIP_CHECK(hipMemcpyAsync(h.data(), nrm.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double e = 0.0;
    for (int c = 0; c < ng; ++c) e = std::max(e, std::sqrt(h[2 * c] / std::max(h[2 * c + 1], 1e-300)));
    return e;
  };
  B.sf_on = false;
  B.merge_off = false;
  const double t_dense = run((double*)kd.p);
  B.merge_off = true;
  run((double*)kx.p);
  B.merge_off = false;
  const double e_dense = worst();
  B.sf_on = true;
  const double t_sf = run((double*)kx.p);
  const double e_sf = worst();
  B.sf_on = t_sf < t_dense && e_sf <= 4.0 * e_dense + 1e-15;
}

void check_batch(const dpg_plan* plan, int bucket, const dpg_batch* batch) {
  if (!plan) throw std::invalid_argument("plan is NULL");
  if (bucket < 0 || bucket >= (int)plan->buckets.size()) throw std::invalid_argument("bucket index out of range");
  if (!batch) throw std::invalid_argument("batch is NULL");
  if (batch->ncells < 0) throw std::invalid_argument("ncells < 0");
  if (batch->ncells > 0 && (!batch->nodes || !batch->parities))
    throw std::invalid_argument("nodes and parities are required");
  if (plan->form.n_coefs > 0 && batch->ncells > 0 && !batch->coefs)
    throw std::invalid_argument("the form needs per-cell coefficients");
  if (batch->load.ncomp < 0 || batch->load.ncomp > 3) throw std::invalid_argument("bad load");
  for (int c = 0; c < batch->load.ncomp; ++c)
    if (batch->load.nmodes[c] < 0 || batch->load.nmodes[c] > DPG_MAX_LOAD_MODES)
      throw std::invalid_argument("bad load mode count");
}

void check_outputs(const dpg_outputs* o) {
  if (!o) throw std::invalid_argument("outputs is NULL");
  if ((o->factor == nullptr) != (o->factor_scale == nullptr))
    throw std::invalid_argument("factor and factor_scale go together");
}

void prepare_workspace(dpg_plan* plan, const BucketPlan& B, int nb, int64_t per_cell_bytes = 0) {
  plan->ws.ensure((size_t)(per_cell_bytes > 0 ? per_cell_bytes : B.per_cell_bytes) * nb);
  plan->info.ensure(sizeof(int32_t) * std::max(nb, 1));
}

int count_failures(const Work& w, int nb, int32_t* out_info, hipStream_t st) {
  std::vector<int32_t> h(nb);
  HIP_CHECK(hipMemcpyAsync(h.data(), w.info, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  int bad = 0;
  for (int i = 0; i < nb; ++i) bad += h[i] != 0;
  if (out_info) HIP_CHECK(hipMemcpy(out_info, h.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice));
  return bad;
}

}  // namespace
}  // namespace dpg

extern "C" {

const char* dpg_version(void) { return "camellia2_amd-dpg 0.1 (gfx950)"; }

const char* dpg_last_error(void) { return dpg::g_last_error.c_str(); }

int dpg_bucket_sizes(const dpg_form_desc* form, const dpg_bucket_desc* bucket, dpg_bucket_info* out) {
  return dpg::guarded([&] {
    if (!form || !bucket || !out) throw std::invalid_argument("NULL argument");
    dpg::validate_form(*form);
    dpg::ElementLayout L = dpg::make_layout(*form, *bucket);
    std::memset(out, 0, sizeof(*out));
    out->N = L.N;
    out->M = L.M;
    out->F = L.F;
    out->T = L.T;
    out->Q = L.Q;
    out->Qs = L.Qs;
    out->QG = L.QG;
    out->nsides = L.rc.nsides;
    out->nverts = L.rc.nverts;
    out->flops_per_cell = dpg::flops_per_cell(*form, L);
  });
}

int dpg_plan_create(const dpg_form_desc* form, const dpg_bucket_desc* buckets, int nbuckets,
                    const dpg_options* options, dpg_plan** out) {
  return dpg::guarded([&] {
    if (!form || !buckets || nbuckets <= 0 || !out) throw std::invalid_argument("NULL argument or no buckets");
    *out = nullptr;
    dpg::validate_form(*form);
    std::unique_ptr<dpg_plan> plan(new dpg_plan());
    plan->form = *form;
    if (options) plan->opt = *options;
    if (plan->opt.precision < DPG_PREC_FP64 || plan->opt.precision > DPG_PREC_INT8SPLIT)
      throw std::invalid_argument("unknown precision");
    if (plan->opt.precision == DPG_PREC_INT8SPLIT &&
        dpg::ip_digits_lds_bytes(form->n_ipterms, form->n_test, form->spatial_dim) > dpg::kMaxLds)
      throw std::invalid_argument("the form has too many test variables x IP terms for the int8 digit kernel's LDS");
    if (plan->opt.refine != 0)
      throw std::invalid_argument("refine: a refinement of G^-1 B cannot correct an error that sits in G itself "
                                  "(the factorisation is FP64); not supported");
    HIP_CHECK(hipGetDevice(&plan->device));
    dpg::static_scratch(plan->device);  // set-up of dpg_assemble_noalloc's workspace on this device
    plan->prof.on = plan->opt.profile != 0;
    plan->buckets.resize(nbuckets);
    for (int i = 0; i < nbuckets; ++i) dpg::build_bucket(plan.get(), buckets[i], plan->buckets[i]);
    *out = plan.release();
  });
}

int dpg_plan_destroy(dpg_plan* plan) {
  return dpg::guarded([&] { delete plan; });
}

int dpg_plan_clone_workspace(const dpg_plan* plan, double workspace_bytes, dpg_plan** out) {
  return dpg::guarded([&] {
    if (!plan || !out) throw std::invalid_argument("NULL argument");
    *out = nullptr;
    if (!(workspace_bytes >= 0.0)) throw std::invalid_argument("workspace_bytes must be >= 0");
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev != plan->device) throw std::invalid_argument("the current device is not the plan's device");
    std::unique_ptr<dpg_plan> c(new dpg_plan());
    c->form = plan->form;
    c->opt = plan->opt;
    if (workspace_bytes > 0.0) c->opt.workspace_bytes = workspace_bytes;
    c->device = plan->device;
    c->buckets = plan->buckets;  // device pointers into the shared tables, tile choices as they are now
    c->tables = plan->tables;
    c->prof.on = plan->prof.on;
    *out = c.release();
  });
}

int dpg_plan_workspace(const dpg_plan* plan, size_t* bytes) {
  return dpg::guarded([&] {
    if (!plan || !bytes) throw std::invalid_argument("NULL argument");
    *bytes = plan->ws.bytes;
  });
}

int dpg_plan_timings(dpg_plan* plan, dpg_timings* out, int reset) {
  return dpg::guarded([&] {
    if (!plan || !out) throw std::invalid_argument("NULL argument");
    *out = plan->prof.acc;
    if (reset) plan->prof.acc = dpg_timings{};
  });
}

int dpg_assemble(dpg_plan* plan, int bucket, const dpg_batch* batch, const dpg_outputs* outputs, void* stream) {
  return dpg::guarded_status([&] {
    dpg::check_batch(plan, bucket, batch);
    dpg::check_outputs(outputs);
    if (batch->ncells == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const dpg::BucketPlan& B = plan->buckets[bucket];
    int C = batch->ncells;
    int maxb = dpg::batch_cells(plan, B, C);
    dpg::prepare_workspace(plan, B, maxb);
    int bad = 0;
    for (int c0 = 0; c0 < C; c0 += maxb) {
      int nb = std::min(maxb, C - c0);
      dpg::Work w = dpg::carve(plan->ws.p, (int32_t*)plan->info.p, B, nb);
      dpg::assemble_batch(plan, B, *batch, c0, nb, w, *outputs, st);
      plan->prof.collect(st, nb);
      // the next batch reuses the workspace in stream order
      if (outputs->info) bad += dpg::count_failures(w, nb, outputs->info + c0, st);
    }
    return bad;
  });
}

int dpg_assemble_noalloc(dpg_plan* plan, int bucket, const dpg_batch* batch, const dpg_outputs* outputs,
                         void* stream) {
  return dpg::guarded_status([&] {
    dpg::check_batch(plan, bucket, batch);
    dpg::check_outputs(outputs);
    if (batch->ncells == 0) return 0;
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev != plan->device) throw std::invalid_argument("the current device is not the plan's device");
    hipStream_t st = (hipStream_t)stream;
    const dpg::BucketPlan& B = plan->buckets[bucket];
    // scratch layout: INFO of a batch (at most 65535 cells), the failure count,
    // then the per-cell workspace regions of the batch
    constexpr size_t kInfoBytes = 65536 * sizeof(int32_t), kCountBytes = 256;
    constexpr size_t kWsOffset = kInfoBytes + kCountBytes;
    const size_t room = (size_t)DPG_NOALLOC_SCRATCH_BYTES - kWsOffset;
    if ((size_t)B.per_cell_bytes > room)
      throw std::invalid_argument("dpg_assemble_noalloc: one cell of this element type needs " +
                                  std::to_string(B.per_cell_bytes) + " bytes of workspace, the static scratch holds " +
                                  std::to_string(room) + " (DPG_NOALLOC_SCRATCH_BYTES); use dpg_assemble");
    int C = batch->ncells;
    int maxb = (int)std::min<size_t>(room / (size_t)B.per_cell_bytes, (size_t)dpg::batch_cells(plan, B, C));
    unsigned char* base = dpg::static_scratch(dev);
    int32_t* info = (int32_t*)base;
    int32_t* count = (int32_t*)(base + kInfoBytes);
    // no per-launch timing events in this path
    struct ProfOff {
      dpg::Profiler& p;
      bool was;
      explicit ProfOff(dpg::Profiler& q) : p(q), was(q.on) { p.on = false; }
      ~ProfOff() { p.on = was; }
    } prof_off(plan->prof);
    int bad = 0;
    for (int c0 = 0; c0 < C; c0 += maxb) {
      int nb = std::min(maxb, C - c0);
      dpg::Work w = dpg::carve(base + kWsOffset, info, B, nb);
      dpg::assemble_batch(plan, B, *batch, c0, nb, w, *outputs, st);
      if (outputs->info) {
        dpg::launch_info_count(info, nb, outputs->info + c0, count, st);
        int32_t h = 0;
        HIP_CHECK(hipMemcpyAsync(&h, count, sizeof(h), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        bad += h;
      }
    }
    return bad;
  });
}

int dpg_energy_error(dpg_plan* plan, int bucket, const dpg_batch* batch, const double* u, double* eta2,
                     void* stream) {
  return dpg::guarded_status([&] {
    dpg::check_batch(plan, bucket, batch);
    if (batch->ncells == 0) return 0;  // an empty batch needs no buffers
    if (!u || !eta2) throw std::invalid_argument("u and eta2 are required");
    hipStream_t st = (hipStream_t)stream;
    const dpg::BucketPlan& B = plan->buckets[bucket];
    int C = batch->ncells;
    int maxb = dpg::batch_cells(plan, B, C);
    dpg::prepare_workspace(plan, B, maxb);
    int bad = 0;
    for (int c0 = 0; c0 < C; c0 += maxb) {
      int nb = std::min(maxb, C - c0);
      dpg::Work w = dpg::carve(plan->ws.p, (int32_t*)plan->info.p, B, nb);
      dpg::stage_factor(plan, B, *batch, c0, nb, w, u, nullptr, st);
      dpg::launch_extract_row(w.A, B.lda, B.nA, B.D - 1, B.D - 1, 1, -1.0, eta2 + c0, nb, st);
      plan->prof.collect(st, nb);
      bad += dpg::count_failures(w, nb, nullptr, st);
    }
    return bad;
  });
}

int dpg_recover_fields(dpg_plan* plan, int bucket, int ncells, const double* K, const double* rhs, double* u,
                       void* stream) {
  return dpg::guarded_status([&] {
    if (!plan) throw std::invalid_argument("plan is NULL");
    if (bucket < 0 || bucket >= (int)plan->buckets.size()) throw std::invalid_argument("bucket index out of range");
    if (ncells < 0) throw std::invalid_argument("ncells < 0");
    if (ncells == 0) return 0;  // an empty batch needs no buffers
    if (!K || !rhs || !u) throw std::invalid_argument("K, rhs and u are required");
    hipStream_t st = (hipStream_t)stream;
    const dpg::BucketPlan& B = plan->buckets[bucket];
    const int F = B.L.F, M = B.L.M;
    if (F == 0) return 0;
    // the field block and one augmented row, in the bucket's workspace layout
    dpg::BucketPlan R = B;
    R.D = F + 1;
    R.zmask[0] = R.zmask[1] = nullptr;  // not the augmented matrix of G: no known zeros
    int maxb = dpg::batch_cells(plan, B, ncells);
    dpg::prepare_workspace(plan, B, maxb);
    int bad = 0;
    for (int c0 = 0; c0 < ncells; c0 += maxb) {
      int nb = std::min(maxb, ncells - c0);
      dpg::Work w = dpg::carve(plan->ws.p, (int32_t*)plan->info.p, B, nb);
      const size_t o = (size_t)c0 * M;
      dpg::launch_recover_setup(K + o * M, rhs + o, u + o, M, F, w.A, B.lda, B.nA, w.s, B.nS, nb, st);
      dpg::launch_fill_i32(w.info, nb, 0, st);
      dpg::factor_columns(plan, R, nb, w, 0, F, 0.0, st);
      dpg::launch_backsolve(w.A, B.lda, B.nA, F, F, w.Ld, B.nLd, w.s, B.nS, u + o, M, nb, st);
      plan->prof.collect(st, nb);
      bad += dpg::count_failures(w, nb, nullptr, st);
    }
    return bad;
  });
}

int dpg_energy_error_factored(dpg_plan* plan, int bucket, int ncells, const double* factor, const double* u,
                              double* eta2, void* stream) {
  return dpg::guarded([&] {
    if (!plan) throw std::invalid_argument("plan is NULL");
    if (bucket < 0 || bucket >= (int)plan->buckets.size()) throw std::invalid_argument("bucket index out of range");
    if (ncells < 0) throw std::invalid_argument("ncells < 0");
    if (ncells == 0) return;
    if (!factor || !u || !eta2) throw std::invalid_argument("factor, u and eta2 are required");
    const dpg::ElementLayout& L = plan->buckets[bucket].L;
    for (int c0 = 0; c0 < ncells; c0 += 65535) {
      const int nb = std::min(65535, ncells - c0);
      dpg::launch_factored_residual(factor + (size_t)c0 * (L.N + L.M + 1) * L.N, (int64_t)(L.N + L.M + 1) * L.N,
                                    L.N, L.M, u + (size_t)c0 * L.M, nullptr, 0, eta2 + c0, nb, (hipStream_t)stream);
    }
  });
}

int dpg_adjoint(dpg_plan* plan, int bucket, const dpg_batch* batch, const double* u, const double* Kbar,
                const double* rhsbar, const double* factor, const double* factor_scale, double* eta2, double* grad,
                void* stream) {
  return dpg::guarded_status([&] {
    dpg::check_batch(plan, bucket, batch);
    if (grad && plan->form.n_coefs <= 0) throw std::invalid_argument("the form has no per-cell coefficients");
    if (plan->opt.precision != DPG_PREC_FP64) throw std::invalid_argument("the adjoint runs in FP64 only");
    if ((factor == nullptr) != (factor_scale == nullptr))
      throw std::invalid_argument("factor and factor_scale go together");
    if (eta2 && !u) throw std::invalid_argument("eta2 needs u");
    if (batch->ncells == 0 || (!grad && !eta2)) return 0;  // nothing to compute
    hipStream_t st = (hipStream_t)stream;
    const dpg::BucketPlan& B = plan->buckets[bucket];
    const dpg::ElementLayout& L = B.L;
    const int N = L.N, M = L.M, S = L.rc.nsides, dim = L.dim, V = L.rc.nverts, nc = plan->form.n_coefs;
    // columns of X: G^-1 B (for Kbar or rhsbar), G^-1 l (rhsbar), G^-1 r (u)
    const int nT = (Kbar || rhsbar) ? M : 0;
    const int cy = rhsbar ? nT : -1;
    const int ce = u ? nT + (rhsbar ? 1 : 0) : -1;
    const int ncol = nT + (rhsbar ? 1 : 0) + (u ? 1 : 0);
    if (grad && dpg::adjoint_lds_bytes(plan->form.n_test, dim, ncol) > 64 * 1024)
      throw std::invalid_argument("adjoint: the element type is too large for the adjoint kernel's LDS");
    const int64_t sF = (int64_t)(N + M + 1) * N;
    const bool stored = factor != nullptr;
    // algorithmic work per cell of the adjoint stage: back solves (N^2 per
    // column), the column functions' prims at every point (2 N npr per column),
    // the Kbar products of the terms whose weight has a coefficient (2 M^2 per
    // IP term and Gram point, 2 M card per trial block and point); bytes: the
    // factor, the cotangents and u read, the gradient written
    double adj_flops = (double)N * N * ncol + 2.0 * N * (dim + 1) * ncol * (L.QG + L.Q + S * L.Qs);
    if (Kbar) {
      const dpg::DevProgram& P = B.prog;
      auto has_coef = [&](int b0, int b1) {
        for (int i = b0; i < b1; ++i)
          if (P.s[i].slot >= 0) return true;
        return false;
      };
      for (int lt = 0; lt < P.n_lt; ++lt)
        if (has_coef(P.lt_beg[lt], P.lt_beg[lt + 1])) adj_flops += 2.0 * M * M * L.QG;
      for (int g = 0; g < P.n_vb; ++g)
        if (has_coef(P.vb_beg[g], P.vb_beg[g + 1])) adj_flops += 2.0 * M * P.nf * L.Q;
      for (int g = 0; g < P.n_sb; ++g)
        if (has_coef(P.sb_beg[g], P.sb_beg[g + 1])) adj_flops += 2.0 * M * P.sb_card[g] * S * L.Qs;
    }
    const double adj_bytes = 8.0 * ((double)sF + N + (Kbar ? (double)M * M : 0.0) + (rhsbar ? M : 0) + (u ? M : 0) +
                                    (double)nc * (dim * V + 1));
    // with a stored factor only X and z are needed
    const int64_t per_cell = stored ? (B.nX + B.nZ) * (int64_t)sizeof(double) : B.per_cell_bytes_adj;
    int C = batch->ncells;
    int maxb = dpg::batch_cells(plan, B, C, per_cell);
    dpg::prepare_workspace(plan, B, maxb, per_cell);
    int bad = 0;
    for (int c0 = 0; c0 < C; c0 += maxb) {
      int nb = std::min(maxb, C - c0);
      dpg::Work w{};
      const double *F, *s;
      int64_t ss;
      if (stored) {
        w.X = (double*)plan->ws.p;
        w.Z = w.X + B.nX * nb;
        F = factor + (size_t)c0 * sF;
        s = factor_scale + (size_t)c0 * N;
        ss = N;
        if (u)
          dpg::launch_factored_residual(F, sF, N, M, u + (size_t)c0 * M, w.Z, B.nZ, eta2 ? eta2 + c0 : nullptr, nb,
                                        st);
      } else {
        w = dpg::carve(plan->ws.p, (int32_t*)plan->info.p, B, nb, true);
        // the forward factorisation with the residual row r = l - B u: its last
        // row then holds z = Lt^-1 S r and its trailing entry -eta2
        dpg::stage_factor(plan, B, *batch, c0, nb, w, u, nullptr, st);
        dpg::launch_extract_factor(w.A, B.lda, B.nA, N, M, w.Ld, B.nLd, w.Fac, B.nFac, nb, st);
        F = w.Fac;
        s = w.s;
        ss = B.nS;
        if (u) {
          dpg::launch_extract_row(w.A, B.lda, B.nA, B.D - 1, 0, N, 1.0, w.Z, nb, st);
          if (eta2) dpg::launch_extract_row(w.A, B.lda, B.nA, B.D - 1, B.D - 1, 1, -1.0, eta2 + c0, nb, st);
        }
      }
      const int64_t fstride = stored ? sF : B.nFac;
      // z of extract_row is packed with stride N; factored_residual wrote it with stride nZ
      const int64_t zstride = stored ? B.nZ : N;
      if (grad) {
        hipEvent_t e = plan->prof.start(st);
        dpg::launch_factor_backsolve(F, fstride, N, M, s, ss, nT, rhsbar ? 1 : 0, u ? w.Z : nullptr, zstride, w.X,
                                     (int64_t)N * ncol, ncol, nb, st);
        dpg::launch_adjoint(dim, B.prog_d, B.tab, N, M, L.QG, L.Q, L.Qs, S, plan->form.n_test,
                            batch->nodes + (size_t)c0 * V * dim, batch->parities + (size_t)c0 * S,
                            batch->coefs + (size_t)c0 * nc, w.X, (int64_t)N * ncol, ncol, nT, cy, ce,
                            Kbar ? Kbar + (size_t)c0 * M * M : nullptr, rhsbar ? rhsbar + (size_t)c0 * M : nullptr,
                            u ? u + (size_t)c0 * M : nullptr, grad + (size_t)c0 * nc, nb, st);
        plan->prof.stop(e, st, DPG_STAGE_ADJOINT, nb * adj_flops, nb * adj_bytes);
      }
      plan->prof.collect(st, nb);
      if (!stored) bad += dpg::count_failures(w, nb, nullptr, st);
    }
    return bad;
  });
}

int dpg_riesz_representation(dpg_plan* plan, int bucket, const dpg_batch* batch, const double* u, double* e,
                             void* stream) {
  return dpg::guarded_status([&] {
    dpg::check_batch(plan, bucket, batch);
    if (plan->opt.precision != DPG_PREC_FP64) throw std::invalid_argument("runs in FP64 only");
    if (batch->ncells == 0) return 0;  // an empty batch needs no buffers
    if (!u || !e) throw std::invalid_argument("u and e are required");
    hipStream_t st = (hipStream_t)stream;
    const dpg::BucketPlan& B = plan->buckets[bucket];
    const int N = B.L.N;
    int C = batch->ncells;
    int maxb = dpg::batch_cells(plan, B, C);
    dpg::prepare_workspace(plan, B, maxb);
    int bad = 0;
    for (int c0 = 0; c0 < C; c0 += maxb) {
      int nb = std::min(maxb, C - c0);
      dpg::Work w = dpg::carve(plan->ws.p, (int32_t*)plan->info.p, B, nb);
      dpg::stage_factor(plan, B, *batch, c0, nb, w, u, nullptr, st);
      dpg::launch_backsolve(w.A, B.lda, B.nA, N, B.D - 1, w.Ld, B.nLd, w.s, B.nS, e + (size_t)c0 * N, N, nb, st);
      plan->prof.collect(st, nb);
      bad += dpg::count_failures(w, nb, nullptr, st);
    }
    return bad;
  });
}

int dpg_plan_get_tiles(const dpg_plan* plan, int bucket, dpg_tiles* out) {
  return dpg::guarded([&] {
    if (!plan || !out) throw std::invalid_argument("NULL argument");
    if (bucket < 0 || bucket >= (int)plan->buckets.size()) throw std::invalid_argument("bucket index out of range");
    const int* t = plan->buckets[bucket].tile;
    *out = dpg_tiles{t[0], t[1], t[2], plan->buckets[bucket].sf_on ? 1 : 0};
  });
}

int dpg_plan_set_tiles(dpg_plan* plan, int bucket, const dpg_tiles* tiles) {
  return dpg::guarded([&] {
    if (!plan || !tiles) throw std::invalid_argument("NULL argument");
    if (bucket < 0 || bucket >= (int)plan->buckets.size()) throw std::invalid_argument("bucket index out of range");
    int v[3] = {tiles->gram, tiles->bmat, tiles->update};
    for (int r = 0; r < 3; ++r)
      if (v[r] != 64 && v[r] != 32 && !(r == 2 && v[r] == 128))
        throw std::invalid_argument("tile edges must be 64 or 32 (128 for the updates)");
    if (tiles->sumfact != 0 && tiles->sumfact != 1) throw std::invalid_argument("sumfact must be 0 or 1");
    if (tiles->sumfact && !plan->buckets[bucket].sf_ok)
      throw std::invalid_argument("the sum-factorised G is not available for this bucket");
    for (int r = 0; r < 3; ++r) plan->buckets[bucket].tile[r] = v[r];
    plan->buckets[bucket].sf_on = tiles->sumfact != 0;
  });
}

int dpg_plan_autotune(dpg_plan* plan, int bucket, const dpg_batch* batch, int sample, void* stream) {
  return dpg::guarded([&] {
    dpg::check_batch(plan, bucket, batch);
    if (sample <= 0) throw std::invalid_argument("sample must be positive");
    hipStream_t st = (hipStream_t)stream;
    dpg::BucketPlan& B = plan->buckets[bucket];
    int ns = std::min(sample, batch->ncells);
    if (ns == 0) return;
    ns = std::min(ns, dpg::batch_cells(plan, B, ns));
    dpg::prepare_workspace(plan, B, ns);
    dpg::Work w = dpg::carve(plan->ws.p, (int32_t*)plan->info.p, B, ns);
    // time with the plan's profiler, keeping the caller's accumulated timings
    bool on = plan->prof.on;
    dpg_timings saved = plan->prof.acc;
    plan->prof.on = true;
    const int stage_of[3] = {DPG_STAGE_GRAM, DPG_STAGE_BMAT, DPG_STAGE_UPDATE};
    // candidate tile edges per role: 64 and 32 for all, and for the updates the
    // 128 x 64 panel-update tile
    const int cand[3][3] = {{64, 64, 64}, {32, 32, 32}, {64, 64, 128}};
    // tiles are timed on the dense G; then, for tensor-product cells in FP64,
    // the sum-factorised G is tried against it (choose_gram_path)
    B.sf_on = false;
    double ms[3][3] = {{0}};
    try {
      for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) B.tile[r] = cand[c][r];
        for (int pass = 0; pass < 2; ++pass) {  // pass 0 warms up
          plan->prof.acc = dpg_timings{};
          dpg::stage_factor(plan, B, *batch, 0, ns, w, nullptr, nullptr, st);
          dpg::stage_schur(plan, B, ns, w, st);
          plan->prof.collect(st, ns);
        }
        for (int r = 0; r < 3; ++r) ms[c][r] = plan->prof.acc.ms[stage_of[r]];
        // the update tile also decides how the panel solve runs (fused or not)
        ms[c][2] += plan->prof.acc.ms[DPG_STAGE_PANEL] + plan->prof.acc.ms[DPG_STAGE_SOLVE];
      }
      for (int r = 0; r < 2; ++r) B.tile[r] = ms[1][r] < ms[0][r] ? 32 : 64;
      int best = 0;
      for (int c = 1; c < 3; ++c)
        if (ms[c][2] < ms[best][2]) best = c;
      B.tile[2] = cand[best][2];
      if (B.sf_ok && plan->opt.precision == DPG_PREC_FP64) dpg::choose_gram_path(plan, B, *batch, ns, w, st);
    } catch (...) {
      B.merge_off = false;
      plan->prof.on = on;
      plan->prof.acc = saved;
      throw;
    }
    plan->prof.on = on;
    plan->prof.acc = saved;
  });
}

}  // extern "C"

struct dpg_graph {
  dpg_plan* plan = nullptr;
  std::vector<std::unique_ptr<dpg::DeviceBuffer>> bufs;
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> events;
  hipStream_t origin = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t bytes = 0;
  ~dpg_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto s : lanes) (void)hipStreamDestroy(s);
    if (origin) (void)hipStreamDestroy(origin);
  }
};

extern "C" {

int dpg_graph_create(dpg_plan* plan, const dpg_work_item* items, int nitems, int lanes, void* stream,
                     dpg_graph** out) {
  return dpg::guarded([&] {
    (void)stream;
    if (!plan || !out || (nitems > 0 && !items)) throw std::invalid_argument("NULL argument");
    if (nitems <= 0) throw std::invalid_argument("no work items");
    if (lanes < 1 || lanes > 16) throw std::invalid_argument("lanes must be in 1..16");
    *out = nullptr;
    for (int i = 0; i < nitems; ++i) dpg::check_batch(plan, items[i].bucket, &items[i].batch);
    std::unique_ptr<dpg_graph> g(new dpg_graph());
    g->plan = plan;
    // longest processing time first: items by decreasing flops, each to the
    // lane with the least work so far
    std::vector<int> order(nitems);
    std::vector<double> cost(nitems);
    for (int i = 0; i < nitems; ++i) {
      order[i] = i;
      cost[i] = (double)items[i].batch.ncells * plan->buckets[items[i].bucket].flops;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    lanes = std::min(lanes, nitems);
    std::vector<std::vector<int>> lane_items(lanes);
    std::vector<double> load(lanes, 0.0);
    for (int i : order) {
      int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      lane_items[l].push_back(i);
      load[l] += cost[i];
    }
    // per-lane workspace: the largest batch any of its items uses
    double budget = plan->opt.workspace_bytes > 0 ? plan->opt.workspace_bytes : 16.0 * (1ull << 30);
    std::vector<int> nb_of(nitems, 0);
    std::vector<dpg::DeviceBuffer*> ws(lanes), inf(lanes);
    for (int l = 0; l < lanes; ++l) {
      size_t need = 0;
      int cells = 1;
      for (int i : lane_items[l]) {
        const dpg::BucketPlan& B = plan->buckets[items[i].bucket];
        int64_t nb = (int64_t)(budget / lanes / (double)B.per_cell_bytes);
        if (plan->opt.max_batch_cells > 0) nb = std::min<int64_t>(nb, plan->opt.max_batch_cells);
        nb = std::max<int64_t>(1, std::min<int64_t>({nb, (int64_t)items[i].batch.ncells, 65535}));
        nb_of[i] = (int)nb;
        need = std::max(need, (size_t)B.per_cell_bytes * nb);
        cells = std::max(cells, (int)nb);
      }
      g->bufs.emplace_back(new dpg::DeviceBuffer());
      g->bufs.back()->ensure(std::max<size_t>(need, 256));
      ws[l] = g->bufs.back().get();
      g->bufs.emplace_back(new dpg::DeviceBuffer());
      g->bufs.back()->ensure(sizeof(int32_t) * cells);
      inf[l] = g->bufs.back().get();
      g->bytes += ws[l]->bytes + inf[l]->bytes;
    }
    HIP_CHECK(hipStreamCreateWithFlags(&g->origin, hipStreamNonBlocking));
    for (int l = 0; l < lanes; ++l) {
      hipStream_t s;
      HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      g->lanes.push_back(s);
    }
    for (int l = 0; l < 2 * lanes + 1; ++l) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      g->events.push_back(e);
    }
    // capture: fork every lane from the origin stream, join them back
    bool prof_on = plan->prof.on;
    plan->prof.on = false;  // no timing events inside the graph
    bool capturing = false;
    try {
      HIP_CHECK(hipStreamBeginCapture(g->origin, hipStreamCaptureModeRelaxed));
      capturing = true;
      HIP_CHECK(hipEventRecord(g->events[0], g->origin));
      for (int l = 0; l < lanes; ++l) {
        hipStream_t st = g->lanes[l];
        HIP_CHECK(hipStreamWaitEvent(st, g->events[0], 0));
        for (int i : lane_items[l]) {
          const dpg_work_item& it = items[i];
          const dpg::BucketPlan& B = plan->buckets[it.bucket];
          int C = it.batch.ncells;
          for (int c0 = 0; c0 < C; c0 += nb_of[i]) {
            int nb = std::min(nb_of[i], C - c0);
            int32_t* info = it.out.info ? it.out.info + c0 : (int32_t*)inf[l]->p;
            dpg::Work w = dpg::carve(ws[l]->p, info, B, nb);
            dpg::assemble_batch(plan, B, it.batch, c0, nb, w, it.out, st);
          }
        }
        HIP_CHECK(hipEventRecord(g->events[1 + l], st));
        HIP_CHECK(hipStreamWaitEvent(g->origin, g->events[1 + l], 0));
      }
      capturing = false;
      HIP_CHECK(hipStreamEndCapture(g->origin, &g->graph));
    } catch (...) {
      plan->prof.on = prof_on;
      if (capturing) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(g->origin, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
      }
      throw;
    }
    plan->prof.on = prof_on;
    HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}

int dpg_graph_launch(dpg_graph* graph, void* stream) {
  return dpg::guarded([&] {
    if (!graph) throw std::invalid_argument("graph is NULL");
    HIP_CHECK(hipGraphLaunch(graph->exec, (hipStream_t)stream));
  });
}

int dpg_graph_workspace(const dpg_graph* graph, size_t* bytes) {
  return dpg::guarded([&] {
    if (!graph || !bytes) throw std::invalid_argument("NULL argument");
    *bytes = graph->bytes;
  });
}

int dpg_graph_destroy(dpg_graph* graph) {
  return dpg::guarded([&] { delete graph; });
}

}  // extern "C"

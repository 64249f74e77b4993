// preprocess.hip -- the host-side preparation of the assembly inputs, on the
// device.
//   dpg_side_parities  side parities of a conforming mesh from its cell ->
//                      vertex ids, the rule of GlobalDofAssignment.cpp:160-185:
//                      +1 on the boundary; of two peer cells sharing a side the
//                      lower cell ID gets +1, the other -1. Sides are matched
//                      by their sorted vertex ids (stable radix sorts).
//   dpg_sort_cells     element-type bucketing (Solution.cpp:1162-1215 groups the
//                      cells by ElementType): a stable sort of the cells by a
//                      type key, the bucket boundaries, and
//   dpg_gather_cells   the gather of the cells' nodes, parities and
//                      coefficients into bucket-contiguous arrays.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>  // before rocPRIM, whose headers use memset
#include <rocprim/rocprim.hpp>
#include <stdexcept>
#include <string>

#include "../../include/dpg_capi.h"
#include "element.hpp"
#include "errors.hpp"

#define PP_CHECK(expr)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace dpg {
namespace {

template <class T>
struct Buf {
  T* p = nullptr;
  explicit Buf(size_t n) {
    if (n) PP_CHECK(hipMalloc(&p, n * sizeof(T)));
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

struct SideMap {
  int V = 0, S = 0, nps = 0;
  int nodes[6][4] = {};
};

SideMap side_map(int topo) {
  RefCell rc = make_ref_cell(topo);
  SideMap m;
  m.V = rc.nverts;
  m.S = rc.nsides;
  m.nps = rc.nodes_per_side;
  for (int s = 0; s < m.S; ++s)
    for (int k = 0; k < m.nps; ++k) m.nodes[s][k] = rc.side_nodes[(size_t)s * m.nps + k];
  return m;
}

// keys of every (cell, side): the side's vertex ids sorted, packed two per
// 64-bit word (hi: two smallest, lo: the others; lines use lo = 0)
__global__ void side_keys(const int32_t* __restrict__ cv, int ncells, SideMap m, uint64_t* __restrict__ hi,
                          uint64_t* __restrict__ lo, uint32_t* __restrict__ idx) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ncells * m.S) return;
  const int c = (int)(q / m.S), s = (int)(q % m.S);
  uint32_t v[4] = {0, 0, 0, 0};
  for (int k = 0; k < m.nps; ++k) v[k] = (uint32_t)cv[(int64_t)c * m.V + m.nodes[s][k]];
  for (int i = 1; i < m.nps; ++i)  // insertion sort of 2 or 4 ids
    for (int j = i; j > 0 && v[j - 1] > v[j]; --j) {
      const uint32_t t = v[j];
      v[j] = v[j - 1];
      v[j - 1] = t;
    }
  hi[q] = ((uint64_t)v[0] << 32) | v[1];
  lo[q] = m.nps == 4 ? (((uint64_t)v[2] << 32) | v[3]) : 0;
  idx[q] = (uint32_t)q;
}

__global__ void gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx, int64_t n,
                           uint64_t* __restrict__ dst) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[idx[q]];
}

// sorted (hi, lo) keys: a side shared by two cells appears twice in a row
__global__ void parities_from_sorted(const uint64_t* __restrict__ hi, const uint64_t* __restrict__ lo,
                                     const uint32_t* __restrict__ idx, int64_t n, int S, int8_t* __restrict__ par,
                                     int32_t* __restrict__ bad) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const bool prev = q > 0 && hi[q - 1] == hi[q] && lo[q - 1] == lo[q];
  const bool next = q + 1 < n && hi[q + 1] == hi[q] && lo[q + 1] == lo[q];
  const uint32_t me = idx[q];
  int8_t p = 1;  // boundary
  if (prev && next) {
    atomicAdd(bad, 1);  // three or more cells on one side: not a conforming mesh
  } else if (prev || next) {
    const uint32_t other = idx[prev ? q - 1 : q + 1];
    const uint32_t cm = me / S, co = other / S;
    if (cm == co) atomicAdd(bad, 1);  // a cell repeats a side
    p = cm < co ? 1 : -1;
  }
  par[me] = p;
}

template <class K>
void sort_pairs(const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, size_t n, unsigned bits, hipStream_t st) {
  size_t tb = 0;
  PP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, n, 0, bits, st));
  Buf<unsigned char> tmp(tb ? tb : 1);
  PP_CHECK(rocprim::radix_sort_pairs(tmp.p, tb, kin, kout, vin, vout, n, 0, bits, st));
}

__global__ void iota_keys(const int32_t* __restrict__ key, int n, uint64_t* __restrict__ k64,
                          uint32_t* __restrict__ idx, int32_t* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (key[i] < 0) atomicAdd(bad, 1);
  k64[i] = (uint64_t)(uint32_t)key[i];
  idx[i] = (uint32_t)i;
}

// offsets[b] = first sorted position of bucket b; first[b] its key
__global__ void bucket_bounds(const uint64_t* __restrict__ sk, int n, const uint32_t* __restrict__ flag_scan,
                              int32_t* __restrict__ offsets, int32_t* __restrict__ keys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || sk[i] != sk[i - 1]) {
    const uint32_t b = flag_scan[i];
    offsets[b] = i;
    keys[b] = (int32_t)sk[i];
  }
}

__global__ void head_flags(const uint64_t* __restrict__ sk, int n, uint32_t* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = (i == 0 || sk[i] != sk[i - 1]) ? 1u : 0u;
}

__global__ void gather_cells(const int32_t* __restrict__ perm, int n, int nd, const double* __restrict__ nodes,
                             double* __restrict__ nodes_out, int S, const int8_t* __restrict__ par,
                             int8_t* __restrict__ par_out, int nc, const double* __restrict__ coefs,
                             double* __restrict__ coefs_out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = nd + S + nc;
  if (q >= (int64_t)n * per) return;
  const int i = (int)(q / per), k = (int)(q % per), c = perm[i];
  if (c < 0 || c >= n) return;  // not a permutation of the cells: nothing read out of range
  if (k < nd) nodes_out[(int64_t)i * nd + k] = nodes[(int64_t)c * nd + k];
  else if (k < nd + S) par_out[(int64_t)i * S + (k - nd)] = par[(int64_t)c * S + (k - nd)];
  else coefs_out[(int64_t)i * nc + (k - nd - S)] = coefs[(int64_t)c * nc + (k - nd - S)];
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace dpg

extern "C" {

int dpg_side_parities(int topology, int ncells, const int32_t* cell_vertices, int8_t* parities, void* stream) {
  return dpg::guarded([&] {
    if (ncells < 0) throw std::invalid_argument("ncells < 0");
    if (ncells == 0) return;
    if (!cell_vertices || !parities) throw std::invalid_argument("NULL argument");
    const dpg::SideMap m = dpg::side_map(topology);
    const int64_t n = (int64_t)ncells * m.S;
    if (n > 0xFFFFFFFFll) throw std::invalid_argument("too many sides");
    hipStream_t st = (hipStream_t)stream;
    dpg::Buf<uint64_t> hi(n), lo(n), hs(n), ls(n);
    dpg::Buf<uint32_t> idx(n), i1(n), i2(n);
    dpg::Buf<int32_t> bad(1);
    PP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
    dpg::side_keys<<<dpg::nblk(n), 256, 0, st>>>(cell_vertices, ncells, m, hi.p, lo.p, idx.p);
    PP_CHECK(hipGetLastError());
    // lexicographic (hi, lo): stable sort by lo, then by hi
    dpg::sort_pairs(lo.p, ls.p, idx.p, i1.p, (size_t)n, 64, st);
    dpg::gather_u64<<<dpg::nblk(n), 256, 0, st>>>(hi.p, i1.p, n, hs.p);
    dpg::sort_pairs(hs.p, hi.p, i1.p, i2.p, (size_t)n, 64, st);  // hi now sorted, i2 the order
    dpg::gather_u64<<<dpg::nblk(n), 256, 0, st>>>(lo.p, i2.p, n, ls.p);
    dpg::parities_from_sorted<<<dpg::nblk(n), 256, 0, st>>>(hi.p, ls.p, i2.p, n, m.S, parities, bad.p);
    PP_CHECK(hipGetLastError());
    int32_t nbad = 0;
    PP_CHECK(hipMemcpyAsync(&nbad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PP_CHECK(hipStreamSynchronize(st));
    if (nbad) throw std::invalid_argument("not a conforming mesh: a side is shared by more than two cells or "
                                          "twice by one cell");
  });
}

int dpg_sort_cells(int ncells, const int32_t* type_key, int32_t* perm, int32_t* offsets, int32_t* bucket_keys,
                   void* stream) {
  return dpg::guarded_status([&] {
    if (ncells < 0) throw std::invalid_argument("ncells < 0");
    if (ncells == 0) return 0;
    if (!type_key || !perm || !offsets || !bucket_keys) throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    const int n = ncells;
    dpg::Buf<uint64_t> k(n), sk(n);
    dpg::Buf<uint32_t> idx(n), sidx(n), flags(n), scan(n);
    dpg::Buf<int32_t> bad(1);
    PP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), st));
    dpg::iota_keys<<<dpg::nblk(n), 256, 0, st>>>(type_key, n, k.p, idx.p, bad.p);
    dpg::sort_pairs(k.p, sk.p, idx.p, sidx.p, (size_t)n, 32, st);
    PP_CHECK(hipMemcpyAsync(perm, sidx.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    dpg::head_flags<<<dpg::nblk(n), 256, 0, st>>>(sk.p, n, flags.p);
    size_t tb = 0;
    PP_CHECK(rocprim::exclusive_scan(nullptr, tb, flags.p, scan.p, 0u, (size_t)n, rocprim::plus<uint32_t>(), st));
    dpg::Buf<unsigned char> tmp(tb ? tb : 1);
    PP_CHECK(rocprim::exclusive_scan(tmp.p, tb, flags.p, scan.p, 0u, (size_t)n, rocprim::plus<uint32_t>(), st));
    dpg::bucket_bounds<<<dpg::nblk(n), 256, 0, st>>>(sk.p, n, scan.p, offsets, bucket_keys);
    PP_CHECK(hipGetLastError());
    uint32_t last = 0, lastflag = 0;
    int32_t nbad = 0;
    PP_CHECK(hipMemcpyAsync(&last, scan.p + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PP_CHECK(hipMemcpyAsync(&lastflag, flags.p + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PP_CHECK(hipMemcpyAsync(&nbad, bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PP_CHECK(hipStreamSynchronize(st));
    if (nbad) throw std::invalid_argument("type keys must be >= 0");
    const int nb = (int)(last + lastflag);
    PP_CHECK(hipMemcpyAsync(offsets + nb, &n, sizeof(int32_t), hipMemcpyHostToDevice, st));
    PP_CHECK(hipStreamSynchronize(st));
    return nb;
  });
}

int dpg_gather_cells(int ncells, const int32_t* perm, int nodes_per_cell, const double* nodes, double* nodes_out,
                     int nsides, const int8_t* parities, int8_t* parities_out, int ncoefs, const double* coefs,
                     double* coefs_out, void* stream) {
  return dpg::guarded([&] {
    if (ncells < 0 || nodes_per_cell < 0 || nsides < 0 || ncoefs < 0) throw std::invalid_argument("negative size");
    if (ncells == 0) return;
    if (!perm || (nodes_per_cell && (!nodes || !nodes_out)) || (nsides && (!parities || !parities_out)) ||
        (ncoefs && (!coefs || !coefs_out)))
      throw std::invalid_argument("NULL argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t work = (int64_t)ncells * (nodes_per_cell + nsides + ncoefs);
    dpg::gather_cells<<<dpg::nblk(work), 256, 0, st>>>(perm, ncells, nodes_per_cell, nodes, nodes_out, nsides,
                                                       parities, parities_out, ncoefs, coefs, coefs_out);
    PP_CHECK(hipGetLastError());
  });
}

}  // extern "C"

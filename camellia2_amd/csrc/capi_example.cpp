// capi_example.cpp -- a C++ caller of the C ABI, the shape of the drop-in
// behind TBF::localStiffnessMatrixAndRHS (BF.h:122-125, called at
// Solution.cpp:1226): build the term tables of a formulation, plan one
// element type, hand a batch of cells (FieldContainer layouts) to the device,
// assemble K, rhs and the trace Schur complement, read them back.
//
//   dpg_capi_example <nx> <ny> <out.bin>
//
// 2D Poisson ultraweak DPG, graph-norm test inner product, H1 order 3 (fields
// p = 2), test order 4, nx x ny quads jittered by 10% of h, an 8-mode sine
// load. Writes K (C, M, M), rhs (C, M), schur (C, T, T) as raw doubles to
// out.bin and prints the sizes, the non-SPD count and the largest asymmetry
// of K. Exit status 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/dpg_capi.h"

#define CHECK_HIP(x)                                                               \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                 \
      std::exit(2);                                                                \
    }                                                                              \
  } while (0)
#define CHECK_DPG(x)                                                               \
  do {                                                                             \
    int rc_ = (x);                                                                 \
    if (rc_ < 0) {                                                                 \
      std::fprintf(stderr, "%s failed (%d): %s\n", #x, rc_, dpg_last_error());     \
      std::exit(3);                                                                \
    }                                                                              \
  } while (0)

template <class T>
T* device_copy(const std::vector<T>& h) {
  T* d = nullptr;
  CHECK_HIP(hipMalloc(&d, sizeof(T) * h.size()));
  CHECK_HIP(hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return d;
}

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s nx ny out.bin\n", argv[0]);
    return 1;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]);
  std::printf("%s\n", dpg_version());

  // the formulation's term tables (PoissonFormulation ULTRAWEAK + graphNorm + (f, v))
  dpg_form_desc form;
  CHECK_DPG(dpg_form_poisson_uw(2, &form));
  dpg_bucket_desc bucket = {DPG_QUAD, 3, 4, 0};
  dpg_bucket_info info;
  CHECK_DPG(dpg_bucket_sizes(&form, &bucket, &info));
  const int C = nx * ny, M = info.M, T = info.T;

  // the batch: physical nodes (C, 4, 2) and side parities (C, 4), as the
  // BasisCaches hold them, and the load
  std::vector<double> nodes((size_t)C * 4 * 2);
  std::vector<int8_t> par((size_t)C * 4);
  CHECK_DPG(dpg_mesh_quad(nx, ny, 0.10, 0xC0FFEE + 1, nodes.data(), par.data()));
  dpg_load_desc load;
  CHECK_DPG(dpg_load_sine_modes(2, 1, 8, 6, 0xC0FFEE + 1 + 100, &load));

  // one plan per element type (the BasisCache role): device tables built once
  dpg_options opt = {DPG_PREC_FP64, 0, 0, 0, 0.0};
  dpg_plan* plan = nullptr;
  CHECK_DPG(dpg_plan_create(&form, &bucket, 1, &opt, &plan));

  double *dK = nullptr, *drhs = nullptr, *dS = nullptr, *dg = nullptr;
  int32_t* dinfo = nullptr;
  CHECK_HIP(hipMalloc(&dK, sizeof(double) * C * M * M));
  CHECK_HIP(hipMalloc(&drhs, sizeof(double) * C * M));
  CHECK_HIP(hipMalloc(&dS, sizeof(double) * C * T * T));
  CHECK_HIP(hipMalloc(&dg, sizeof(double) * C * T));
  CHECK_HIP(hipMalloc(&dinfo, sizeof(int32_t) * C));
  double* dnodes = device_copy(nodes);
  int8_t* dpar = device_copy(par);
  hipStream_t st;
  CHECK_HIP(hipStreamCreate(&st));

  dpg_batch batch = {C, 0, dnodes, dpar, nullptr, load};
  dpg_outputs out = {dK, drhs, nullptr, dS, dg, dinfo, nullptr, nullptr};
  const int nbad = dpg_assemble(plan, 0, &batch, &out, st);  // synchronises (info requested)
  CHECK_DPG(nbad);

  std::vector<double> K((size_t)C * M * M), rhs((size_t)C * M), S((size_t)C * T * T);
  CHECK_HIP(hipMemcpy(K.data(), dK, sizeof(double) * K.size(), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(rhs.data(), drhs, sizeof(double) * rhs.size(), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(S.data(), dS, sizeof(double) * S.size(), hipMemcpyDeviceToHost));
  double asym = 0.0;
  for (int c = 0; c < C; ++c)
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < i; ++j)
        asym = std::fmax(asym, std::fabs(K[((size_t)c * M + i) * M + j] - K[((size_t)c * M + j) * M + i]));
  std::printf("cells %d N %d M %d T %d non-SPD %d max |K - K^T| %.3e\n", C, info.N, M, T, nbad, asym);

  FILE* fo = std::fopen(argv[3], "wb");
  if (!fo) return 4;
  std::fwrite(K.data(), sizeof(double), K.size(), fo);
  std::fwrite(rhs.data(), sizeof(double), rhs.size(), fo);
  std::fwrite(S.data(), sizeof(double), S.size(), fo);
  std::fclose(fo);

  CHECK_HIP(hipStreamDestroy(st));
  for (void* p : {(void*)dK, (void*)drhs, (void*)dS, (void*)dg, (void*)dinfo, (void*)dnodes, (void*)dpar})
    CHECK_HIP(hipFree(p));
  CHECK_DPG(dpg_plan_destroy(plan));
  return (nbad == 0 && asym == 0.0) ? 0 : 5;
}

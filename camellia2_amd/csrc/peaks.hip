// dpg_peaks: sustained rates of the GPU it runs on, the measured counterparts
// of the published peaks a roofline is drawn against.
//   fp64 MFMA    v_mfma_f64_16x16x4 on register operands only (NCH independent
//                accumulator chains per wave), 1..8 waves per SIMD
//   fp64 vector  independent v_fma_f64 chains on register operands
//   HBM          streaming copy (read + write) and streaming read of 2 GiB
// Best of 5 timed launches each (HIP events). Prints one JSON line.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                      \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

typedef double d4 __attribute__((ext_vector_type(4)));

// out holds gridDim.x * blockDim.x * NCH * 4 doubles
template <int NCH>
__global__ __launch_bounds__(256) void mfma_f64(double* __restrict__ out, int iters, double a0, double b0) {
  d4 acc[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
  const double a = a0 + threadIdx.x * 1e-3, b = b0 - threadIdx.x * 1e-3;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  const size_t base = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * NCH * 4;
#pragma unroll
  for (int c = 0; c < NCH; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) out[base + c * 4 + v] = acc[c][v];
}

// out holds gridDim.x * blockDim.x doubles
__global__ __launch_bounds__(256) void fma_f64(double* __restrict__ out, int iters, double b, double c) {
  constexpr int CH = 8;
  double x[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) x[k] = 1.0 + 1e-3 * (threadIdx.x + k);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < CH; ++k) x[k] = fma(x[k], b, c);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < CH; ++k) s += x[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy_f64(double2* __restrict__ dst, const double2* __restrict__ src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// out holds gridDim.x * blockDim.x doubles
__global__ __launch_bounds__(256) void read_f64(const double2* __restrict__ src, size_t n, double* __restrict__ out) {
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = src[i];
    s += v.x + v.y;
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <class F>
static float best_ms(F launch) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();  // warm-up
  CHECK(hipGetLastError());
  float best = 1e30f;
  for (int r = 0; r < 5; ++r) {
    CHECK(hipEventRecord(e0));
    launch();
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) best = ms;
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return best;
}

int main() {
  int dev = 0;
  CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, dev));
  const int cus = prop.multiProcessorCount;
  const int max_blocks = cus * 8;  // up to 8 workgroups of 4 waves per CU
  double* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)max_blocks * 256 * 16 * 4 * sizeof(double)));

  // fp64 MFMA: 2*16*16*4 flops per instruction and wave
  double mfma_best = 0.0;
  int mfma_wps = 0;
  for (int per_cu : {1, 2, 4, 8}) {
    const int blocks = cus * per_cu, iters = 20000;
    // 16 chains per wave: fewer sustain less here (4: 49, 8: 49, 16: 56 TFLOP/s)
    const float ms = best_ms([&] { mfma_f64<16><<<blocks, 256>>>(out, iters / 4, 1.0, 2.0); });
    const double tf = (double)blocks * 4 * (iters / 4) * 16 * 2.0 * 16 * 16 * 4 / (ms * 1e-3) / 1e12;
    if (tf > mfma_best) {
      mfma_best = tf;
      mfma_wps = per_cu;  // waves per SIMD: 4 waves per workgroup over 4 SIMDs
    }
  }
  // fp64 vector FMA: 2 flops per lane and instruction
  double fma_best = 0.0;
  for (int per_cu : {2, 4, 8}) {
    const int blocks = cus * per_cu, iters = 20000;
    const float ms = best_ms([&] { fma_f64<<<blocks, 256>>>(out, iters, 0.999999, 1e-7); });
    const double tf = (double)blocks * 256 * iters * 8 * 2.0 / (ms * 1e-3) / 1e12;
    if (tf > fma_best) fma_best = tf;
  }
  // HBM: 2 GiB source and destination
  const size_t bytes = (size_t)2 << 30, n2 = bytes / sizeof(double2);
  double2 *src = nullptr, *dst = nullptr;
  CHECK(hipMalloc(&src, bytes));
  CHECK(hipMalloc(&dst, bytes));
  CHECK(hipMemset(src, 0, bytes));
  const int sblocks = cus * 8;
  const float copy_ms = best_ms([&] { copy_f64<<<sblocks, 256>>>(dst, src, n2); });
  const float read_ms = best_ms([&] { read_f64<<<sblocks, 256>>>(src, n2, out); });
  const double copy_gbs = 2.0 * bytes / (copy_ms * 1e-3) / 1e9, read_gbs = (double)bytes / (read_ms * 1e-3) / 1e9;
  std::printf("{\"device\": \"%s\", \"gcn_arch\": \"%s\", \"cus\": %d, \"clock_mhz\": %d, "
              "\"fp64_mfma_tflops\": %.2f, \"fp64_mfma_waves_per_simd\": %d, \"fp64_vector_fma_tflops\": %.2f, "
              "\"hbm_copy_gbs\": %.1f, \"hbm_read_gbs\": %.1f, \"bytes\": %zu}\n",
              prop.name, prop.gcnArchName, cus, prop.clockRate / 1000, mfma_best, mfma_wps, fma_best, copy_gbs,
              read_gbs, bytes);
  CHECK(hipFree(src));
  CHECK(hipFree(dst));
  CHECK(hipFree(out));
  return 0;
}

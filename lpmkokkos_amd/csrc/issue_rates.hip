// Instruction issue costs on one GPU: cycles per wave64 instruction per SIMD for the
// operations the pair kernels are made of (FP32/FP64 FMA, v_rcp_f32/f64, the FP64<->FP32
// conversions, FP32 MFMA), alone and mixed.  These set the attainable rate of the pair
// kernels: a pair kernel's instruction counts (rocprofv3 SQ_INSTS_VALU*) times these costs,
// over the SIMD-cycles of a launch, is the fraction of the issue limit it reaches.
//
// Each kernel runs register-only loops of independent chains, 8 waves per SIMD, and writes one
// value per thread to a buffer sized to its grid.  Cycles are at the device's reported clock.
//
//     make -C lpmkokkos_amd/csrc issue_rates && ./lpmkokkos_amd/issue_rates
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

#define CK(x)                                                                           \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                     \
    }                                                                                   \
  } while (0)

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kIters = 16384;
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_fma32(float* out, float seed) {
  float x[8];
  for (int k = 0; k < 8; ++k) x[k] = seed + threadIdx.x * 1e-3f + k;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], 0.999f, 1e-3f);
  float s = 0;
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_rcp32(float* out, float seed) {
  float x[8];
  for (int k = 0; k < 8; ++k) x[k] = seed + threadIdx.x * 1e-3f + k;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = __builtin_amdgcn_rcpf(x[k]);
  float s = 0;
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// 8 v_rcp_f32 + NF FMA per iteration, independent chains
template <int NF>
__global__ __launch_bounds__(kThreads) void k_rcp32_fma(float* out, float seed) {
  float x[8], y[NF];
  for (int k = 0; k < 8; ++k) x[k] = seed + threadIdx.x * 1e-3f + k;
  for (int k = 0; k < NF; ++k) y[k] = seed + k;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = __builtin_amdgcn_rcpf(x[k]);
#pragma unroll
    for (int k = 0; k < NF; ++k) y[k] = fmaf(y[k], 0.999f, 1e-3f);
  }
  float s = 0;
  for (int k = 0; k < 8; ++k) s += x[k];
  for (int k = 0; k < NF; ++k) s += y[k];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_mfma4(float* out, float seed) {
  v4f c[8];
  for (int k = 0; k < 8; ++k) c[k] = v4f{seed, seed, seed, (float)k};
  const float a = 0.5f + threadIdx.x * 1e-4f;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, 1e-3f, c[k], 0, 0, 0);
  float s = 0;
  for (int k = 0; k < 8; ++k) s += c[k][0] + c[k][1] + c[k][2] + c[k][3];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// 8 MFMA 4x4x1 + 16 FMA per iteration
__global__ __launch_bounds__(kThreads) void k_mfma4_fma(float* out, float seed) {
  v4f c[8];
  float y[16];
  for (int k = 0; k < 8; ++k) c[k] = v4f{seed, seed, seed, (float)k};
  for (int k = 0; k < 16; ++k) y[k] = seed + k + threadIdx.x * 1e-3f;
  const float a = 0.5f + threadIdx.x * 1e-4f;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, 1e-3f, c[k], 0, 0, 0);
#pragma unroll
    for (int k = 0; k < 16; ++k) y[k] = fmaf(y[k], 0.999f, 1e-3f);
  }
  float s = 0;
  for (int k = 0; k < 8; ++k) s += c[k][0] + c[k][1] + c[k][2] + c[k][3];
  for (int k = 0; k < 16; ++k) s += y[k];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_mfma16(float* out, float seed) {
  v4f c[4];
  for (int k = 0; k < 4; ++k) c[k] = v4f{seed, seed, seed, (float)k};
  const float a = 0.5f + threadIdx.x * 1e-4f;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, 1e-3f, c[k], 0, 0, 0);
  float s = 0;
  for (int k = 0; k < 4; ++k) s += c[k][0] + c[k][1] + c[k][2] + c[k][3];
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_fma64(float* out, float seed) {
  double x[8];
  for (int k = 0; k < 8; ++k) x[k] = seed + threadIdx.x * 1e-3 + k;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = fma(x[k], 0.999, 1e-3);
  double s = 0;
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * kThreads + threadIdx.x] = (float)s;
}

__global__ __launch_bounds__(kThreads) void k_rcp64(float* out, float seed) {
  double x[8];
  for (int k = 0; k < 8; ++k) x[k] = seed + threadIdx.x * 1e-3 + k;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = __builtin_amdgcn_rcp(x[k]);
  double s = 0;
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * kThreads + threadIdx.x] = (float)s;
}

// FP64 -> FP32 -> FP64 round trip and an FP64 add per chain step (the add keeps the
// conversions from folding)
__global__ __launch_bounds__(kThreads) void k_cvt(float* out, float seed) {
  double x[8];
  for (int k = 0; k < 8; ++k) x[k] = seed + threadIdx.x * 1e-3 + k;
  for (int it = 0; it < kIters; ++it)
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (double)(float)x[k] + 1e-3;
  double s = 0;
  for (int k = 0; k < 8; ++k) s += x[k];
  out[blockIdx.x * kThreads + threadIdx.x] = (float)s;
}

}  // namespace

int main() {
  int dev = 0;
  CK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  CK(hipGetDeviceProperties(&p, dev));
  const int cus = p.multiProcessorCount;
  const double clk = p.clockRate * 1e3;
  std::printf("%s: %d CUs, clock %.0f MHz; 8 waves per SIMD; cycles per wave64 instruction per SIMD\n", p.gcnArchName,
              cus, clk / 1e6);
  const int blocks = cus * 8;  // 8 workgroups of 4 waves per CU
  float* out = nullptr;
  CK(hipMalloc(&out, (size_t)blocks * kThreads * sizeof(float)));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  // insts = wave instructions per loop iteration (the unit the cost is quoted per)
  auto run = [&](const char* name, void (*k)(float*, float), double insts, const char* unit) {
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, 0, out, 1.5f);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, 0, out, 1.5f);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double wave_insts = blocks * 4.0 * kIters * insts;
    const double simd_cycles = ms * 1e-3 * clk * cus * 4;
    std::printf("  %-34s %7.3f ms  %6.2f cycles per %s\n", name, ms, simd_cycles / wave_insts, unit);
    std::fflush(stdout);
  };
  for (int w = 0; w < 20; ++w) hipLaunchKernelGGL(k_fma32, dim3(blocks), dim3(kThreads), 0, 0, out, 1.5f);  // clocks up
  CK(hipDeviceSynchronize());
  run("FP32 FMA", k_fma32, 8, "FMA");
  run("v_rcp_f32", k_rcp32, 8, "rcp");
  run("8 v_rcp_f32 + 8 FMA", k_rcp32_fma<8>, 8, "(rcp + FMA)");
  run("8 v_rcp_f32 + 32 FMA", k_rcp32_fma<32>, 8, "(rcp + 4 FMA)");
  run("MFMA f32 4x4x1 (16 blocks)", k_mfma4, 8, "MFMA");
  run("MFMA f32 4x4x1 + 2 FMA", k_mfma4_fma, 8, "(MFMA + 2 FMA)");
  run("MFMA f32 16x16x4", k_mfma16, 4, "MFMA");
  run("FP64 FMA", k_fma64, 8, "FMA");
  run("v_rcp_f64", k_rcp64, 8, "rcp");
  run("FP64->FP32->FP64 + FP64 add", k_cvt, 8, "(2 conversions + add)");
  run("FP32 FMA (again)", k_fma32, 8, "FMA");
  CK(hipFree(out));
  return 0;
}

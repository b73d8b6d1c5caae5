// mbr_calib.hip — calibration kernels for reading rocprofv3 counters and
// HIP-event times of the render kernels against the MI355X vector peak.
//
// Each kernel runs a known number of VALU instructions per wave: an unrolled
// loop of independent FMA chains (FP32 or FP64), or of v_rsq_f32
// transcendentals, over enough waves to fill every SIMD.  The program
// prints, per kernel, the instructions issued (from the source's count),
// the HIP-event time, and the implied issue rate, as one JSON line each;
// under `rocprofv3 --pmc SQ_INSTS_VALU ...` the same launches give the
// counter readings to set beside those counts.
//
//   mbr_calib [waves_per_simd]      (default 8)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kChains = 8;    // independent chains per lane (hides latency)
constexpr int kUnroll = 16;   // FMAs per chain per loop pass

#define CHECK(x)                                                                             \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) {                                                                  \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                          \
      std::exit(1);                                                                          \
    }                                                                                        \
  } while (0)

// kChains x kUnroll FMAs per pass, `passes` passes: the loop body is all
// VALU except the scalar loop counter.
template <class T>
__global__ void __launch_bounds__(256) k_fma(T *out, int passes, T a, T b) {
  T acc[kChains];
  for (int c = 0; c < kChains; c++) acc[c] = (T)(threadIdx.x + c);
  for (int i = 0; i < passes; i++) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
#pragma unroll
      for (int c = 0; c < kChains; c++) acc[c] = acc[c] * a + b;
  }
  T s = 0;
  for (int c = 0; c < kChains; c++) s += acc[c];
  if (s == (T)-12345.0) out[blockIdx.x] = s;  // keeps the chains live; never taken for these inputs
}

// kChains x kUnroll v_rsq_f32 per pass.
__global__ void __launch_bounds__(256) k_rsq(float *out, int passes) {
  float acc[kChains];
  for (int c = 0; c < kChains; c++) acc[c] = 1.0f + (float)(threadIdx.x + c);
  for (int i = 0; i < passes; i++) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
#pragma unroll
      for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_rsqf(acc[c]);
  }
  float s = 0;
  for (int c = 0; c < kChains; c++) s += acc[c];
  if (s == -12345.0f) out[blockIdx.x] = s;
}

template <class K>
double time_ms(K launch) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();  // warm-up
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0, nullptr));
  launch();
  CHECK(hipEventRecord(e1, nullptr));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return ms;
}

}  // namespace

int main(int argc, char **argv) {
  const int waves_per_simd = argc > 1 ? std::atoi(argv[1]) : 8;
  int cus = 0, clock_khz = 0;
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  CHECK(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, 0));
  // 4 SIMDs per CU, 4 waves per 256-thread workgroup
  const int blocks = cus * 4 * waves_per_simd / 4;
  const long long waves = (long long)blocks * 4;
  const int passes = 2048;
  const double per_wave = (double)passes * kChains * kUnroll;  // VALU ops of the loop per wave
  float *out = nullptr;
  CHECK(hipMalloc((void **)&out, (size_t)blocks * sizeof(double)));
  std::printf("{\"device_cus\": %d, \"clock_mhz_attr\": %.0f, \"waves_per_simd\": %d, \"waves\": %lld}\n", cus,
              clock_khz / 1e3, waves_per_simd, waves);
  struct Row {
    const char *name;
    double ms;
    int flops_per_op;
  } rows[3];
  rows[0] = {"k_fma<float>", time_ms([&] {
               hipLaunchKernelGGL(k_fma<float>, dim3(blocks), dim3(256), 0, nullptr, out, passes, 0.999f, 1e-3f);
             }), 2};
  rows[1] = {"k_fma<double>", time_ms([&] {
               hipLaunchKernelGGL(k_fma<double>, dim3(blocks), dim3(256), 0, nullptr, (double *)out, passes, 0.999,
                                  1e-3);
             }), 2};
  rows[2] = {"k_rsq", time_ms([&] {
               hipLaunchKernelGGL(k_rsq, dim3(blocks), dim3(256), 0, nullptr, out, passes);
             }), 1};
  for (const Row &r : rows) {
    const double insts = per_wave * (double)waves;         // wave-level VALU instructions of the loops
    const double simd_cycles = insts * 4.0;                 // a wave64 op occupies a 16-lane SIMD for 4 cycles
    const double ghz = simd_cycles / (cus * 4.0) / (r.ms * 1e-3) / 1e9;  // clock implied if issue never stalls
    const double tflops = insts * 64.0 * r.flops_per_op / (r.ms * 1e-3) / 1e12;
    std::printf("{\"kernel\": \"%s\", \"ms\": %.4f, \"valu_insts_wave_level\": %.6e, "
                "\"implied_clock_ghz_at_4_cycles_per_op\": %.3f, \"lane_ops_per_s\": %.4e, \"tflops\": %.2f}\n",
                r.name, r.ms, insts, ghz, insts * 64.0 / (r.ms * 1e-3), tflops);
  }
  CHECK(hipFree(out));
  return 0;
}

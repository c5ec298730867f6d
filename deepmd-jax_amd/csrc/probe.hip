// Issue-overlap probe for gfx950 (built as its own library, libdpmd_probe.so; not part of the
// energy/force path).  It times a chain of matrix-core instructions and a chain of FP32 VALU
// FMAs alone, together in one wave, and in separate waves of one block, so the cost model of
// the embedding kernels (FP32 MFMA time + VALU time, or their maximum) can be read off the
// hardware: if the two kinds of work overlap, "separate waves" takes about the longer of the
// two alone; if they share an issue resource it takes about their sum.
//
// Every wave runs the same instruction counts in every configuration; each wave writes one
// float per lane into an output buffer of exactly grid x block floats.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kMfmaChains = 4;   // independent accumulators per wave
constexpr int kValuChains = 8;   // independent FMA chains per lane
constexpr int kValuRepeat = 4;   // FMAs per chain per iteration

enum ProbeKind { MFMA_F32 = 0, MFMA_BF16 = 1 };

// AGPR forms (KIND + 2): the same instructions with the accumulators pinned to AccVGPRs by the
// asm constraint (the builtins let the compiler choose ArchVGPRs, which it does here).  The
// chains read their own accumulator back only as the next instruction's SrcC of the same
// opcode, which the hardware forwards; the VALU reads them after the loop, past the wait below.
enum ProbeKindAcc { MFMA_F32_AGPR = 2, MFMA_BF16_AGPR = 3 };

template <int KIND>
__device__ __forceinline__ void mfma_step(floatx4 (&acc)[kMfmaChains], float a, float b, bf16x8 ah, bf16x8 bh) {
#pragma unroll
  for (int k = 0; k < kMfmaChains; ++k) {
    if constexpr (KIND == MFMA_F32) acc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[k], 0, 0, 0);
    else if constexpr (KIND == MFMA_BF16) acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[k], 0, 0, 0);
    else if constexpr (KIND == MFMA_F32_AGPR)
      asm volatile("v_mfma_f32_16x16x4_f32 %0, %1, %2, %0" : "+a"(acc[k]) : "v"(a), "v"(b));
    else
      asm volatile("v_mfma_f32_16x16x32_bf16 %0, %1, %2, %0" : "+a"(acc[k]) : "v"(ah), "v"(bh));
  }
}

__device__ __forceinline__ void valu_step(float (&v)[kValuChains], float x, float y) {
#pragma unroll
  for (int r = 0; r < kValuRepeat; ++r)
#pragma unroll
    for (int k = 0; k < kValuChains; ++k) v[k] = __builtin_fmaf(v[k], x, y);
}

// DO_M / DO_V: the wave runs the matrix / VALU chain.  SPLIT (8-wave blocks, each wave with the
// same counts as alone): 1 = waves 0..3 run the matrix chain and waves 4..7 the VALU chain,
// 2 = even waves the matrix chain and odd waves the VALU chain.  Which SIMD a wave lands on is
// the dispatcher's choice; the two splits pair the kinds differently, so the faster of them is
// the one that gives every SIMD both kinds.
template <int KIND, bool DO_M, bool DO_V, int SPLIT>
__global__ void __launch_bounds__(512) issue_probe_kernel(float* out, int iters, float seed) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const bool m_wave = SPLIT == 1 ? wave < 4 : (wave & 1) == 0;
  const bool run_m = SPLIT ? m_wave : DO_M;
  const bool run_v = SPLIT ? !m_wave : DO_V;
  const float a = seed + 1e-3f * (float)(tid & 63);
  const float b = 1.f - 1e-3f * (float)(tid & 15);
  bf16x8 ah, bh;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    ah[j] = (__bf16)(a + 1e-2f * (float)j);
    bh[j] = (__bf16)(b - 1e-2f * (float)j);
  }
  floatx4 acc[kMfmaChains];
#pragma unroll
  for (int k = 0; k < kMfmaChains; ++k) acc[k] = floatx4{0.f, 0.f, 0.f, 0.f};
  float v[kValuChains];
#pragma unroll
  for (int k = 0; k < kValuChains; ++k) v[k] = a + (float)k;
  const float x = 0.999f, y = 1e-4f * b;
  for (int it = 0; it < iters; ++it) {
    if (run_m) mfma_step<KIND>(acc, a, b, ah, bh);
    if (run_v) valu_step(v, x, y);
  }
  // the last matrix results are read by VALU moves: wait out the longest MFMA-write ->
  // VALU-read distance (the compiler does not see into the asm forms)
  if constexpr (KIND >= MFMA_F32_AGPR)
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7\n\ts_nop 7\n\ts_nop 7"
                 : "+a"(acc[0]), "+a"(acc[1]), "+a"(acc[2]), "+a"(acc[3]));
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kMfmaChains; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
#pragma unroll
  for (int k = 0; k < kValuChains; ++k) s += v[k];
  out[(size_t)blockIdx.x * blockDim.x + tid] = s;
}

template <int KIND, bool DO_M, bool DO_V, int SPLIT>
hipError_t launch(float* out, int blocks, int iters) {
  hipLaunchKernelGGL((issue_probe_kernel<KIND, DO_M, DO_V, SPLIT>), dim3(blocks), dim3(SPLIT ? 512 : 256), 0, 0,
                     out, iters, 0.5f);
  return hipGetLastError();
}

hipError_t launch_case(int which, float* out, int blocks, int iters) {
  switch (which) {
    case 0: return launch<MFMA_F32, true, false, 0>(out, blocks, iters);
    case 1: return launch<MFMA_F32, false, true, 0>(out, blocks, iters);
    case 2: return launch<MFMA_F32, true, true, 0>(out, blocks, iters);
    case 3: return launch<MFMA_F32, false, false, 1>(out, blocks, iters);
    case 4: return launch<MFMA_BF16, true, false, 0>(out, blocks, iters);
    case 5: return launch<MFMA_BF16, true, true, 0>(out, blocks, iters);
    case 6: return launch<MFMA_BF16, false, false, 1>(out, blocks, iters);
    case 7: return launch<MFMA_F32, false, false, 2>(out, blocks, iters);
    case 8: return launch<MFMA_BF16, false, false, 2>(out, blocks, iters);
    case 9: return launch<MFMA_F32_AGPR, true, false, 0>(out, blocks, iters);
    case 10: return launch<MFMA_F32_AGPR, true, true, 0>(out, blocks, iters);
    case 11: return launch<MFMA_F32_AGPR, false, false, 2>(out, blocks, iters);
    case 12: return launch<MFMA_BF16_AGPR, true, false, 0>(out, blocks, iters);
    case 13: return launch<MFMA_BF16_AGPR, true, true, 0>(out, blocks, iters);
    case 14: return launch<MFMA_BF16_AGPR, false, false, 2>(out, blocks, iters);
    case 15: return launch<MFMA_F32_AGPR, false, false, 1>(out, blocks, iters);
    case 16: return launch<MFMA_BF16_AGPR, false, false, 1>(out, blocks, iters);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

// which: 0 FP32 MFMA alone, 1 VALU alone, 2 FP32 MFMA + VALU in one wave, 3 FP32 MFMA and VALU
// in separate waves of one block (halves), 4 BF16 MFMA alone, 5 BF16 MFMA + VALU in one wave,
// 6 BF16 MFMA and VALU in separate waves (halves), 7 and 8: cases 3 and 6 split by wave parity;
// 9-14: cases 0, 2, 7, 4, 5, 8 with the accumulators in AccVGPRs; 15, 16: cases 3, 6 likewise.
// blocks: grid size (2 x 256 CUs puts two blocks on every CU).  Writes the average kernel time
// over `reps` launches (after one warm-up) to *ms.  Returns 0, or -1 with a message on stderr.
extern "C" int dpmd_probe_issue(int which, int blocks, int iters, int reps, float* ms) {
  if (which < 0 || which > 16 || blocks <= 0 || blocks > (1 << 16) || iters <= 0 || reps <= 0 || !ms) {
    std::fprintf(stderr, "dpmd_probe_issue: bad arguments\n");
    return -1;
  }
  float* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -1;
  const size_t n = (size_t)blocks * 512;  // the larger block size
  if (hipMalloc((void**)&out, n * sizeof(float)) != hipSuccess) {
    std::fprintf(stderr, "dpmd_probe_issue: hipMalloc failed\n");
    return -1;
  }
  if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
      launch_case(which, out, blocks, iters) == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
      hipEventRecord(e0, 0) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; ++r) ok = launch_case(which, out, blocks, iters) == hipSuccess;
    float t = 0.f;
    if (ok && hipEventRecord(e1, 0) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&t, e0, e1) == hipSuccess) {
      *ms = t / (float)reps;
      rc = 0;
    }
  }
  if (rc) std::fprintf(stderr, "dpmd_probe_issue: %s\n", hipGetErrorString(hipGetLastError()));
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(out);
  return rc;
}

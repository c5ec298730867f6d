// Per-type fitting net forward + backward as six MFMA GEMMs with fused epilogues.
//
// Restates deepmd_jax/utils.py:153-166 (fitting_net) for three equal hidden widths and
// the energy head of deepmd_jax/dpmodel.py:79-81:
//   y1 = tanh(G K0 + b0); y2 = tanh(y1 K1 + b1)*dt1 + y1; y3 = tanh(y2 K2 + b2)*dt2 + y2
//   e  = y3 k3 + b3
// and its reverse pass for dE/dG (no weight gradients):
//   dz3 = k3 dt2 (1-t3^2); dy2 = k3 + dz3 K2^T; dz2 = dy2 dt1 (1-t2^2)
//   dy1 = dy2 + dz2 K1^T;  dz1 = dy1 (1-y1^2);  dG = dz1 K0^T
// Rows are atoms of one type (padded to 64).  B is given transposed (N x K row-major).
#include "internal.h"
#include <algorithm>
#include <string.h>

namespace dpmd {

typedef float floatx4 __attribute__((ext_vector_type(4)));

enum FitEpi { EPI_TANH = 0, EPI_RES = 1, EPI_E3 = 2, EPI_BWD2 = 3, EPI_BWD1 = 4, EPI_STORE = 5 };

struct GemmArgs {
  const float* A;   // M x K, row stride lda
  int lda;
  const float* Bt;  // N x K, row stride ldb
  int ldb;
  int M, N, K;
  float* C;         // primary output, row stride ldc
  int ldc;
  float* C2;        // secondary output (same stride)
  const float* aux1;
  const float* aux2;
  const float* bias;
  const float* dt;
  const float* k3;
  float b3;
  float* eout;      // per-row energy
};

constexpr int KC = 16;
constexpr int LDS_STRIDE = KC + 4;

template <int NT, int EPI>
__global__ void __launch_bounds__(256, 2) gemm_f32_kernel(const GemmArgs g) {
  __shared__ float As[64 * LDS_STRIDE];
  __shared__ float Bs[16 * NT * LDS_STRIDE];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int q = lane >> 4, n = lane & 15;
  const int m0 = blockIdx.x * 64;
  const int n0 = blockIdx.y * 16 * NT;
  constexpr int BN = 16 * NT;
  constexpr int BV = BN * KC / 4;  // float4 per B chunk

  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};

  // staging registers
  float4 ra;
  float4 rb[(BV + 255) / 256];
  auto load_chunk = [&](int k0) {
    {
      const int row = tid / 4, c4 = tid % 4;
      ra = *reinterpret_cast<const float4*>(g.A + (size_t)(m0 + row) * g.lda + k0 + 4 * c4);
    }
#pragma unroll
    for (int v = 0; v < (BV + 255) / 256; ++v) {
      const int idx = tid + 256 * v;
      if (idx < BV) {
        const int row = idx / 4, c4 = idx % 4;
        rb[v] = *reinterpret_cast<const float4*>(g.Bt + (size_t)(n0 + row) * g.ldb + k0 + 4 * c4);
      }
    }
  };
  auto store_chunk = [&]() {
    {
      const int row = tid / 4, c4 = tid % 4;
      *reinterpret_cast<float4*>(&As[row * LDS_STRIDE + 4 * c4]) = ra;
    }
#pragma unroll
    for (int v = 0; v < (BV + 255) / 256; ++v) {
      const int idx = tid + 256 * v;
      if (idx < BV) {
        const int row = idx / 4, c4 = idx % 4;
        *reinterpret_cast<float4*>(&Bs[row * LDS_STRIDE + 4 * c4]) = rb[v];
      }
    }
  };

  load_chunk(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    __syncthreads();
    store_chunk();
    __syncthreads();
    if (k0 + KC < g.K) load_chunk(k0 + KC);
    // lane (row m = n, k-slot q) uses k = 4q + kk
    const float4 av = *reinterpret_cast<const float4*>(&As[(16 * wv + n) * LDS_STRIDE + 4 * q]);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const float4 bv = *reinterpret_cast<const float4*>(&Bs[(16 * nt + n) * LDS_STRIDE + 4 * q]);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv.x, acc[nt], 0, 0, 0);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv.y, acc[nt], 0, 0, 0);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv.z, acc[nt], 0, 0, 0);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv.w, acc[nt], 0, 0, 0);
    }
  }

  // epilogue: lane holds C[m0 + 16wv + 4q + i][n0 + 16nt + n]
  float esum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int col = n0 + 16 * nt + n;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t row = (size_t)(m0 + 16 * wv + 4 * q + i);
      const float cv = acc[nt][i];
      const size_t o = row * g.ldc + col;
      if (EPI == EPI_TANH) {
        g.C[o] = tanhf(cv + g.bias[col]);
      } else if (EPI == EPI_RES) {
        const float t = tanhf(cv + g.bias[col]);
        g.C2[o] = t;
        g.C[o] = t * g.dt[col] + g.aux1[o];
      } else if (EPI == EPI_E3) {
        const float t = tanhf(cv + g.bias[col]);
        const float y3 = t * g.dt[col] + g.aux1[o];
        esum[i] += y3 * g.k3[col];
        g.C[o] = g.k3[col] * g.dt[col] * (1.f - t * t);
      } else if (EPI == EPI_BWD2) {
        const float dy2 = g.k3[col] + cv;
        const float t2 = g.aux1[o];
        g.C2[o] = dy2;
        g.C[o] = dy2 * g.dt[col] * (1.f - t2 * t2);
      } else if (EPI == EPI_BWD1) {
        const float dy1 = g.aux1[o] + cv;
        const float y1 = g.aux2[o];
        g.C[o] = dy1 * (1.f - y1 * y1);
      } else {
        g.C[o] = cv;
      }
    }
  }
  if (EPI == EPI_E3) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v = esum[i];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      if (n == 0) g.eout[m0 + 16 * wv + 4 * q + i] = v + g.b3;
    }
  }
}

template <int EPI>
static int32_t gemm_dispatch(hipStream_t st, const GemmArgs& g) {
  if (g.M % 64 || g.K % KC || g.N % 16) {
    set_error("gemm shape %d x %d x %d not tiled", g.M, g.N, g.K);
    return DPMD_E_STATE;
  }
  const int nt16 = g.N / 16;
  // EPI_E3 needs the whole row in one block
  auto go = [&](auto ntc) -> int32_t {
    constexpr int NT = decltype(ntc)::value;
    dim3 grid(g.M / 64, nt16 / NT);
    hipLaunchKernelGGL((gemm_f32_kernel<NT, EPI>), grid, dim3(256), 0, st, g);
    DPMD_HIP(hipGetLastError());
    return DPMD_OK;
  };
  if (EPI == EPI_E3) {
    if (nt16 == 15) return go(std::integral_constant<int, 15>());
    if (nt16 == 8) return go(std::integral_constant<int, 8>());
    if (nt16 == 16) return go(std::integral_constant<int, 16>());
    set_error("fitting width %d has no fused energy epilogue", g.N);
    return DPMD_E_UNSUPPORTED;
  }
  if (nt16 % 15 == 0) return go(std::integral_constant<int, 15>());
  if (nt16 % 10 == 0) return go(std::integral_constant<int, 10>());
  if (nt16 % 8 == 0) return go(std::integral_constant<int, 8>());
  if (nt16 % 12 == 0) return go(std::integral_constant<int, 12>());
  if (nt16 % 16 == 0) return go(std::integral_constant<int, 16>());
  set_error("gemm width %d has no tile instantiation", g.N);
  return DPMD_E_UNSUPPORTED;
}

// Fitting-weight block offsets in Model::fit_off[t][*]
enum { FK0T = 0, FK1T, FK2T, FK0, FK1, FK2, FB0, FB1, FB2, FDT1, FDT2, FK3, FB3 };

int32_t fit_launch(System* s, int precision) {
  (void)precision;
  Model* M = s->model;
  const int F = s->fpad;
  const int D = M->axis * M->wout;
  float* Y1 = s->d_fitbuf;
  const size_t act = (size_t)s->mrows * F;
  float* T2 = Y1 + act;
  float* Y2 = T2 + act;
  float* DZ3 = Y2 + act;
  float* DY2 = DZ3 + act;
  float* DZ2 = DY2 + act;
  float* DZ1 = DZ2 + act;
  for (int t = 0; t < s->ntypes; ++t) {
    const int rows = s->rowbase[t + 1] - s->rowbase[t];
    if (rows == 0) continue;
    const size_t r0 = s->rowbase[t];
    const float* W = M->d_fit;
    const size_t* o = M->fit_off[t];
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = rows;
    // 1: y1 = tanh(G K0 + b0)
    g.A = (const float*)s->d_G + r0 * D; g.lda = D; g.K = D;
    g.Bt = W + o[FK0T]; g.ldb = D; g.N = F;
    g.C = Y1 + r0 * F; g.ldc = F; g.bias = W + o[FB0];
    DPMD_TRY(gemm_dispatch<EPI_TANH>(s->stream, g));
    // 2: t2 = tanh(y1 K1 + b1), y2 = t2 dt1 + y1
    g.A = Y1 + r0 * F; g.lda = F; g.K = F;
    g.Bt = W + o[FK1T]; g.ldb = F;
    g.C = Y2 + r0 * F; g.C2 = T2 + r0 * F; g.aux1 = Y1 + r0 * F;
    g.bias = W + o[FB1]; g.dt = W + o[FDT1];
    DPMD_TRY(gemm_dispatch<EPI_RES>(s->stream, g));
    // 3: t3, y3, e, dz3
    g.A = Y2 + r0 * F;
    g.Bt = W + o[FK2T];
    g.C = DZ3 + r0 * F; g.C2 = nullptr; g.aux1 = Y2 + r0 * F;
    g.bias = W + o[FB2]; g.dt = W + o[FDT2]; g.k3 = W + o[FK3];
    g.b3 = M->fit_b3[t];
    g.eout = s->d_eatom + r0;
    DPMD_TRY(gemm_dispatch<EPI_E3>(s->stream, g));
    // 4: dy2 = k3 + dz3 K2^T, dz2 = dy2 dt1 (1 - t2^2)
    g.A = DZ3 + r0 * F;
    g.Bt = W + o[FK2];
    g.C = DZ2 + r0 * F; g.C2 = DY2 + r0 * F; g.aux1 = T2 + r0 * F;
    g.dt = W + o[FDT1];
    DPMD_TRY(gemm_dispatch<EPI_BWD2>(s->stream, g));
    // 5: dy1 = dy2 + dz2 K1^T, dz1 = dy1 (1 - y1^2)
    g.A = DZ2 + r0 * F;
    g.Bt = W + o[FK1];
    g.C = DZ1 + r0 * F; g.C2 = nullptr; g.aux1 = DY2 + r0 * F; g.aux2 = Y1 + r0 * F;
    DPMD_TRY(gemm_dispatch<EPI_BWD1>(s->stream, g));
    // 6: dG = dz1 K0^T
    g.A = DZ1 + r0 * F;
    g.Bt = W + o[FK0]; g.ldb = F; g.K = F;
    g.N = D;
    g.C = s->d_dG + r0 * D; g.ldc = D;
    DPMD_TRY(gemm_dispatch<EPI_STORE>(s->stream, g));
  }
  return DPMD_OK;
}

}  // namespace dpmd

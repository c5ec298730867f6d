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
// FP32: v_mfma_f32_16x16x4f32, K staged in 16-wide slices.  BF16 / FP16: A converted
// while staging into LDS, B pre-converted, v_mfma_f32_16x16x32_{bf16,f16}, FP32
// accumulation and FP32 epilogues.
#include "internal.h"
#include <algorithm>
#include <string.h>

namespace dpmd {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int uintx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

enum FitEpi { EPI_TANH = 0, EPI_RES = 1, EPI_E3 = 2, EPI_BWD2 = 3, EPI_BWD1 = 4, EPI_STORE = 5, EPI_BWD2A = 6 };

struct GemmArgs {
  const float* A;   // M x K fp32, row stride lda
  const uint16_t* Ah;  // or M x K already in the 16-bit operand format (16-bit GEMMs), stride lda
  int lda;
  const void* Bt;   // N x K (fp32 or 16-bit), row stride ldb elements
  const void* Btl;  // split BF16 form: the low parts of Bt (same layout)
  int ldb;
  int M, N, K;
  float* C;         // primary output, row stride ldc
  uint16_t* C16;    // EPI_STORE: FP16 output in place of C (same element layout), or null
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

// tanh(x) = 1 - 2 / (1 + 2^(2x log2 e)), branch free; absolute error ~2e-7
__device__ __forceinline__ float tanh_fit(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return __builtin_fmaf(-2.f, __builtin_amdgcn_rcpf(1.f + e), 1.f);
}

// One output element: C[row][col] from the accumulator value cv.  Returns the energy-head
// contribution y3 k3 for EPI_E3 (0 otherwise).
template <int EPI>
__device__ __forceinline__ float epi_one(const GemmArgs& g, size_t row, int col, float cv) {
  const size_t o = row * g.ldc + col;
  if (EPI == EPI_TANH) {
    g.C[o] = tanh_fit(cv + g.bias[col]);
  } else if (EPI == EPI_RES) {
    const float t = tanh_fit(cv + g.bias[col]);
    g.C2[o] = t;
    g.C[o] = t * g.dt[col] + g.aux1[o];
  } else if (EPI == EPI_E3) {
    const float t = tanh_fit(cv + g.bias[col]);
    const float y3 = t * g.dt[col] + g.aux1[o];
    g.C[o] = g.k3[col] * g.dt[col] * (1.f - t * t);
    return y3 * g.k3[col];
  } else if (EPI == EPI_BWD2) {
    const float dy2 = g.k3[col] + cv;
    const float t2 = g.aux1[o];
    g.C2[o] = dy2;
    g.C[o] = dy2 * g.dt[col] * (1.f - t2 * t2);
  } else if (EPI == EPI_BWD2A) {
    // as EPI_BWD2 with a per-row upstream gradient aux2 in place of the energy head k3
    const float dy2 = g.aux2[o] + cv;
    const float t2 = g.aux1[o];
    g.C2[o] = dy2;
    g.C[o] = dy2 * g.dt[col] * (1.f - t2 * t2);
  } else if (EPI == EPI_BWD1) {
    const float dy1 = g.aux1[o] + cv;
    const float y1 = g.aux2[o];
    g.C[o] = dy1 * (1.f - y1 * y1);
  } else {
    // EPI_STORE writes dE/dG: FP32, or FP16 in the plain 16-bit forms (whose GEMM normally writes
    // it through LDS in 16-byte row pieces, see gemm_h_kernel)
    if (g.C16) g.C16[o] = __builtin_bit_cast(uint16_t, (_Float16)cv);
    else g.C[o] = cv;
  }
  return 0.f;
}

// lane holds C[m0 + 16wv + 4q + i][n0 + 16nt + n]
template <int NT, int EPI>
__device__ __forceinline__ void epilogue(const GemmArgs& g, const floatx4 (&acc)[NT], int m0, int n0, int wv, int q,
                                         int n) {
  float esum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      esum[i] += epi_one<EPI>(g, (size_t)(m0 + 16 * wv + 4 * q + i), n0 + 16 * nt + n, acc[nt][i]);
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

// Column tiles per block of the 240-wide hidden-layer GEMMs whose epilogue is per element:
// three column blocks (5 tiles each) put three times as many blocks on the CUs as one
// 240-wide block (measured: fitting time -2.4 %); the energy-head GEMM keeps whole rows.
constexpr int kNtHidden = 5;
constexpr int KC32 = 16;             // fp32 K granularity (slices of 16 or 32)
constexpr int KC16 = 32;             // 16-bit K granularity (slices of 32 or 64)

// Block tile (row block mb, column block nb) under an XCD-aware order.  The dispatcher deals
// consecutive block ids (x fastest) round-robin over the 8 XCDs, each with its own L2; in
// grid order the column blocks of one row panel reach their XCD a whole column pass apart,
// after the panel has left that L2, so every column block fetched the panel from HBM again.
// Renumbered, each XCD takes a contiguous run of tiles with the column blocks of a row
// panel adjacent, and the panel is fetched once per XCD.  A bijection for any grid size.
__device__ __forceinline__ void xcd_tile(int& mb, int& nb) {
  const int nn = gridDim.y;
  const int total = gridDim.x * nn;
  const int bid = blockIdx.x + gridDim.x * blockIdx.y;
  const int xcd = bid & 7, local = bid >> 3;
  const int per = total >> 3, rem = total & 7;
  const int t = xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  mb = t / nn;
  nb = t - mb * nn;
}

// FP32 GEMM: block tile 64 rows x 16 NT columns, four waves of 16 rows each over all NT
// column tiles; K staged in KC-wide slices, KC = 32 when K allows (half the block barriers of
// KC = 16: fitting time -4 %; 64 for the narrow tiles measured slower), else 16.
template <int NT, int EPI, int KC>
__global__ void __launch_bounds__(256, 2) gemm_f32_kernel(const GemmArgs g) {
  constexpr int LS = KC + 4;                     // LDS row stride (floats)
  __shared__ float As[64 * LS];
  __shared__ float Bs[16 * NT * LS];
  constexpr int BN = 16 * NT;
  constexpr int C4 = KC / 4;                     // float4 per row of a slice
  constexpr int NAP = 64 * C4 / 256;             // A float4 per thread
  constexpr int BV = BN * C4;                    // float4 per B slice
  constexpr int NB = (BV + 255) / 256;
  static_assert(NAP >= 1 && NAP <= 2 && NB <= 8, "staging unrolled for at most 2 A / 8 B pieces");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int q = lane >> 4, n = lane & 15;
  int mb, nb;
  xcd_tile(mb, nb);
  const int m0 = mb * 64;
  const int n0 = nb * BN;
  const float* Bt = (const float*)g.Bt;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  floatx4 ra[NAP], rb[NB];
  auto live = [&](int piece) { return piece + 1 < NB || tid + 256 * piece < BV; };
  auto fetch = [&](int k) {
#pragma unroll
    for (int p = 0; p < NAP; ++p) {
      const int idx = tid + 256 * p;
      ra[p] = *reinterpret_cast<const floatx4*>(g.A + (size_t)(m0 + idx / C4) * g.lda + k + 4 * (idx % C4));
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int idx = tid + 256 * p;
      if (live(p)) rb[p] = *reinterpret_cast<const floatx4*>(Bt + (size_t)(n0 + idx / C4) * g.ldb + k + 4 * (idx % C4));
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NAP; ++p) {
      const int idx = tid + 256 * p;
      *reinterpret_cast<floatx4*>(&As[(idx / C4) * LS + 4 * (idx % C4)]) = ra[p];
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int idx = tid + 256 * p;
      if (live(p)) *reinterpret_cast<floatx4*>(&Bs[(idx / C4) * LS + 4 * (idx % C4)]) = rb[p];
    }
    __syncthreads();
    if (k0 + KC < g.K) fetch(k0 + KC);
#pragma unroll
    for (int sub = 0; sub < KC / 16; ++sub) {
      const floatx4 av = *reinterpret_cast<const floatx4*>(&As[(16 * wv + n) * LS + 16 * sub + 4 * q]);
      // column tiles in pairs: the two accumulators alternate (no MFMA waits on the one before
      // it), and the next pair's B operands are read while this pair's MFMAs run
      auto ldb = [&](int nt) { return *reinterpret_cast<const floatx4*>(&Bs[(16 * nt + n) * LS + 16 * sub + 4 * q]); };
      floatx4 b0 = ldb(0), b1 = ldb(NT > 1 ? 1 : 0);
#pragma unroll
      for (int nt = 0; nt < NT; nt += 2) {
        const floatx4 c0 = b0, c1 = b1;
        if (nt + 2 < NT) b0 = ldb(nt + 2);
        if (nt + 3 < NT) b1 = ldb(nt + 3);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], c0[i], acc[nt], 0, 0, 0);
          if (nt + 1 < NT) acc[nt + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], c1[i], acc[nt + 1], 0, 0, 0);
        }
      }
    }
  }
  epilogue<NT, EPI>(g, acc, m0, n0, wv, q, n);
}

template <int PREC>
struct Half8;
template <>
struct Half8<1> {
  typedef bf16x8 T;
  typedef __bf16 E;
  static __device__ __forceinline__ floatx4 mfma(T a, T b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Half8<2> {
  typedef halfx8 T;
  typedef _Float16 E;
  static __device__ __forceinline__ floatx4 mfma(T a, T b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
// PREC 3: the split BF16 form.  Operands are BF16 high parts and BF16 low parts of the
// residual, x = hi + lo + O(2^-16 |x|); a product is hi*hi + hi*lo + lo*hi on BF16 MFMAs with
// FP32 accumulation (the lo*lo term is below the residual).
template <>
struct Half8<3> : Half8<1> {};
// the low part of the split: bf16(x - hi), hi = bf16(x) (round to nearest even)
__device__ __forceinline__ __bf16 bf16_low(float x, __bf16 hi) { return (__bf16)(x - (float)hi); }

// 16-bit GEMM: block tile 64 rows x 16 NT columns, four waves as 2 x 2, each owning 32
// rows (two 16-row tiles) and half of the column tiles, so every B fragment read from LDS
// feeds two MFMAs (the LDS read rate, not the MFMA rate, bounds a 16-row wave tile).  K is
// staged in 64-wide slices when K allows (two MFMA k-steps per block barrier: BF16 fitting
// time -10 % against 32-wide slices), else 32.
template <int NT, int EPI, int PREC, int KC, bool A16 = false>
__global__ void __launch_bounds__(256, 2) gemm_h_kernel(const GemmArgs g) {
  typedef typename Half8<PREC>::T HT;
  typedef typename Half8<PREC>::E HE;
  constexpr bool SPLIT = PREC == 3;   // low-part tiles beside the high parts
  static_assert(!(SPLIT && A16), "the split form reads FP32 A");
  constexpr int LS = KC + 8;          // LDS row stride (elements)
  __shared__ __attribute__((aligned(16))) uint16_t As[64 * LS];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[16 * NT * LS];
  __shared__ __attribute__((aligned(16))) uint16_t Asl[SPLIT ? 64 * LS : 8];
  __shared__ __attribute__((aligned(16))) uint16_t Bsl[SPLIT ? 16 * NT * LS : 8];
  __shared__ float Es[64];
  constexpr int BN = 16 * NT;
  constexpr int NH = (NT + 1) / 2;   // column tiles per wave
  constexpr int P8 = KC / 8;         // 8-element pieces per row of a slice
  constexpr int NAP = 64 * P8 / 256; // A pieces per thread
  constexpr int BV = BN * P8;        // 16-byte pieces per B slice
  constexpr int NB = (BV + 255) / 256;
  static_assert(NAP >= 1 && NAP <= 2 && NB <= 8, "staging unrolled for at most 2 A / 8 B pieces per thread");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int wm = wv & 1, wn = wv >> 1;
  const int q = lane >> 4, n = lane & 15;
  int mb, nb;
  xcd_tile(mb, nb);
  const int m0 = mb * 64;
  const int n0 = nb * BN;
  const uint16_t* Bt = (const uint16_t*)g.Bt;
  floatx4 acc[2][NH];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int t = 0; t < NH; ++t) acc[mi][t] = floatx4{0.f, 0.f, 0.f, 0.f};
  floatx4 ra[2 * NAP];
  uintx4 rh[NAP];
  uintx4 rb[NB];
  uintx4 rbl[SPLIT ? NB : 1];
  const uint16_t* Btl = (const uint16_t*)g.Btl;
  auto live = [&](int piece) { return piece + 1 < NB || tid + 256 * piece < BV; };
  auto fetch = [&](int k) {
#pragma unroll
    for (int p = 0; p < NAP; ++p) {
      const int idx = tid + 256 * p;
      if constexpr (A16) {
        rh[p] = *reinterpret_cast<const uintx4*>(g.Ah + (size_t)(m0 + idx / P8) * g.lda + k + 8 * (idx % P8));
      } else {
        const float* ap = g.A + (size_t)(m0 + idx / P8) * g.lda + k + 8 * (idx % P8);
        ra[2 * p] = *reinterpret_cast<const floatx4*>(ap);
        ra[2 * p + 1] = *reinterpret_cast<const floatx4*>(ap + 4);
      }
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int idx = tid + 256 * p;
      const size_t off = (size_t)(n0 + idx / P8) * g.ldb + k + 8 * (idx % P8);
      if (live(p)) {
        rb[p] = *reinterpret_cast<const uintx4*>(Bt + off);
        if constexpr (SPLIT) rbl[p] = *reinterpret_cast<const uintx4*>(Btl + off);
      }
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NAP; ++p) {
      const int idx = tid + 256 * p;
      if constexpr (A16) {
        *reinterpret_cast<uintx4*>(&As[(idx / P8) * LS + 8 * (idx % P8)]) = rh[p];
      } else if constexpr (SPLIT) {
        const floatx4 x0 = ra[2 * p], x1 = ra[2 * p + 1];
        HT h, l;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const __bf16 h0 = (__bf16)x0[j], h1 = (__bf16)x1[j];
          h[j] = h0;
          h[4 + j] = h1;
          l[j] = bf16_low(x0[j], h0);
          l[4 + j] = bf16_low(x1[j], h1);
        }
        *reinterpret_cast<HT*>(&As[(idx / P8) * LS + 8 * (idx % P8)]) = h;
        *reinterpret_cast<HT*>(&Asl[(idx / P8) * LS + 8 * (idx % P8)]) = l;
      } else {
        const floatx4 x0 = ra[2 * p], x1 = ra[2 * p + 1];
        HT h;
        h[0] = (HE)x0.x; h[1] = (HE)x0.y; h[2] = (HE)x0.z; h[3] = (HE)x0.w;
        h[4] = (HE)x1.x; h[5] = (HE)x1.y; h[6] = (HE)x1.z; h[7] = (HE)x1.w;
        *reinterpret_cast<HT*>(&As[(idx / P8) * LS + 8 * (idx % P8)]) = h;
      }
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int idx = tid + 256 * p;
      if (live(p)) {
        *reinterpret_cast<uintx4*>(&Bs[(idx / P8) * LS + 8 * (idx % P8)]) = rb[p];
        if constexpr (SPLIT) *reinterpret_cast<uintx4*>(&Bsl[(idx / P8) * LS + 8 * (idx % P8)]) = rbl[p];
      }
    }
    __syncthreads();
    if (k0 + KC < g.K) fetch(k0 + KC);
#pragma unroll
    for (int sub = 0; sub < KC / 32; ++sub) {
      const HT a0 = *reinterpret_cast<const HT*>(&As[(32 * wm + n) * LS + 32 * sub + 8 * q]);
      const HT a1 = *reinterpret_cast<const HT*>(&As[(32 * wm + 16 + n) * LS + 32 * sub + 8 * q]);
      HT a0l, a1l;
      if constexpr (SPLIT) {
        a0l = *reinterpret_cast<const HT*>(&Asl[(32 * wm + n) * LS + 32 * sub + 8 * q]);
        a1l = *reinterpret_cast<const HT*>(&Asl[(32 * wm + 16 + n) * LS + 32 * sub + 8 * q]);
      }
#pragma unroll
      for (int nt = 0; nt < NH; ++nt) {
        const int tile = wn * NH + nt;
        if (NT % 2 == 0 || tile < NT) {
          const HT bv = *reinterpret_cast<const HT*>(&Bs[(16 * tile + n) * LS + 32 * sub + 8 * q]);
          acc[0][nt] = Half8<PREC>::mfma(a0, bv, acc[0][nt]);
          acc[1][nt] = Half8<PREC>::mfma(a1, bv, acc[1][nt]);
          if constexpr (SPLIT) {
            const HT bl = *reinterpret_cast<const HT*>(&Bsl[(16 * tile + n) * LS + 32 * sub + 8 * q]);
            acc[0][nt] = Half8<PREC>::mfma(a0, bl, acc[0][nt]);
            acc[1][nt] = Half8<PREC>::mfma(a1, bl, acc[1][nt]);
            acc[0][nt] = Half8<PREC>::mfma(a0l, bv, acc[0][nt]);
            acc[1][nt] = Half8<PREC>::mfma(a1l, bv, acc[1][nt]);
          }
        }
      }
    }
  }
  // epilogue: lane holds C[m0 + 32 wm + 16 mi + 4q + i][n0 + 16 (wn NH + nt) + n]
  if constexpr (EPI == EPI_STORE && KC >= 56) {
    if (g.C16) {
      // FP16 output (dE/dG of the plain 16-bit forms): the block's 64 x BN tile is written to LDS
      // row-major (the B staging space, 16 NT rows of KC + 8 >= 64 elements), then stored in
      // 16-byte pieces, whole rows of BN elements per block (2-byte stores from the MFMA layout
      // wrote 32-byte row pieces and ran slower than the FP32 output)
      uint16_t* Ct = Bs;
      __syncthreads();
#pragma unroll
      for (int nt = 0; nt < NH; ++nt) {
        const int tile = wn * NH + nt;
        if (NT % 2 == 0 || tile < NT) {
#pragma unroll
          for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int i = 0; i < 4; ++i)
              Ct[(32 * wm + 16 * mi + 4 * q + i) * BN + 16 * tile + n] = __builtin_bit_cast(uint16_t, (_Float16)acc[mi][nt][i]);
        }
      }
      __syncthreads();
      constexpr int PR = BN / 8;  // 16-byte pieces per row
      for (int pc = tid; pc < 64 * PR; pc += 256) {
        const int row = pc / PR, pe = pc - row * PR;
        const uintx4 v = *reinterpret_cast<const uintx4*>(&Ct[row * BN + 8 * pe]);
        __builtin_nontemporal_store(v, reinterpret_cast<uintx4*>(g.C16 + (size_t)(m0 + row) * g.ldc + n0 + 8 * pe));
      }
      return;
    }
  }
  float esum[2][4];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int i = 0; i < 4; ++i) esum[mi][i] = 0.f;
#pragma unroll
  for (int nt = 0; nt < NH; ++nt) {
    const int tile = wn * NH + nt;
    if (NT % 2 == 0 || tile < NT) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          esum[mi][i] += epi_one<EPI>(g, (size_t)(m0 + 32 * wm + 16 * mi + 4 * q + i), n0 + 16 * tile + n,
                                      acc[mi][nt][i]);
    }
  }
  if (EPI == EPI_E3) {
    // row sums: over the 16 columns of a lane group, then over the two column halves
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = esum[mi][i];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        esum[mi][i] = v;
      }
    if (wn == 1 && n == 0) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) Es[32 * wm + 16 * mi + 4 * q + i] = esum[mi][i];
    }
    __syncthreads();
    if (wn == 0 && n == 0) {
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = 32 * wm + 16 * mi + 4 * q + i;
          g.eout[m0 + r] = (esum[mi][i] + Es[r]) + g.b3;
        }
    }
  }
}

template <int NT, int EPI>
static int32_t go(hipStream_t st, const GemmArgs& g, int prec) {
  dim3 grid(g.M / 64, g.N / (16 * NT));
  const bool k64 = g.K % 64 == 0;
  if constexpr (EPI == EPI_TANH) {
    // layer 0 of a 16-bit form reads G as the embedding kernel stored it (16-bit)
    if (g.Ah) {
      if (prec == DPMD_PREC_BF16 && k64) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 1, 64, true>), grid, dim3(256), 0, st, g);
      else if (prec == DPMD_PREC_BF16) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 1, 32, true>), grid, dim3(256), 0, st, g);
      else if (prec == DPMD_PREC_FP16 && k64) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 2, 64, true>), grid, dim3(256), 0, st, g);
      else if (prec == DPMD_PREC_FP16) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 2, 32, true>), grid, dim3(256), 0, st, g);
      else {
        set_error("a 16-bit G needs a 16-bit GEMM");
        return DPMD_E_STATE;
      }
      DPMD_HIP(hipGetLastError());
      return DPMD_OK;
    }
  }
  if (prec == DPMD_PREC_BF16 && k64) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 1, 64>), grid, dim3(256), 0, st, g);
  else if (prec == DPMD_PREC_BF16) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 1, 32>), grid, dim3(256), 0, st, g);
  else if (prec == DPMD_PREC_BF16X3) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 3, 32>), grid, dim3(256), 0, st, g);
  else if (prec == DPMD_PREC_FP16 && k64) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 2, 64>), grid, dim3(256), 0, st, g);
  else if (prec == DPMD_PREC_FP16) hipLaunchKernelGGL((gemm_h_kernel<NT, EPI, 2, 32>), grid, dim3(256), 0, st, g);
  else if (g.K % 32 == 0) hipLaunchKernelGGL((gemm_f32_kernel<NT, EPI, 32>), grid, dim3(256), 0, st, g);
  else hipLaunchKernelGGL((gemm_f32_kernel<NT, EPI, 16>), grid, dim3(256), 0, st, g);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

template <int EPI>
static int32_t gemm_dispatch(hipStream_t st, const GemmArgs& g, int prec) {
  const int kc = prec == DPMD_PREC_FP32 ? KC32 : KC16;
  if (g.M % 64 || g.K % kc || g.N % 16) {
    set_error("gemm shape %d x %d x %d not tiled", g.M, g.N, g.K);
    return DPMD_E_STATE;
  }
  const int nt16 = g.N / 16;
  if (EPI == EPI_E3) {  // the fused energy head needs the whole row in one block
    if (nt16 == 15) return go<15, EPI>(st, g, prec);
    if (nt16 == 8) return go<8, EPI>(st, g, prec);
    if (nt16 == 16) return go<16, EPI>(st, g, prec);
    set_error("fitting width %d has no fused energy epilogue", g.N);
    return DPMD_E_UNSUPPORTED;
  }
  if (EPI != EPI_STORE && EPI != EPI_TANH && nt16 == 15) return go<kNtHidden, EPI>(st, g, prec);
  if (nt16 % 15 == 0) return go<15, EPI>(st, g, prec);
  if (nt16 % 10 == 0) return go<10, EPI>(st, g, prec);
  if (nt16 % 8 == 0) return go<8, EPI>(st, g, prec);
  if (nt16 % 12 == 0) return go<12, EPI>(st, g, prec);
  if (nt16 % 16 == 0) return go<16, EPI>(st, g, prec);
  if (nt16 % 7 == 0) return go<7, EPI>(st, g, prec);
  set_error("gemm width %d has no tile instantiation", g.N);
  return DPMD_E_UNSUPPORTED;
}

// Fitting-weight block offsets in Model::fit_off[t][*] (elements)
enum { FK0T = 0, FK1T, FK2T, FK0, FK1, FK2, FB0, FB1, FB2, FDT1, FDT2, FK3, FB3 };

// ---------------------------------------------------------------------------------
// The hidden layers in one kernel: GEMMs 2-5 of fit_launch (the two hidden forward layers
// with the energy head, and their reverse pass down to dE/dz1) for 64 atoms per block.
// Orientation is transposed: MFMA rows are features, MFMA columns 16 atoms of one wave, so
// a layer's output, register (t, i) of lane group q = feature 16 t + 4 q + i, is the next
// layer's B operand as it stands (the K order of a sum is free; the weight operand is read
// in the same order).  Activations, t2 and dy2 stay in registers; only the weights (A,
// features x features) move, staged through LDS in K slices shared by the block's four
// waves.  HBM traffic per atom: y1 read twice, dz1 written, one energy.
struct HiddenArgs {
  const float* Y1;   // rows x ld: y1 = tanh(G K0 + b0)
  int ld;            // row stride of the activations and of the weight blocks (elements)
  const void* W[4];  // A operands, row-major [m][k], stride ld: K1^T, K2^T (out x in), K2, K1 (in x out)
  const void* Wl[4]; // split BF16 form: the low parts of W (same layout)
  const float* b1;
  const float* b2;
  const float* dt1;
  const float* dt2;
  const float* k3;
  float b3;
  float* eout;       // per-row energy
  float* DZ1;        // rows x ld: dz1 = dE/d(G K0 + b0)
};

// Live state per lane: NT x 4 accumulators, the layer input, and t2 (later dy2): 12 NT
// registers before weight operands.  Up to 128-wide layers two blocks share a CU (256
// registers per lane; the FP16 form's conversions need more); wider ones run one block per
// CU with the full 512.
// Block barrier ordering LDS only: __syncthreads() also orders global memory, which makes
// every wave wait for all of its global loads in flight, the weight prefetch included.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int NT, int PREC>
__global__ void __launch_bounds__(256, (NT <= 8 && PREC <= 1) ? 2 : 1) fit_hidden_kernel(const HiddenArgs h) {
  constexpr bool H16 = PREC != 0;
  constexpr bool SPLIT = PREC == 3;  // low-part weight slices beside the high parts
  constexpr int NS = H16 ? (NT + 1) / 2 : NT;   // K slices per layer: 32 features (16-bit) or 16
  constexpr int LSW = H16 ? 40 : 20;            // LDS row stride of a staged slice (elements)
  constexpr int ESZ = H16 ? 2 : 4;
  constexpr int SLICE = 16 * NT * LSW * ESZ;    // bytes
  constexpr int CH = 16 * NT * 4;               // 16-byte pieces per slice (4 per weight row)
  constexpr int NCH = (CH + 255) / 256;
  __shared__ __attribute__((aligned(16))) unsigned char Ws[2][SLICE];
  __shared__ __attribute__((aligned(16))) unsigned char Wsl[SPLIT ? 2 : 1][SPLIT ? SLICE : 16];
  __shared__ __attribute__((aligned(16))) unsigned char spill_slot[32];  // stores of pieces past CH
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int q = lane >> 4, n = lane & 15;
  const int ld = h.ld;
  const size_t row = (size_t)blockIdx.x * 64 + 16 * wv + n;  // this lane's atom

  // slice s of a weight block into registers / from registers into LDS buffer b.  FP32: 16
  // columns, LDS row = the 16 values in order.  16-bit: 32 columns, LDS row permuted so lane
  // group q finds the 8 operands of its k-slots (columns 4q..4q+3 and 16+4q..16+4q+3) together.
  // Branch free: only the last piece can lie past CH (256 (NCH - 1) < CH); there the load
  // re-reads piece CH - 1 and the store goes to spill_slot.
  auto piece = [&](int p) {
    const int c = tid + 256 * p;
    return (p + 1 < NCH || CH % 256 == 0) ? c : min(c, CH - 1);
  };
  auto stage_load = [&](const void* Wg, int s, uintx4 (&st)[NCH]) {
#pragma unroll
    for (int p = 0; p < NCH; ++p) {
      const int c = piece(p);
      const int r = c >> 2, u = c & 3;
      const unsigned char* src = (const unsigned char*)Wg + ((size_t)r * ld + (size_t)s * (H16 ? 32 : 16)) * ESZ + 16 * u;
      st[p] = *reinterpret_cast<const uintx4*>(src);
    }
  };
  auto stage_store = [&](int b, const uintx4 (&st)[NCH], bool low = false) {
#pragma unroll
    for (int p = 0; p < NCH; ++p) {
      const int c = piece(p);
      const int r = c >> 2, u = c & 3;
      const bool past = (p + 1 == NCH && CH % 256 != 0) && tid + 256 * p >= CH;
      if constexpr (H16) {
        unsigned char* base = Ws[b];
        if constexpr (SPLIT) base = low ? Wsl[b] : Ws[b];
        uint16_t* d = reinterpret_cast<uint16_t*>(base) + r * LSW + 16 * (u & 1) + 4 * (u >> 1);
        if (p + 1 == NCH && CH % 256 != 0) d = past ? reinterpret_cast<uint16_t*>(spill_slot) : d;
        *reinterpret_cast<uint2*>(d) = make_uint2(st[p][0], st[p][1]);
        *reinterpret_cast<uint2*>(d + 8) = make_uint2(st[p][2], st[p][3]);
      } else {
        float* d = reinterpret_cast<float*>(Ws[b]) + r * LSW + 4 * u;
        if (p + 1 == NCH && CH % 256 != 0) d = past ? reinterpret_cast<float*>(spill_slot) : d;
        *reinterpret_cast<uintx4*>(d) = st[p];
      }
    }
  };
  // The four layers' weight slices form one sequence g = l NS + s, loaded PF slices ahead
  // into registers (a 16-bit slice is only NT MFMAs of work, far shorter than an L2 round
  // trip) and stored into the LDS buffer g & 1 at the end of slice g - 1; the prefetch runs
  // across layer boundaries, so the epilogues overlap the next layer's first loads.
  constexpr int PF = H16 ? 4 : 1;
  constexpr int NG = 4 * NS;
  uintx4 st[PF][NCH];
  uintx4 stl[SPLIT ? PF : 1][NCH];
  // acc[ot] = sum_k W_l[16 ot + .][k] act[k][.] over layer l's K slices
  auto layer = [&](auto lc, floatx4 (&acc)[NT], const float (&bop)[NT][4]) {
    constexpr int l = decltype(lc)::value;
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int g = l * NS + s;
      if (g + PF < NG) {
        stage_load(h.W[(g + PF) / NS], (g + PF) % NS, st[g % PF]);
        if constexpr (SPLIT) stage_load(h.Wl[(g + PF) / NS], (g + PF) % NS, stl[g % PF]);
      }
      // keep every slice's loads, MFMAs and stores in their slice (the scheduler would
      // otherwise hoist the prefetch of many slices at once and wait for all of them)
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (H16) {
        typedef typename Half8<PREC>::T HT;
        typedef typename Half8<PREC>::E HE;
        HT b, bl;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int t = 2 * s + j / 4;
          const float x = t < NT ? bop[t < NT ? t : 0][j % 4] : 0.f;
          b[j] = (HE)x;
          if constexpr (SPLIT) bl[j] = bf16_low(x, b[j]);
        }
        const uint16_t* Wl = reinterpret_cast<const uint16_t*>(Ws[g & 1]);
#pragma unroll
        for (int ot = 0; ot < NT; ++ot) {
          const HT av = *reinterpret_cast<const HT*>(Wl + (16 * ot + n) * LSW + 8 * q);
          acc[ot] = Half8<PREC>::mfma(av, b, acc[ot]);
          if constexpr (SPLIT) {
            const HT avl = *reinterpret_cast<const HT*>(reinterpret_cast<const uint16_t*>(Wsl[g & 1]) +
                                                        (16 * ot + n) * LSW + 8 * q);
            acc[ot] = Half8<PREC>::mfma(av, bl, acc[ot]);
            acc[ot] = Half8<PREC>::mfma(avl, b, acc[ot]);
          }
        }
      } else {
        const float* Wl = reinterpret_cast<const float*>(Ws[g & 1]);
        // output tiles in pairs: the two accumulators alternate (no MFMA waits on the one before it)
        auto lda = [&](int ot) { return *reinterpret_cast<const floatx4*>(Wl + (16 * ot + n) * LSW + 4 * q); };
        floatx4 a0 = lda(0), a1 = lda(NT > 1 ? 1 : 0);
#pragma unroll
        for (int ot = 0; ot < NT; ot += 2) {
          const floatx4 c0 = a0, c1 = a1;
          if (ot + 2 < NT) a0 = lda(ot + 2);
          if (ot + 3 < NT) a1 = lda(ot + 3);
          // the next pair's reads stay ahead of these MFMAs (the wide layers run one wave per
          // SIMD, which nothing else hides an LDS wait behind)
          if constexpr (NT > 8) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            acc[ot] = __builtin_amdgcn_mfma_f32_16x16x4f32(c0[i], bop[s][i], acc[ot], 0, 0, 0);
            if (ot + 1 < NT) acc[ot + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(c1[i], bop[s][i], acc[ot + 1], 0, 0, 0);
          }
        }
      }
      if (g + 1 < NG) {
        stage_store((g + 1) & 1, st[(g + 1) % PF]);
        if constexpr (SPLIT) stage_store((g + 1) & 1, stl[(g + 1) % PF], true);
        lds_barrier();
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  auto vec4 = [&](const float* p, int t) { return *reinterpret_cast<const floatx4*>(p + 16 * t + 4 * q); };

  float bop[NT][4], keep[NT][4];
  floatx4 acc[NT];
  const float* y1row = h.Y1 + row * ld;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const floatx4 v = vec4(y1row, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) bop[t][i] = v[i];
  }
#pragma unroll
  for (int p = 0; p < PF; ++p) {
    stage_load(h.W[p / NS], p % NS, st[p]);
    if constexpr (SPLIT) stage_load(h.Wl[p / NS], p % NS, stl[p]);
  }
  stage_store(0, st[0]);
  if constexpr (SPLIT) stage_store(0, stl[0], true);
  lds_barrier();
  // t2 = tanh(y1 K1 + b1), y2 = t2 dt1 + y1
  layer(std::integral_constant<int, 0>(), acc, bop);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const floatx4 bb = vec4(h.b1, t), dd = vec4(h.dt1, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float t2 = tanh_fit(acc[t][i] + bb[i]);
      keep[t][i] = t2;
      bop[t][i] = t2 * dd[i] + bop[t][i];
    }
  }
  // t3 = tanh(y2 K2 + b2), y3 = t3 dt2 + y2, e = y3 k3 + b3, dz3 = k3 dt2 (1 - t3^2)
  layer(std::integral_constant<int, 1>(), acc, bop);
  float esum = 0.f;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const floatx4 bb = vec4(h.b2, t), dd = vec4(h.dt2, t), kk = vec4(h.k3, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float t3 = tanh_fit(acc[t][i] + bb[i]);
      const float y3 = t3 * dd[i] + bop[t][i];
      esum += y3 * kk[i];
      bop[t][i] = kk[i] * dd[i] * (1.f - t3 * t3);
    }
  }
  esum += __shfl_xor(esum, 16);
  esum += __shfl_xor(esum, 32);
  if (q == 0) h.eout[row] = esum + h.b3;
  // dy2 = k3 + dz3 K2^T, dz2 = dy2 dt1 (1 - t2^2)
  layer(std::integral_constant<int, 2>(), acc, bop);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const floatx4 kk = vec4(h.k3, t), dd = vec4(h.dt1, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float dy2 = kk[i] + acc[t][i];
      bop[t][i] = dy2 * dd[i] * (1.f - keep[t][i] * keep[t][i]);
      keep[t][i] = dy2;
    }
  }
  // dy1 = dy2 + dz2 K1^T, dz1 = dy1 (1 - y1^2)
  layer(std::integral_constant<int, 3>(), acc, bop);
  float* dzrow = h.DZ1 + row * ld;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const floatx4 y1 = vec4(y1row, t);
    floatx4 o;
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = (keep[t][i] + acc[t][i]) * (1.f - y1[i] * y1[i]);
    *reinterpret_cast<floatx4*>(dzrow + 16 * t + 4 * q) = o;
  }
}

template <int NT>
static int32_t hidden_go(hipStream_t st, const HiddenArgs& h, int rows, int prec) {
  const dim3 grid(rows / 64);
  if (prec == DPMD_PREC_BF16) hipLaunchKernelGGL((fit_hidden_kernel<NT, 1>), grid, dim3(256), 0, st, h);
  else if (prec == DPMD_PREC_FP16) hipLaunchKernelGGL((fit_hidden_kernel<NT, 2>), grid, dim3(256), 0, st, h);
  else if (prec == DPMD_PREC_BF16X3) hipLaunchKernelGGL((fit_hidden_kernel<NT, 3>), grid, dim3(256), 0, st, h);
  else hipLaunchKernelGGL((fit_hidden_kernel<NT, 0>), grid, dim3(256), 0, st, h);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// Launches the fused hidden layers when the width has an instantiation; returns false (and
// launches nothing) otherwise, for the per-GEMM path.
static bool hidden_fused(hipStream_t st, const HiddenArgs& h, int FN, int rows, int prec, int32_t* rc) {
  if (rows % 64) return false;
  switch (FN / 16) {
    case 8: *rc = hidden_go<8>(st, h, rows, prec); return true;
    case 15: *rc = hidden_go<15>(st, h, rows, prec); return true;
    case 16: *rc = hidden_go<16>(st, h, rows, prec); return true;
    default: return false;
  }
}

int32_t fit_launch(System* s, int precision) {
  Model* M = s->model;
  const int FN = s->fpad;          // output width (multiple of 16)
  const int FK = s->fstride;       // activation row stride = K of hidden GEMMs (multiple of 32)
  const int D = M->axis * M->wout;
  float* Y1 = s->d_fitbuf;
  const size_t act = (size_t)s->mrows * FK;
  float* T2 = Y1 + act;
  float* Y2 = T2 + act;
  float* DZ3 = Y2 + act;
  float* DY2 = DZ3 + act;
  float* DZ2 = DY2 + act;
  float* DZ1 = DZ2 + act;
  const float* Wf = M->d_fit;
  // 16-bit weight copies: BF16 (also the high parts of the split form), FP16; the split form
  // adds the BF16 low parts
  const bool split = precision == DPMD_PREC_BF16X3;
  const bool dg16 = precision == DPMD_PREC_BF16 || precision == DPMD_PREC_FP16;
  s->dg16 = dg16 ? 1 : 0;
  const uint16_t* Wh = precision == DPMD_PREC_FP32 ? nullptr : M->d_fit_h[split ? 0 : precision - 1];
  const uint16_t* Wlo = split ? M->d_fit_h[2] : nullptr;
  if ((precision != DPMD_PREC_FP32 && !Wh) || (split && !Wlo)) {
    set_error("16-bit fitting weights missing");
    return DPMD_E_STATE;
  }
  for (int t = 0; t < s->ntypes; ++t) {
    const int rows = s->rowbase[t + 1] - s->rowbase[t];
    if (rows == 0) continue;
    const size_t r0 = s->rowbase[t];
    const size_t* o = M->fit_off[t];
    auto B = [&](int blk) -> const void* {
      return Wh ? (const void*)(Wh + o[blk]) : (const void*)(Wf + o[blk]);
    };
    auto BL = [&](int blk) -> const void* { return Wlo ? (const void*)(Wlo + o[blk]) : nullptr; };
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = rows;
    // 1: y1 = tanh(G K0 + b0); a plain 16-bit form's G was stored 16-bit by the embedding
    // kernel, the FP32 and split forms read it FP32
    if (precision == DPMD_PREC_FP32 || split) g.A = (const float*)s->d_G + r0 * D;
    else g.Ah = (const uint16_t*)s->d_G + r0 * D;
    g.lda = D; g.K = D;
    g.Bt = B(FK0T); g.Btl = BL(FK0T); g.ldb = D; g.N = FN;
    g.C = Y1 + r0 * FK; g.ldc = FK; g.bias = Wf + o[FB0];
    DPMD_TRY(gemm_dispatch<EPI_TANH>(s->stream, g, precision));
    // 2-5 in one kernel where the width has an instantiation
    HiddenArgs hd;
    memset(&hd, 0, sizeof(hd));
    hd.Y1 = Y1 + r0 * FK;
    hd.ld = FK;
    hd.W[0] = B(FK1T);
    hd.W[1] = B(FK2T);
    hd.W[2] = B(FK2);
    hd.W[3] = B(FK1);
    hd.Wl[0] = BL(FK1T);
    hd.Wl[1] = BL(FK2T);
    hd.Wl[2] = BL(FK2);
    hd.Wl[3] = BL(FK1);
    hd.b1 = Wf + o[FB1];
    hd.b2 = Wf + o[FB2];
    hd.dt1 = Wf + o[FDT1];
    hd.dt2 = Wf + o[FDT2];
    hd.k3 = Wf + o[FK3];
    hd.b3 = M->fit_b3[t];
    hd.eout = s->d_eatom + r0;
    hd.DZ1 = DZ1 + r0 * FK;
    int32_t hrc = DPMD_OK;
    if (s->fit_fused && hidden_fused(s->stream, hd, FN, rows, precision, &hrc)) {
      DPMD_TRY(hrc);
    } else {
      // 2: t2 = tanh(y1 K1 + b1), y2 = t2 dt1 + y1
      g.A = Y1 + r0 * FK; g.Ah = nullptr; g.lda = FK; g.K = FK;
      g.Bt = B(FK1T); g.Btl = BL(FK1T); g.ldb = FK;
      g.C = Y2 + r0 * FK; g.C2 = T2 + r0 * FK; g.aux1 = Y1 + r0 * FK;
      g.bias = Wf + o[FB1]; g.dt = Wf + o[FDT1];
      DPMD_TRY(gemm_dispatch<EPI_RES>(s->stream, g, precision));
      // 3: t3, y3, e, dz3
      g.A = Y2 + r0 * FK;
      g.Bt = B(FK2T); g.Btl = BL(FK2T);
      g.C = DZ3 + r0 * FK; g.C2 = nullptr; g.aux1 = Y2 + r0 * FK;
      g.bias = Wf + o[FB2]; g.dt = Wf + o[FDT2]; g.k3 = Wf + o[FK3];
      g.b3 = M->fit_b3[t];
      g.eout = s->d_eatom + r0;
      DPMD_TRY(gemm_dispatch<EPI_E3>(s->stream, g, precision));
      // 4: dy2 = k3 + dz3 K2^T, dz2 = dy2 dt1 (1 - t2^2)
      g.A = DZ3 + r0 * FK;
      g.Bt = B(FK2); g.Btl = BL(FK2);
      g.C = DZ2 + r0 * FK; g.C2 = DY2 + r0 * FK; g.aux1 = T2 + r0 * FK;
      g.dt = Wf + o[FDT1];
      DPMD_TRY(gemm_dispatch<EPI_BWD2>(s->stream, g, precision));
      // 5: dy1 = dy2 + dz2 K1^T, dz1 = dy1 (1 - y1^2)
      g.A = DZ2 + r0 * FK;
      g.Bt = B(FK1); g.Btl = BL(FK1);
      g.C = DZ1 + r0 * FK; g.C2 = nullptr; g.aux1 = DY2 + r0 * FK; g.aux2 = Y1 + r0 * FK;
      DPMD_TRY(gemm_dispatch<EPI_BWD1>(s->stream, g, precision));
    }
    // 6: dG = dz1 K0^T; the plain 16-bit forms store it FP16 (the embedding backward reads it
    // once per centre; its GEMM operands were 16-bit already)
    g.A = DZ1 + r0 * FK; g.Ah = nullptr; g.lda = FK;
    g.Bt = B(FK0); g.Btl = BL(FK0); g.ldb = FK; g.K = FK;
    g.N = D;
    g.C = s->d_dG + r0 * D; g.ldc = D;
    g.C16 = dg16 ? reinterpret_cast<uint16_t*>(s->d_dG) + r0 * D : nullptr;
    DPMD_TRY(gemm_dispatch<EPI_STORE>(s->stream, g, precision));
  }
  return DPMD_OK;
}

// ---------------------------------------------------------------------------------
// Atomic vector (Wannier-centroid) model: fitting_net(use_final=False) (deepmd_jax/utils.py:
// 153-166) gives the last hidden layer f, the output is out_norm sum_w f_w T3[:, w]
// (deepmd_jax/dpmodel.py:83-98).  The ten activation buffers share the fitting scratch.

int32_t fit_atomic_buffers(System* s, const Model* M, AtomicFitBufs& b) {
  const int FK = M->fit_k;
  if (!s->d_wc_fit) { set_error("centroid model buffers missing"); return DPMD_E_STATE; }
  const size_t act = (size_t)s->mrows * FK;
  float* p = s->d_wc_fit;
  float** dst[10] = {&b.Y1, &b.T2, &b.Y2, &b.T3, &b.Y3, &b.DY3, &b.DZ3, &b.DY2, &b.DZ2, &b.DZ1};
  for (int k = 0; k < 10; ++k) *dst[k] = p + k * act;
  b.FK = FK;
  return DPMD_OK;
}

int32_t fit_atomic_fwd(System* s, const AtomicFitBufs& b) {
  Model* M = s->model;
  const int FN = M->fit_n, FK = b.FK;
  const int D = M->axis * M->wout;
  const float* Wf = M->d_fit;
  for (int t = 0; t < s->ntypes; ++t) {
    const int rows = s->rowbase[t + 1] - s->rowbase[t];
    if (rows == 0 || !(M->nsel_mask >> t & 1)) continue;
    const size_t r0 = s->rowbase[t];
    const size_t* o = M->fit_off[t];
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = rows;
    g.A = (const float*)s->d_G + r0 * D; g.lda = D; g.K = D;
    g.Bt = Wf + o[FK0T]; g.ldb = D; g.N = FN;
    g.C = b.Y1 + r0 * FK; g.ldc = FK; g.bias = Wf + o[FB0];
    DPMD_TRY(gemm_dispatch<EPI_TANH>(s->stream, g, DPMD_PREC_FP32));
    g.A = b.Y1 + r0 * FK; g.lda = FK; g.K = FK;
    g.Bt = Wf + o[FK1T]; g.ldb = FK;
    g.C = b.Y2 + r0 * FK; g.C2 = b.T2 + r0 * FK; g.aux1 = b.Y1 + r0 * FK;
    g.bias = Wf + o[FB1]; g.dt = Wf + o[FDT1];
    DPMD_TRY(gemm_dispatch<EPI_RES>(s->stream, g, DPMD_PREC_FP32));
    g.A = b.Y2 + r0 * FK;
    g.Bt = Wf + o[FK2T];
    g.C = b.Y3 + r0 * FK; g.C2 = b.T3 + r0 * FK; g.aux1 = b.Y2 + r0 * FK;
    g.bias = Wf + o[FB2]; g.dt = Wf + o[FDT2];
    DPMD_TRY(gemm_dispatch<EPI_RES>(s->stream, g, DPMD_PREC_FP32));
  }
  return DPMD_OK;
}

// DZ3 (= DY3 dt2 (1 - t3^2), formed by the caller) -> dG of the selected types; the rows of
// other types get dG = 0
int32_t fit_atomic_bwd(System* s, const AtomicFitBufs& b) {
  Model* M = s->model;
  s->dg16 = 0;
  const int FK = b.FK;
  const int D = M->axis * M->wout;
  const float* Wf = M->d_fit;
  for (int t = 0; t < s->ntypes; ++t) {
    const int rows = s->rowbase[t + 1] - s->rowbase[t];
    if (rows == 0) continue;
    const size_t r0 = s->rowbase[t];
    if (!(M->nsel_mask >> t & 1)) {
      DPMD_HIP(hipMemsetAsync(s->d_dG + r0 * D, 0, sizeof(float) * rows * D, s->stream));
      continue;
    }
    const size_t* o = M->fit_off[t];
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = rows;
    g.lda = FK; g.K = FK; g.ldb = FK; g.N = M->fit_n; g.ldc = FK;
    // dy2 = dy3 + dz3 K2^T, dz2 = dy2 dt1 (1 - t2^2)
    g.A = b.DZ3 + r0 * FK;
    g.Bt = Wf + o[FK2];
    g.C = b.DZ2 + r0 * FK; g.C2 = b.DY2 + r0 * FK; g.aux1 = b.T2 + r0 * FK; g.aux2 = b.DY3 + r0 * FK;
    g.dt = Wf + o[FDT1];
    DPMD_TRY(gemm_dispatch<EPI_BWD2A>(s->stream, g, DPMD_PREC_FP32));
    // dy1 = dy2 + dz2 K1^T, dz1 = dy1 (1 - y1^2)
    g.A = b.DZ2 + r0 * FK;
    g.Bt = Wf + o[FK1];
    g.C = b.DZ1 + r0 * FK; g.C2 = nullptr; g.aux1 = b.DY2 + r0 * FK; g.aux2 = b.Y1 + r0 * FK;
    DPMD_TRY(gemm_dispatch<EPI_BWD1>(s->stream, g, DPMD_PREC_FP32));
    // dG = dz1 K0^T
    g.A = b.DZ1 + r0 * FK;
    g.Bt = Wf + o[FK0]; g.ldb = FK; g.K = FK;
    g.N = D;
    g.C = s->d_dG + r0 * D; g.ldc = D;
    DPMD_TRY(gemm_dispatch<EPI_STORE>(s->stream, g, DPMD_PREC_FP32));
  }
  return DPMD_OK;
}

}  // namespace dpmd

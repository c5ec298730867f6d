// Fused environment matrix + type-pair embedding MLP + reducer (forward) and its
// hand-written backward, on gfx950 MFMA (v_mfma_f32_16x16x4f32).
//
// Math restated from deepmd-jax:
//   pair geometry  deepmd_jax/utils.py:20-23,79-82   x = minimg(r_j - r_i), r = |x|
//   switch         deepmd_jax/utils.py:29-32          s = (1-3t^2+2t^3)/(r+1e-15), t=r/rc
//   env rows       deepmd_jax/dpmodel.py:37-47, utils.py:51-52
//                  R = [sh, sqrt3 sh u, 3 sh (u_c^2-1/3), 3 sh sqrt2 u_c u_{c-1}], u = x/(r+1e-16)
//   embedding      deepmd_jax/utils.py:107-125        h1=tanh(st K0+b0); h_{l+1}=tanh(h_l K_l+b_l)+repeat(h_l)
//   reducer        deepmd_jax/dpmodel.py:50-51        T = sum_j R^T h3 / Nnbrs
//   descriptor G   deepmd_jax/dpmodel.py:74-78
// Neighbours beyond rcut have s = 0 and R = 0 exactly, so each centre's skin list is
// compacted to the pairs with r < rcut before the MLP.
//
// Orientation: every layer is computed transposed, features x neighbours, on 16 x 16
// tiles.  A 16-neighbour chunk of one centre occupies the 16 columns; lane (q, n)
// holds features {16*tile + 4q + i} of neighbour n.  That accumulator layout is the
// B-operand layout of the next layer (K = features), so the MLP chain runs without any
// data movement; the repeat-residual parent of register r is register r/Z in the same
// lane because of the feature layout (EmbedLayout).  Only the reducer and the
// backward's dT tables go through LDS.
#include "internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <type_traits>

namespace dpmd {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// 16 blocks of 4x4x1: lane l is row (A) and column (B) l % 4 of block l / 4; register j of the
// result is row j of the lane's block and column (a quarter of mfma4's cost)
__device__ __forceinline__ floatx4 mfma4x4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane swaps of gfx950 (VALU, no LDS): permlane32_swap(x, y) exchanges the upper 32 lanes of x
// with the lower 32 lanes of y, permlane16_swap(x, y) the odd 16-lane rows of x with the even
// rows of y; each returns the new (x, y).
__device__ __forceinline__ float swap32_lo(float x, float y) {
  return __uint_as_float(__builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false)[0]);
}
__device__ __forceinline__ float swap16_lo(float x, float y) {
  return __uint_as_float(__builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false)[0]);
}

// Sum over the four 16-lane rows (lane groups): every lane receives its column's sum
// ((x0 + x2) + (x1 + x3), the same order in every row).
__device__ __forceinline__ float row_sum4(float x) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  const float y = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(y), __float_as_uint(y), false, false);
  return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Lane (q, n) receives register q of lane (0, n): four values that only lane group 0 holds,
// spread over the groups as one K step of a 16x16x4 product.  fold2: groups 0, 1 receive
// registers 0, 1 of group 0 (groups 2, 3 those of group 2).
__device__ __forceinline__ float fold4(float r0, float r1, float r2, float r3) {
  return swap16_lo(swap32_lo(r0, r2), swap32_lo(r1, r3));
}
__device__ __forceinline__ float fold2(float r0, float r1) { return swap16_lo(r0, r1); }

// ---------------------------------------------------------------------------------
// Host-side layout and packing

// 2 log2(e): forward weight tables and biases are stored multiplied by it (see tanh_s)
constexpr float kTanhScale = 2.8853900817779268f;

bool EmbedLayout::init(const int* widths) {
  for (int l = 0; l < 3; ++l) w[l] = widths[l];
  if (w[0] <= 0 || w[0] > 32) return false;
  z[0] = 1;
  for (int l = 1; l < 3; ++l) {
    if (w[l] == w[l - 1]) z[l] = 1;
    else if (w[l] == 2 * w[l - 1]) z[l] = 2;
    else return false;
  }
  int acc = 0;
  for (int q = 0; q < 4; ++q) {
    size[0][q] = w[0] / 4 + (q < w[0] % 4 ? 1 : 0);
    base[0][q] = acc;
    acc += size[0][q];
  }
  for (int l = 1; l < 3; ++l)
    for (int q = 0; q < 4; ++q) {
      size[l][q] = z[l] * size[l - 1][q];
      base[l][q] = z[l] * base[l - 1][q];
    }
  for (int l = 0; l < 3; ++l) {
    int mx = 0;
    for (int q = 0; q < 4; ++q) mx = std::max(mx, size[l][q]);
    tiles[l] = (mx + 3) / 4;
  }
  // fold form: sizes (s + 1, s, s, s) in layer 1, and in layers 2 and 3 (S + d, S, S, S) with
  // S a multiple of 4 and d <= 4, so the last tile of those layers is group 0's d extra rows
  const int s1 = size[0][1];
  fold = size[0][0] == s1 + 1 && size[0][2] == s1 && size[0][3] == s1;
  for (int l = 1; l < 3 && fold; ++l) {
    const int S = size[l][1], d = size[l][0] - S;
    fold = S % 4 == 0 && d >= 1 && d <= 4 && size[l][2] == S && size[l][3] == S && tiles[l] == S / 4 + 1;
  }
  return true;
}

int EmbedLayout::logical(int layer, int slot) const {
  int tile = slot / 16, rem = slot % 16, q = rem / 4, i = rem % 4;
  int r = 4 * tile + i;
  if (tile >= tiles[layer] || r >= size[layer][q]) return -1;
  return base[layer][q] + r;
}

EmbedNetOffsets embed_net_offsets(const EmbedLayout& L) {
  EmbedNetOffsets o;
  int t1 = L.tiles[0], t2 = L.tiles[1], t3 = L.tiles[2];
  int p = 0;
  o.k0 = p; p += 16 * t1;
  o.b0 = p; p += 16 * t1;
  o.b1 = p; p += 16 * t2;
  o.b2 = p; p += 16 * t3;
  p = (p + 63) / 64 * 64;
  o.w1a = p; p += t2 * 4 * t1 * 64;
  o.w2a = p; p += t3 * 4 * t2 * 64;
  o.w2b = p; p += t2 * 4 * t3 * 64;
  o.w1b = p; p += t1 * 4 * t2 * 64;
  o.f2a = p; p += t3 * (4 * (t2 - 1) + 1) * 64;
  o.f2b = p; p += t2 * (4 * (t3 - 1) + 1) * 64;
  o.f1b = p; p += t1 * (4 * (t2 - 1) + 1) * 64;
  o.f1a = p; p += t2 * 4 * t1 * 64;
  o.s3a = p; p += 4 * t2 * 64;
  o.total = p;
  return o;
}

// K0 (1,w0), K1 (w0,w1), K2 (w1,w2) in flax (in,out) row-major layout.
void embed_pack_net(const EmbedLayout& L, const float* K0, const float* B0, const float* K1,
                    const float* B1, const float* K2, const float* B2, float* out) {
  EmbedNetOffsets o = embed_net_offsets(L);
  memset(out, 0, sizeof(float) * o.total);
  int t1 = L.tiles[0], t2 = L.tiles[1], t3 = L.tiles[2];
  // forward operands (first-layer weights, biases, w1a, w2a) carry the tanh scale; the
  // backward tables (w2b, w1b) multiply gradients and stay unscaled
  for (int s = 0; s < 16 * t1; ++s) {
    int a = L.logical(0, s);
    if (a >= 0) { out[o.k0 + s] = kTanhScale * K0[a]; out[o.b0 + s] = kTanhScale * B0[a]; }
  }
  for (int s = 0; s < 16 * t2; ++s) {
    int c = L.logical(1, s);
    if (c >= 0) out[o.b1 + s] = kTanhScale * B1[c];
  }
  for (int s = 0; s < 16 * t3; ++s) {
    int w = L.logical(2, s);
    if (w >= 0) out[o.b2 + s] = kTanhScale * B2[w];
  }
  const int w1 = L.w[1], w2 = L.w[2];
  for (int lane = 0; lane < 64; ++lane) {
    int m = lane % 16, q = lane / 16;
    // layer-2 forward: A[m=out slot][k=in slot]
    for (int ot = 0; ot < t2; ++ot)
      for (int st = 0; st < 4 * t1; ++st) {
        int a = L.logical(0, 16 * (st / 4) + 4 * q + st % 4);
        int c = L.logical(1, 16 * ot + m);
        out[o.w1a + (ot * 4 * t1 + st) * 64 + lane] = (a >= 0 && c >= 0) ? kTanhScale * K1[a * w1 + c] : 0.f;
      }
    for (int ot = 0; ot < t3; ++ot)
      for (int st = 0; st < 4 * t2; ++st) {
        int c = L.logical(1, 16 * (st / 4) + 4 * q + st % 4);
        int w = L.logical(2, 16 * ot + m);
        out[o.w2a + (ot * 4 * t2 + st) * 64 + lane] = (c >= 0 && w >= 0) ? kTanhScale * K2[c * w2 + w] : 0.f;
      }
    // backward: A[m=in slot][k=out slot]
    for (int ct = 0; ct < t2; ++ct)
      for (int st = 0; st < 4 * t3; ++st) {
        int c = L.logical(1, 16 * ct + m);
        int w = L.logical(2, 16 * (st / 4) + 4 * q + st % 4);
        out[o.w2b + (ct * 4 * t3 + st) * 64 + lane] = (c >= 0 && w >= 0) ? K2[c * w2 + w] : 0.f;
      }
    for (int at = 0; at < t1; ++at)
      for (int st = 0; st < 4 * t2; ++st) {
        int a = L.logical(0, 16 * at + m);
        int c = L.logical(1, 16 * (st / 4) + 4 * q + st % 4);
        out[o.w1b + (at * 4 * t2 + st) * 64 + lane] = (a >= 0 && c >= 0) ? K1[a * w1 + c] : 0.f;
      }
  }
  if (!L.fold) return;
  // Fold form.  The last tile of layers 2 and 3 holds lane group 0's extra rows only (e2, e3: its
  // first slot).  A folded K step gives lane group q group 0's register q of such a tile, so its
  // A operand pairs row m with feature (group 0, register 4 (tiles - 1) + q).  A thin tile (the
  // last layer-2 tile as an output) runs on v_mfma_f32_4x4x1_16b_f32, which treats lane l as row
  // (A) and column (B) l % 4 of block l / 4: step r multiplies each lane's register r (feature
  // (q, r) of neighbour n, or the folded feature) by A = weight[thin row n % 4][that feature], and
  // block (q, n / 4) accumulates lane group q's part of the thin rows; a sum over the lane groups
  // ends it.  Tables (tiles of [step][lane]):
  //   f1a: layer-2 forward, K = layer-1 registers: w1a's full tiles, then the thin tile;
  //   f2a: layer-3 forward, K = layer-2 registers with the last tile folded;
  //   f2b: layer-3 backward, K = layer-3 registers with the last tile folded: full tiles, thin tile;
  //   f1b: layer-2 backward, K = layer-2 registers with the last tile folded;
  //   s3a: the last layer-3 tile (rows of group 0) in 4x4 blocks, K = layer-2 registers.
  const int e2 = 16 * (t2 - 1), e3 = 16 * (t3 - 1);
  const int s2 = 4 * (t2 - 1) + 1, s3 = 4 * (t3 - 1) + 1;  // K steps of the folded tables
  for (int lane = 0; lane < 64; ++lane) {
    const int m = lane % 16, q = lane / 16, m4 = lane % 4;
    const int c_last = L.logical(1, e2 + q), w_last = L.logical(2, e3 + q);  // folded features
    const int c_thin = L.logical(1, e2 + m4), w_thin = L.logical(2, e3 + m4);  // thin rows
    for (int at = 0; at < t2; ++at)
      for (int st = 0; st < 4 * t1; ++st) {
        float v;
        if (at < t2 - 1) {
          v = out[o.w1a + (at * 4 * t1 + st) * 64 + lane];
        } else {
          const int a = L.logical(0, 16 * (st / 4) + 4 * q + st % 4);
          v = (a >= 0 && c_thin >= 0) ? kTanhScale * K1[a * w1 + c_thin] : 0.f;
        }
        out[o.f1a + (at * 4 * t1 + st) * 64 + lane] = v;
      }
    for (int ot = 0; ot < t3; ++ot) {
      const int w = L.logical(2, 16 * ot + m);
      for (int st = 0; st < s2 - 1; ++st) out[o.f2a + (ot * s2 + st) * 64 + lane] = out[o.w2a + (ot * 4 * t2 + st) * 64 + lane];
      out[o.f2a + (ot * s2 + s2 - 1) * 64 + lane] = (c_last >= 0 && w >= 0) ? kTanhScale * K2[c_last * w2 + w] : 0.f;
    }
    for (int ct = 0; ct < t2; ++ct) {
      for (int st = 0; st < s3; ++st) {
        float v;
        if (ct < t2 - 1) {
          if (st < s3 - 1) {
            v = out[o.w2b + (ct * 4 * t3 + st) * 64 + lane];
          } else {
            const int c = L.logical(1, 16 * ct + m);
            v = (c >= 0 && w_last >= 0) ? K2[c * w2 + w_last] : 0.f;
          }
        } else {
          const int w = st < s3 - 1 ? L.logical(2, 16 * (st / 4) + 4 * q + st % 4) : w_last;
          v = (c_thin >= 0 && w >= 0) ? K2[c_thin * w2 + w] : 0.f;
        }
        out[o.f2b + (ct * s3 + st) * 64 + lane] = v;
      }
    }
    for (int at = 0; at < t1; ++at) {
      const int a = L.logical(0, 16 * at + m);
      for (int st = 0; st < s2 - 1; ++st) out[o.f1b + (at * s2 + st) * 64 + lane] = out[o.w1b + (at * 4 * t2 + st) * 64 + lane];
      out[o.f1b + (at * s2 + s2 - 1) * 64 + lane] = (a >= 0 && c_last >= 0) ? K1[a * w1 + c_last] : 0.f;
    }
    for (int r = 0; r < 4 * t2; ++r) {
      const int c = L.logical(1, 16 * (r / 4) + 4 * q + r % 4);
      out[o.s3a + r * 64 + lane] = (c >= 0 && w_thin >= 0) ? kTanhScale * K2[c * w2 + w_thin] : 0.f;
    }
  }
}

// ---------------------------------------------------------------------------------
// Device helpers

struct EmbedArgs {
  const float4* pos;           // (NA) wrapped fp32 positions, type-sorted per segment
  const SegInfo* seg;
  const int* blk_seg;          // block -> segment
  const int* nbr[kMaxT];       // (NA, cap_j) segment- and type-local indices, padded with N
  const int* nbr_cnt[kMaxT];   // (NA)
  int cap[kMaxT];
  int rowbase[kMaxT];
  int ntypes;
  float rcut, inv_rcut;
  float mean[kMaxT], inv_std[kMaxT];
  const float* net;            // packed nets, index (i*ntypes+j)*net_stride
  int net_stride;
  int base3[4], size3[4];
  int W, A;
  float inv_nnbrs;
  const float* tbias;
  float* T;                    // (NA, 10, W)
  float* G;                    // (mrows, A*W)
  const float* dG;             // (mrows, A*W)
  float* force;                // (R*N, 3)
  double* partial;             // (gridDim.x, 9) virial partials
  int* paircnt;                // (NA)
  int list_cap;                // LDS list entries per wave
  int net_bytes;
  const void* hnet;            // 16-bit operand tables (bf16 or fp16), per net hnet_stride bytes
  int hnet_stride, hnet_bytes;
  int lds_wave_floats;
  const float* tab;            // tabulated nets [net][interval][6][16*T3] (compressed model)
  int tab_n, tab_stride, tab_bytes;
  float tab_srmin[kMaxT * kMaxT], tab_delta[kMaxT * kMaxT];
  // backward of an atomic (centroid) model: dE/dT3[c][w] += wc_norm wc_g[atom][c] wc_y3[row][w]
  const float* wc_g;           // (NA, 3) or null
  const float* wc_y3;          // (mrows, wc_ldy)
  int wc_ldy;
  float wc_norm;
  int centre_mask;             // bit t: centres of type t are evaluated (an atomic model's nsel)
  // deterministic forces (backward): pair gradients by skin slot and centre parts instead of
  // force atomics; pgrad row of centre g and neighbour type j starts at (g ksum + koff[j]) 3
  float* pgrad;                // (NA, ksum, 3) or null (atomics)
  float* fcen;                 // (NA, 3)
  int ksum, koff[kMaxT];
  int g16;                     // forward: G stored as FP32 (0), BF16 (1) or FP16 (2) for the fitting GEMMs
  int dg16;                    // backward: dG stored as FP32 (0) or FP16 (1, the plain 16-bit fitting forms)
};

// Minimum image of d = x_j - x_i for positions wrapped into [0, L) (so |d| < L): the
// representative in [-L/2, L/2), as ((d + L/2) mod L) - L/2 gives it, branch free.
__device__ __forceinline__ float minimg(float d, float L) {
  const float h = L * 0.5f;
  d = d >= h ? d - L : d;
  return d < -h ? d + L : d;
}

__device__ __forceinline__ float safe_c(float x) { return fabsf(x) > 1e-15f ? x : 1e-15f; }

// |x| with components below 1e-15 replaced by 1e-15 (deepmd_jax/utils.py:81)
__device__ __forceinline__ float pair_r(float x, float y, float z) {
  float a = safe_c(x), b = safe_c(y), c = safe_c(z);
  return __builtin_amdgcn_sqrtf(a * a + b * b + c * c);
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

struct Geo {
  float x[3], u[3], r, inv_rr, sh, st, dsdr;
};

// Switch and environment row (deepmd_jax/utils.py:29-32, dpmodel.py:37-47) with
// reciprocals in place of divisions (1-ulp hardware rcp / sqrt).
__device__ __forceinline__ void env_forward(const float4 e, bool valid, float inv_rcut, float mean,
                                            float inv_std, Geo& g, float R[10]) {
  g.x[0] = valid ? e.x : 0.f;
  g.x[1] = valid ? e.y : 0.f;
  g.x[2] = valid ? e.z : 0.f;
  const float r = valid ? pair_r(g.x[0], g.x[1], g.x[2]) : 0.f;
  const float t = r * inv_rcut;
  const float sw = 1.f - 3.f * t * t + 2.f * t * t * t;
  const bool on = (r > 1e-14f) && (t < 1.f);
  const float inv = rcp(r + 1e-15f);
  const float s = on ? inv * sw : 0.f;
  g.r = r;
  g.dsdr = on ? ((-6.f * t + 6.f * t * t) * inv_rcut) * inv - sw * inv * inv : 0.f;
  g.sh = s * inv_std;
  g.st = valid ? (s - mean) * inv_std : 0.f;
  g.inv_rr = rcp(r + 1e-16f);
#pragma unroll
  for (int c = 0; c < 3; ++c) g.u[c] = g.x[c] * g.inv_rr;
  const float s3 = 1.7320508075688772f * g.sh, s33 = 3.f * g.sh;
  R[0] = g.sh;
#pragma unroll
  for (int c = 0; c < 3; ++c) R[1 + c] = s3 * g.u[c];
#pragma unroll
  for (int c = 0; c < 3; ++c) R[4 + c] = s33 * (g.u[c] * g.u[c] - (1.f / 3.f));
#pragma unroll
  for (int c = 0; c < 3; ++c) R[7 + c] = s33 * (1.4142135623730951f * g.u[c] * g.u[(c + 2) % 3]);
}

// dE/dx of one pair given dE/dR (10) and dE/dst.
__device__ __forceinline__ void env_backward(const Geo& g, const float dR[10], float dst, float inv_std,
                                             float out[3]) {
  const float s3 = 1.7320508075688772f, r2 = 1.4142135623730951f;
  float dsh = dR[0];
  float du[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int cm = (c + 2) % 3, cp = (c + 1) % 3;
    dsh += s3 * dR[1 + c] * g.u[c] + 3.f * dR[4 + c] * (g.u[c] * g.u[c] - (1.f / 3.f)) +
           3.f * r2 * dR[7 + c] * g.u[c] * g.u[cm];
    du[c] = s3 * g.sh * dR[1 + c] + 6.f * g.sh * g.u[c] * dR[4 + c] +
            3.f * r2 * g.sh * (dR[7 + c] * g.u[cm] + dR[7 + cp] * g.u[cp]);
  }
  const float ds = (dsh + dst) * inv_std;
  const float irr = g.inv_rr;
  float dr = ds * g.dsdr;
#pragma unroll
  for (int c = 0; c < 3; ++c) dr -= du[c] * g.x[c] * (irr * irr);
  const float inv_r = g.r > 0.f ? rcp(g.r) : 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float drdx = fabsf(g.x[c]) > 1e-15f ? g.x[c] * inv_r : 0.f;
    out[c] = du[c] * irr + dr * drdx;
  }
}

// The forward weight tables and biases are stored multiplied by kTanhScale = 2 log2(e)
// (embed_pack_net), so a pre-activation arrives as y = 2 log2(e) z and
// tanh(z) = 1 - 2 / (1 + 2^y) with the hardware exp2 and reciprocal, branch free (four
// instructions, two of them transcendental).  Absolute error ~2e-7 over the real line,
// saturating to +-1; near 0 the error stays absolute (no relative accuracy is needed:
// every activation enters a sum with O(1) terms).
__device__ __forceinline__ float tanh_s(float y) {
  const float e = __builtin_amdgcn_exp2f(y);
  return __builtin_fmaf(-2.f, __builtin_amdgcn_rcpf(1.f + e), 1.f);
}

// Interval of a tabulated net (deepmd_jax/utils.py:130-135): x = s~ - srmin,
// idx = trunc(x / delta), x0 = x - idx*delta; the coefficient row is idx brought into
// [0, n) by jax's gather rule (negative wraps once, then clamps) while x0 keeps the
// unclamped idx.  The quotient is bounded before the conversion (no overflow).
__device__ __forceinline__ int tab_interval(float st, float srmin, float delta, int n, float& x0) {
  const float x = st - srmin;
  const float qf = fminf(fmaxf(x / delta, -1073741824.f), 1073741824.f);
  const int idx = (int)qf;
  x0 = x - (float)idx * delta;
  return idx < 0 ? max(idx + n, 0) : min(idx, n - 1);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct H8;
template <>
struct H8<1> {
  typedef bf16x8 T;
  typedef __bf16 E;
  static __device__ __forceinline__ floatx4 mfma(T a, T b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct H8<2> {
  typedef halfx8 T;
  typedef _Float16 E;
  static __device__ __forceinline__ floatx4 mfma(T a, T b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// 16-bit operand bits of v (round to nearest even)
template <int PREC>
__device__ __forceinline__ uint16_t h16_bits(float v) {
  return __builtin_bit_cast(uint16_t, (typename H8<PREC>::E)v);
}

// Per-wave LDS of the 16-bit forms' FP16 dE/dT operand copies (floats): dTa [tile pair][lane
// group][x 0..10][8] for the dR products, dTb [slot][x 0..15] for the dh3 products.
// output width of the fold form: first width 4 (V1 - 1) + 1 (V1 - 1 features in each lane group, one
// more in group 0), widened Z2 and Z3 times by the repeat residuals
constexpr int fold_width(int V1, int Z2, int Z3) { return Z2 * Z3 * (4 * (V1 - 1) + 1); }
constexpr int embed_dt_floats(int T3) { return (10 * (16 * T3 + 1) + 3) / 4 * 4; }
constexpr int embed_dt16_floats(int T3) { return ((11 * ((T3 + 1) / 2) * 32 + 16 * T3 * 16) / 2 + 3) / 4 * 4; }

// B operand of a 16x16x32 step from 8 consecutive fp32 registers r0..r0+7 (zero past n)
template <int PREC, int NREG, typename GET>
__device__ __forceinline__ typename H8<PREC>::T pack8(GET get, int s2) {
  typedef typename H8<PREC>::E E;
  typename H8<PREC>::T v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (8 * s2 + j < NREG) ? (E)get(8 * s2 + j) : (E)0.f;
  return v;
}

// acc[ot] += sum_{s in [S0, S0+SN)} A(ot, s) * b(s), A read from the packed [ot][s][lane]
// table with row length S.  Weight operands are prefetched PF steps ahead; the
// scheduling barrier keeps the compiler from hoisting the whole table into registers.
typedef __amdgpu_buffer_rsrc_t rsrc_t;

// MFMA steps by which weight operands run ahead of their use
#ifndef DPMD_WEIGHT_PREFETCH
#define DPMD_WEIGHT_PREFETCH 2
#endif
#ifndef DPMD_WEIGHT_PREFETCH_H
#define DPMD_WEIGHT_PREFETCH_H 2
#endif

// weight read: byte offset = voff (per lane) + soff (wave-uniform)
__device__ __forceinline__ float bld(rsrc_t rs, int voff, int soff) {
  return __builtin_bit_cast(float, (unsigned)__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}

// tabulated-net coefficients: four consecutive slots of one power of one interval
__device__ __forceinline__ floatx4 tab_load(rsrc_t rs, int voff, int soff) {
  return __builtin_bit_cast(floatx4, (u32x4)__builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
}

// acc[0..OT) += A tiles (OT of them, S steps apart) x B steps S0 .. S0 + SN (bfun(step)), on
// 16x16x4 MFMAs; the last OS tiles on 4x4x1 blocks
template <int OT, int S, int S0, int SN, int OS = 0, typename BF, int NA>
__device__ __forceinline__ void mfma_chain(rsrc_t rs, int sbase, int vl4, BF bfun, floatx4 (&acc)[NA]) {
  static_assert(OT <= NA, "more tiles than accumulators");
  constexpr int PF = DPMD_WEIGHT_PREFETCH;
  float wb[PF + 1][OT];
#pragma unroll
  for (int p = 0; p < PF && p < SN; ++p)
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) wb[p][ot] = bld(rs, vl4, sbase + (ot * S + S0 + p) * 256);
#pragma unroll
  for (int s = 0; s < SN; ++s) {
    if (s + PF < SN) {
#pragma unroll
      for (int ot = 0; ot < OT; ++ot) wb[(s + PF) % (PF + 1)][ot] = bld(rs, vl4, sbase + (ot * S + S0 + s + PF) * 256);
    }
    const float b = bfun(S0 + s);
#pragma unroll
    for (int ot = 0; ot < OT; ++ot)
      acc[ot] = ot >= OT - OS ? mfma4x4(wb[s % (PF + 1)][ot], b, acc[ot]) : mfma4(wb[s % (PF + 1)][ot], b, acc[ot]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// acc += one tile of SN steps on 4x4x1 blocks alone: every weight load issued first and the
// steps spread over two accumulators (a lone 4x4 step is short next to a load or its own result)
template <int SN, typename BF>
__device__ __forceinline__ void thin_chain(rsrc_t rs, int sbase, int vl4, BF bfun, floatx4& acc) {
  float w[SN];
#pragma unroll
  for (int s = 0; s < SN; ++s) w[s] = bld(rs, vl4, sbase + s * 256);
  floatx4 acc1 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < SN; ++s) {
    if (s % 2 == 0) acc = mfma4x4(w[s], bfun(s), acc);
    else acc1 = mfma4x4(w[s], bfun(s), acc1);
  }
  acc += acc1;
}

// 16-bit A tables are [ot][s2][lane][8]: one 16-byte load per lane per 16x16x32 MFMA.
template <int PREC, int OT, int S2, int S0, int SN, typename BF>
__device__ __forceinline__ void mfma_chain_h(rsrc_t rs, int sbase, int vl16, BF bfun, floatx4 (&acc)[OT]) {
  typedef typename H8<PREC>::T HT;
  constexpr int PF = DPMD_WEIGHT_PREFETCH_H;
  HT wb[PF + 1][OT];
#pragma unroll
  for (int p = 0; p < PF && p < SN; ++p)
#pragma unroll
    for (int ot = 0; ot < OT; ++ot)
      wb[p][ot] = __builtin_bit_cast(HT, (u32x4)__builtin_amdgcn_raw_buffer_load_b128(
                                             rs, vl16, sbase + (ot * S2 + S0 + p) * 1024, 0));
#pragma unroll
  for (int s = 0; s < SN; ++s) {
    if (s + PF < SN) {
#pragma unroll
      for (int ot = 0; ot < OT; ++ot)
        wb[(s + PF) % (PF + 1)][ot] = __builtin_bit_cast(
            HT, (u32x4)__builtin_amdgcn_raw_buffer_load_b128(rs, vl16, sbase + (ot * S2 + S0 + s + PF) * 1024, 0));
    }
    const HT b = bfun(S0 + s);
#pragma unroll
    for (int ot = 0; ot < OT; ++ot) acc[ot] = H8<PREC>::mfma(wb[s % (PF + 1)][ot], b, acc[ot]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// A wave-uniform zero the compiler cannot see through: adding it to the per-chunk
// weight base keeps weight loads (and their addresses) inside the chunk loop.
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}

// f(integral_constant<G0>, integral_constant<size>) over tile groups [G0, G0+GS) < T
template <int G0, int T, int GS, typename F>
__device__ __forceinline__ void for_groups(F&& f) {
  if constexpr (G0 < T) {
    constexpr int SZ = (G0 + GS <= T) ? GS : (T - G0);
    f(std::integral_constant<int, G0>(), std::integral_constant<int, SZ>());
    for_groups<G0 + GS, T, GS>(f);
  }
}

constexpr int kGroup = 2;  // layer-3 tiles processed together
// Row stride (floats) of the tabulated backward's [neighbour][slot] tile of a group: a multiple
// of 4 (16-byte rows), 36 mod 64 so that 16 rows' 16-byte accesses hit distinct banks.
constexpr int kTS = 36;
static_assert(16 * kGroup <= kTS && kTS % 4 == 0, "a tile group's slots fit one 16-byte aligned row");

// Environment table of a super-chunk (16 SCH neighbours of a compacted row, processed as SCH
// chunks of 16): row l holds neighbour c0 + l's ten environment values in columns 0..9, zeros
// in columns 10 and 11 (the K padding of the dh3 products) and s~ in column kESc; then four
// zero floats for operand reads that run past the last row.  Each neighbour's environment is
// formed once (lane l % sn) instead of once per lane group.
constexpr int kES = 13;  // odd row stride: a column's 16 rows fall in distinct banks
constexpr int kESc = 12;
constexpr int es_floats(int sch) { return 16 * sch * kES + 4; }
// Row stride (floats) of the per-chunk [16 neighbours][env row] tiles (the FP32 forward, the
// backward): odd, so the 16 neighbours of a column land in distinct banks; columns 10..15 hold
// zeros.
constexpr int kRS = 17;
// LDS of the forward's environment rows: a 64-neighbour table, or the FP32 MLP form's tile (the
// table measured slower there)
constexpr int fwd_env_floats(int prec, bool tab) { return prec == 0 && !tab ? 16 * kRS : es_floats(4); }
// ... and of the backward's: a two-chunk table in the FP32 MLP form, the per-chunk tile in the
// 16-bit and tabulated forms (there the table saved less than its LDS round trips cost)
constexpr int bwd_env_floats(int prec, bool tab) { return prec == 0 && !tab ? es_floats(2) : 16 * kRS; }
__device__ __forceinline__ void env_table(const EmbedArgs& a, const float4* list, float* Et, int c0, int m, int sn,
                                          int lane, int ti) {
  const int l = lane % sn;
  const bool vl = c0 + l < m;
  const float4 e = vl ? list[c0 + l] : make_float4(0.f, 0.f, 0.f, 0.f);
  Geo g;
  float R[10];
  env_forward(e, vl, a.inv_rcut, a.mean[ti], a.inv_std[ti], g, R);
  if (lane < sn) {
    float* row = Et + l * kES;
#pragma unroll
    for (int x = 0; x < 10; ++x) row[x] = R[x];
    row[10] = 0.f;
    row[11] = 0.f;
    row[kESc] = g.st;
  }
  if (lane < 4) Et[sn * kES + lane] = 0.f;
}

// A pair's gradient gx = dE/dx_ij into the forces (atomics on the neighbour, the centre's part
// summed in fc) or, with deterministic forces, into the centre's pair-gradient row by skin slot
// (e.w); the virial -x (x) gx.  `on`: this lane owns a valid pair.
__device__ __forceinline__ void pair_out(const EmbedArgs& a, float* prow, bool on, const float4 e, const Geo& g,
                                         const float gx[3], float fc[3], float vir[9]) {
  if (on) {
    const int gj = __float_as_int(e.w);  // skin slot when prow is set
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (prow) prow[3 * gj + k] = gx[k];
      else atomicAdd(a.force + (size_t)gj * 3 + k, -gx[k]);
      fc[k] += gx[k];
    }
#pragma unroll
    for (int aa = 0; aa < 3; ++aa)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) vir[aa * 3 + bb] -= g.x[aa] * gx[bb];
  }
}

// Per-net first-layer weights and the three bias vectors (physical slot order) of all
// ntypes^2 nets, copied once per block into LDS: [net][k0 | b0 | b1 | b2].
// Thin tile (4x4-block form): lane group 0's rows of a tile from the per-group partial sums
// of the 4x4 blocks (sum over the four lane groups); the other groups' (padding) slots are zero
template <int D>
__device__ __forceinline__ void thin_rows(floatx4& acc, int q) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float v = 0.f;
    if (j < D) v = row_sum4(acc[j]);
    acc[j] = q == 0 ? v : 0.f;
  }
}

template <int T1, int T2, int T3>
__device__ __forceinline__ float* load_small_tables(const EmbedArgs& a, float* smem) {
  constexpr int SM = 16 * (2 * T1 + T2 + T3);
  const int nn = a.ntypes * a.ntypes;
  for (int i = threadIdx.x; i < nn * SM; i += blockDim.x) smem[i] = a.net[(size_t)(i / SM) * a.net_stride + i % SM];
  __syncthreads();
  return smem + (nn * SM + 3) / 4 * 4;
}

// Compact the skin row of neighbour type j to the pairs within rcut, in ascending
// index order, into the wave's LDS list (float4: x, y, z, global index bits).  With zrow
// (deterministic backward) the fourth word is the pair's skin slot instead, and the slots
// of dropped pairs get a zero gradient in zrow.
__device__ __forceinline__ int compact_row(const EmbedArgs& a, const SegInfo& S, int c, int j, float4 pc,
                                           float4* list, int lane, float* zrow = nullptr) {
  const int gc = S.base + c;
  const int cnt = a.nbr_cnt[j][gc];
  const int* row = a.nbr[j] + (size_t)gc * a.cap[j];
  // row entries through a buffer of cnt entries: lanes past the end read index 0 (a valid atom
  // of type j, masked below), so no lane branches around its loads, and the next piece's
  // indices are read while this piece's positions are gathered
  const rsrc_t rrs = __builtin_amdgcn_make_buffer_rsrc((void*)row, (short)0, cnt * 4, 0x00020000);
  const int gbase = S.base + S.off[j];
  int idx_next = __builtin_amdgcn_raw_buffer_load_b32(rrs, 4 * lane, 0, 0);
  int m = 0;
  for (int e0 = 0; e0 < cnt; e0 += 64) {
    const int e = e0 + lane;
    const int idx = idx_next;
    idx_next = __builtin_amdgcn_raw_buffer_load_b32(rrs, 4 * (e + 64), 0, 0);
    const int g = gbase + idx;
    const float4 pj = a.pos[g];
    float dx, dy, dz;
    if (S.tri) {
      // general cell: fractional minimum image (deepmd_jax/utils.py:24-27)
      float sf[3], d[3] = {pj.x - pc.x, pj.y - pc.y, pj.z - pc.z};
      cell_mul(S.Hi, d, sf);
#pragma unroll
      for (int k = 0; k < 3; ++k) sf[k] -= rintf(sf[k]);
      cell_mul(S.H, sf, d);
      dx = d[0], dy = d[1], dz = d[2];
    } else {
      dx = minimg(pj.x - pc.x, S.box[0]);
      dy = minimg(pj.y - pc.y, S.box[1]);
      dz = minimg(pj.z - pc.z, S.box[2]);
    }
    const float r = pair_r(dx, dy, dz);
    const bool in_row = e < cnt;
    const bool keep = in_row && (r > 1e-14f) && (r / a.rcut < 1.f);
    const float4 out = make_float4(dx, dy, dz, __int_as_float(zrow ? e : g));
    if (zrow && in_row && !keep) {
      zrow[3 * e] = 0.f;
      zrow[3 * e + 1] = 0.f;
      zrow[3 * e + 2] = 0.f;
    }
    const unsigned long long bal = __ballot(keep);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (keep) list[m + pre] = out;
    m += __popcll(bal);
  }
  return m;
}

// ---------------------------------------------------------------------------------
// Forward: T (10 x W) and G (A x W) per centre.  One wave per centre.

// V1: registers of layer 1 that hold a feature in some lane group (the rest are padding
// in every group); layer 2 then has V2 = Z2 V1 such registers.  MFMA k-steps and tanh
// evaluations over registers that are padding everywhere are skipped.
template <int T1, int T2, int T3, int Z2, int Z3, int PREC, int V1, bool TAB, bool FOLD>
__global__ void __launch_bounds__(256, 2) embed_fwd_kernel(const EmbedArgs a) {
  constexpr int V2 = Z2 * V1;
  static_assert(V1 <= 4 * T1 && V2 <= 4 * T2, "valid registers exceed the tiles");
  // fold form (FP32 MLP only): the last layer-2 tile (D2 rows of lane group 0) and the last
  // layer-3 tile (D3 rows) on 4x4x1 blocks; the last layer-2 tile's registers enter the
  // layer-3 products as one folded K step
  static_assert(!FOLD || (PREC == 0 && !TAB && Z2 * Z3 <= 4 && T2 >= 2 && T3 >= 2), "fold form");
  constexpr int D2 = Z2, D3 = Z2 * Z3;
  extern __shared__ float smem[];
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const rsrc_t wrs = __builtin_amdgcn_make_buffer_rsrc((void*)a.net, (short)0, a.net_bytes, 0x00020000);
  const rsrc_t hrs = __builtin_amdgcn_make_buffer_rsrc((void*)(PREC ? a.hnet : a.net), (short)0,
                                                       PREC ? a.hnet_bytes : a.net_bytes, 0x00020000);
  constexpr int S21 = (4 * T1 + 7) / 8, S22 = (4 * T2 + 7) / 8, S23 = (4 * T3 + 7) / 8;
  constexpr int H_W1A = 0, H_W2A = H_W1A + T2 * S21 * 1024;
  (void)S23;
  const int q = lane >> 4, n = lane & 15;
  float* small = smem;
  float* wbase = load_small_tables<T1, T2, T3>(a, smem) + wv * a.lds_wave_floats;
  float4* list = reinterpret_cast<float4*>(wbase);
  // ETAB: an environment table per super-chunk of SCH chunks (16-bit and tabulated forms); the
  // FP32 MLP form stages each chunk's R tile from every lane's own environment row instead (the
  // table measured slower there)
  constexpr bool ETAB = !(PREC == 0 && !TAB);
  constexpr int SCH = ETAB ? 4 : 1, SN = 16 * SCH, RS = ETAB ? kES : kRS;
  float* Et = wbase + 4 * a.list_cap;    // environment table [SN neighbours][RS]
  float* Hl = Et + fwd_env_floats(PREC, TAB);  // one tile group of h3: [16 neighbours][HG]
  constexpr int HG = 16 * kGroup + 1;
  // 16-bit MLP forms: the T reducer on FP16 MFMAs (R, h3: bounded values, FP16's 11-bit
  // significand; BF16 operands here would double the form's force error), h3 of a tile
  // group held FP16 as [slot][nbr] in Hl's space, row stride kSH
  constexpr bool H16 = PREC != 0 && !TAB;
  typedef typename H8<2>::T HT16;
  typedef typename H8<2>::E HE16;
  constexpr int kSH = 24;
  static_assert(16 * kGroup * kSH * 2 <= 16 * HG * 4, "16-bit h3 tile group exceeds its LDS");
  uint16_t* HlT = reinterpret_cast<uint16_t*>(Hl);

  const SegInfo& S = a.seg[a.blk_seg[blockIdx.x]];
  const int c = (blockIdx.x - S.blk0) * 4 + wv;
  if (c >= S.natoms) return;
  const int gc = S.base + c;
  int ti = 0;
  while (ti + 1 < a.ntypes && c >= S.off[ti + 1]) ++ti;
  if (!(a.centre_mask >> ti & 1)) return;
  const float4 pc = a.pos[gc];

  floatx4 Tacc[T3];
#pragma unroll
  for (int t = 0; t < T3; ++t) Tacc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  int pairs = 0;

  for (int j = 0; j < a.ntypes; ++j) {
    const int m = compact_row(a, S, c, j, pc, list, lane);
    pairs += m;
    wave_sync();
    // byte offsets of the net's tables (layout of embed_net_offsets)
    const int netb = __builtin_amdgcn_readfirstlane((ti * a.ntypes + j) * a.net_stride * 4);
    constexpr int O_K0 = 0, O_B0 = 16 * T1, O_B1 = 32 * T1, O_B2 = 32 * T1 + 16 * T2;
    constexpr int P0 = ((16 * T1 * 2 + 16 * T2 + 16 * T3) + 63) / 64 * 64;
    constexpr int O_W1A = P0, O_W2A = O_W1A + T2 * 4 * T1 * 64;
    constexpr int O_W2B = O_W2A + T3 * 4 * T2 * 64, O_W1B = O_W2B + T2 * 4 * T3 * 64;
    constexpr int SF2 = 4 * (T2 - 1) + 1, SF3 = 4 * (T3 - 1) + 1;  // K steps of the fold-form copies
    constexpr int O_F2A = O_W1B + T1 * 4 * T2 * 64, O_F2B = O_F2A + T3 * SF2 * 64;
    constexpr int O_F1B = O_F2B + T2 * SF3 * 64, O_F1A = O_F1B + T1 * SF2 * 64;
    constexpr int O_S3A = O_F1A + T2 * 4 * T1 * 64;
    const int vl4 = lane * 4, vl16 = lane * 16;
    const int hnetb = PREC ? __builtin_amdgcn_readfirstlane((ti * a.ntypes + j) * a.hnet_stride) : 0;
    for (int c0 = 0; c0 < m; c0 += SN) {
    // the super-chunk's environment rows, one neighbour per lane
    if constexpr (ETAB) {
      env_table(a, list, Et, c0, m, SN, lane, ti);
      wave_sync();
    }
    for (int cs = 0; cs < SCH && c0 + 16 * cs < m; ++cs) {
      const int nb = netb + opaque_zero();
      const int hb = hnetb + opaque_zero();
      const float* sm = small + (ti * a.ntypes + j) * (16 * (2 * T1 + T2 + T3)) + 4 * q + (nb - netb);
      auto k0 = [&](int r) { return sm[O_K0 + 16 * (r / 4) + (r % 4)]; };
      auto b0 = [&](int r) { return sm[O_B0 + 16 * (r / 4) + (r % 4)]; };
      auto b1 = [&](int r) { return sm[O_B1 + 16 * (r / 4) + (r % 4)]; };
      auto b2 = [&](int r) { return sm[O_B2 + 16 * (r / 4) + (r % 4)]; };
      // this chunk's R tile: rows are neighbours; columns x >= 10 that the reducer reads give
      // output rows of T that are not stored
      const float* Rl = Et + 16 * cs * RS;
      float st;
      if constexpr (ETAB) {
        st = Rl[n * RS + kESc];
      } else {
        const bool valid = c0 + n < m;
        const float4 e = valid ? list[c0 + n] : make_float4(0.f, 0.f, 0.f, 0.f);
        Geo g;
        float R[10];
        env_forward(e, valid, a.inv_rcut, a.mean[ti], a.inv_std[ti], g, R);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int x = 4 * q + i;  // lane group q stores env rows 4q..4q+3 of its neighbour
          float v = 0.f;
#pragma unroll
          for (int y = 0; y < 10; ++y) v = x == y ? R[y] : v;
          Et[n * RS + x] = v;
        }
        st = g.st;
      }
      // 16-bit forms: the reducer's A operand R[nbr 8q + j][x = n] (lane groups 0, 1; the
      // 16 neighbours fill half of the K = 32 of a 16x16x32 step)
      HT16 aR{};
      if constexpr (H16) {
#pragma unroll
        for (int j = 0; j < 8; ++j) aR[j] = (HE16)(q < 2 ? Rl[(8 * (q & 1) + j) * RS + n] : 0.f);
      }
      // reducer for one group of layer-3 tiles held in Hl:
      // T[x][slot] += sum_nbr R[nbr][x] h3[nbr][slot]
      auto reduce_group = [&](auto g0c, auto szc) {
        constexpr int g0 = decltype(g0c)::value, SZ = decltype(szc)::value;
        wave_sync();
        if constexpr (H16) {
          // h3 held 16-bit as [slot][nbr] (HlT): one 16-byte read per tile; lane groups 2, 3
          // re-read the finite values of groups 0, 1 against their zero A operand
#pragma unroll
          for (int k = 0; k < SZ; ++k) {
            const HT16 bH = *reinterpret_cast<const HT16*>(HlT + (16 * k + n) * kSH + 8 * (q & 1));
            Tacc[g0 + k] = H8<2>::mfma(aR, bH, Tacc[g0 + k]);
          }
        } else {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const float av = Rl[(4 * kk + q) * RS + n];
#pragma unroll
            for (int k = 0; k < SZ; ++k) Tacc[g0 + k] = mfma4(av, Hl[(4 * kk + q) * HG + 16 * k + n], Tacc[g0 + k]);
          }
        }
        wave_sync();
      };
      if constexpr (TAB) {
        // tabulated net: h3 = sum_c coeff[row][c] x0^(5-c) (Horner), four slots per load.
        // The gathers run with lane 4m + p on neighbour m, slots 4p..4p+3 of a tile: the four
        // consecutive lanes of a neighbour read its 64 contiguous bytes, one cache access where
        // lane (q, n) = 16q + n made four (the L1 handles about one access per clock, which
        // bounded these kernels); h3 lands in LDS in the same [neighbour][slot] layout.
        const int net = ti * a.ntypes + j;
        float x0;
        const int row = tab_interval(st, a.tab_srmin[net], a.tab_delta[net], a.tab_n, x0);
        const int gm = lane >> 2, gp = lane & 3;  // gather neighbour and quarter of this lane
        const int rowm = __shfl(row, gm);
        const float x0m = __shfl(x0, gm);
        const int voff = rowm * (6 * 64 * T3) + 16 * gp;
        const int tb = __builtin_amdgcn_readfirstlane(net * a.tab_stride * 4) + opaque_zero();
        const rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc((void*)a.tab, (short)0, a.tab_bytes, 0x00020000);
        for_groups<0, T3, kGroup>([&](auto g0c, auto szc) {
          constexpr int g0 = decltype(g0c)::value, SZ = decltype(szc)::value;
#pragma unroll
          for (int k = 0; k < SZ; ++k) {
            floatx4 cf[6];
#pragma unroll
            for (int cc = 0; cc < 6; ++cc) cf[cc] = tab_load(trs, voff + 64 * (g0 + k), tb + cc * 64 * T3);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              float v = cf[0][i];
#pragma unroll
              for (int cc = 1; cc < 6; ++cc) v = __builtin_fmaf(v, x0m, cf[cc][i]);
              Hl[gm * HG + 16 * k + 4 * gp + i] = v;
            }
          }
          reduce_group(g0c, szc);
        });
      } else {
        float h1[4 * T1];
#pragma unroll
        for (int r = 0; r < 4 * T1; ++r) h1[r] = r < V1 ? tanh_s(st * k0(r) + b0(r)) : 0.f;
        // layer accumulators start at the (scaled) bias: one VALU add fewer per activation
        floatx4 acc2[T2];
#pragma unroll
        for (int t = 0; t < T2; ++t) acc2[t] = floatx4{b1(4 * t), b1(4 * t + 1), b1(4 * t + 2), b1(4 * t + 3)};
        if constexpr (PREC == 0) {
          if constexpr (FOLD) {
            // full tiles and the thin last tile (4x4 blocks) in one chain
            mfma_chain<T2, 4 * T1, 0, V1, 1>(wrs, nb + 4 * O_F1A, vl4, [&](int s_) { return h1[s_]; }, acc2);
            thin_rows<D2>(acc2[T2 - 1], q);
          } else {
            mfma_chain<T2, 4 * T1, 0, V1>(wrs, nb + 4 * O_W1A, vl4, [&](int s_) { return h1[s_]; }, acc2);
          }
        } else {
          mfma_chain_h<PREC, T2, S21, 0, S21>(hrs, hb + H_W1A, vl16, [&](int s2) {
            return pack8<PREC, 4 * T1>([&](int r) { return h1[r]; }, s2); }, acc2);
        }
        float h2[4 * T2];
#pragma unroll
        for (int r = 0; r < 4 * T2; ++r) {
          float v = 0.f;
          if (r < V2) {
            v = tanh_s(acc2[r / 4][r % 4]);
            if (r / Z2 < 4 * T1) v += h1[r / Z2];
          }
          h2[r] = v;
        }
        // folded K step: group 0's last-tile layer-2 registers spread over the lane groups
        const float x2 = !FOLD ? 0.f : (D2 == 2 ? fold2(h2[4 * (T2 - 1)], h2[4 * (T2 - 1) + 1]) : h2[4 * (T2 - 1)]);
        for_groups<0, T3, kGroup>([&](auto g0c, auto szc) {
          constexpr int g0 = decltype(g0c)::value, SZ = decltype(szc)::value;
          floatx4 acc3[SZ];
#pragma unroll
          for (int t = 0; t < SZ; ++t) acc3[t] = floatx4{b2(4 * (g0 + t)), b2(4 * (g0 + t) + 1), b2(4 * (g0 + t) + 2), b2(4 * (g0 + t) + 3)};
          if constexpr (PREC == 0) {
            if constexpr (FOLD) {
              // the group's full tiles on the folded K steps; the last tile (last group) in 4x4 blocks
              constexpr bool LAST = g0 + SZ == T3;
              if constexpr (!LAST || SZ > 1)
                mfma_chain<LAST ? SZ - 1 : SZ, SF2, 0, SF2>(wrs, nb + 4 * (O_F2A + g0 * SF2 * 64), vl4,
                                                           [&](int s_) { return s_ < SF2 - 1 ? h2[s_] : x2; }, acc3);
              if constexpr (LAST) {
                thin_chain<V2>(wrs, nb + 4 * O_S3A, vl4, [&](int s_) { return h2[s_]; }, acc3[SZ - 1]);
                thin_rows<D3>(acc3[SZ - 1], q);
              }
            } else {
              mfma_chain<SZ, 4 * T2, 0, V2>(wrs, nb + 4 * (O_W2A + g0 * 4 * T2 * 64), vl4,
                                            [&](int s_) { return h2[s_]; }, acc3);
            }
          } else
            mfma_chain_h<PREC, SZ, S22, 0, S22>(hrs, hb + H_W2A + g0 * S22 * 1024, vl16, [&](int s2) {
              return pack8<PREC, 4 * T2>([&](int r) { return h2[r]; }, s2); }, acc3);
#pragma unroll
          for (int k = 0; k < SZ; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int r = 4 * (g0 + k) + i;
              float v = tanh_s(acc3[k][i]);
              if (r / Z3 < 4 * T2) v += h2[r / Z3];
              if constexpr (H16) HlT[(16 * k + 4 * q + i) * kSH + n] = h16_bits<2>(v);
              else Hl[n * HG + 16 * k + 4 * q + i] = v;
            }
          reduce_group(g0c, szc);
        });
      }
    }
    }
  }
  // T in logical order into LDS, then T and G to global
  float* Tl = wbase;
  // the fold form's layout fixes the output width (first width 4 (V1 - 1) + 1, doubled Z2 Z3 times)
  const int W = FOLD ? fold_width(V1, Z2, Z3) : a.W, A = a.A;
#pragma unroll
  for (int nt = 0; nt < T3; ++nt) {
    const int qq = n / 4, ii = n % 4, r = 4 * nt + ii;
    const int w = r < a.size3[qq] ? a.base3[qq] + r : -1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = 4 * q + i;
      if (x < 10 && w >= 0) Tl[x * W + w] = Tacc[nt][i] * a.inv_nnbrs;
    }
  }
  wave_sync();
  float* Tout = a.T + (size_t)gc * 10 * W;
  for (int e = lane; e < 10 * W; e += 64) Tout[e] = Tl[e];
  const int grow = a.rowbase[ti] + S.row[ti] + (c - S.off[ti]);
  float* Gout = a.G + (size_t)grow * A * W;
  const float* tb = a.tbias;
  const float r2 = 1.4142135623730951f;
  // the 16-bit fitting GEMMs read G in their operand format: rounded here once (the same
  // round-to-nearest-even conversion the GEMM would apply on staging), half the bytes
  uint16_t* Gh = reinterpret_cast<uint16_t*>(a.G) + (size_t)grow * A * W;
  auto gstore = [&](int e, float v) {
    if (a.g16 == 1) Gh[e] = __builtin_bit_cast(uint16_t, (__bf16)v);
    else if (a.g16 == 2) Gh[e] = __builtin_bit_cast(uint16_t, (_Float16)v);
    else Gout[e] = v;
  };
  if (A <= 16) {
    // G[a][w] = sum_x X[x][a] Y[x][w] (dpmodel.py:74-78) with Y = (TW, T3, T6) rows of T and
    // X[x][a] = TW[a], T3[c][a], then the six second-order factors
    // G2[k][a] = 3to6(T3[:, A + a])_k + T6[k][A + a]; on MFMA: M = a (16), K = x (10 of 12),
    // N = w in 16-wide tiles
    float* Xl = Tl + 10 * W;  // [12][16], rows 10, 11 and columns a >= A zero
    for (int e = lane; e < 12 * 16; e += 64) {
      const int x = e >> 4, aa = e & 15;
      float v = 0.f;
      if (x < 10 && aa < A) {
        if (x == 0) {
          v = Tl[aa] + tb[aa];
        } else if (x < 4) {
          v = Tl[x * W + aa];
        } else if (x < 7) {
          const float t3 = Tl[(x - 3) * W + A + aa];
          v = t3 * t3 + Tl[x * W + A + aa];
        } else {
          const int k = x - 7;
          v = r2 * Tl[(1 + k) * W + A + aa] * Tl[(1 + (k + 2) % 3) * W + A + aa] + Tl[x * W + A + aa];
        }
      }
      Xl[e] = v;
    }
    wave_sync();
    float xa[3];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) xa[kk] = Xl[(4 * kk + q) * 16 + n];
    const int ntile = (W + 15) / 16;
    for (int t = 0; t < ntile; ++t) {
      const int w = 16 * t + n;
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const int x = 4 * kk + q;
        float y = 0.f;
        if (x < 10 && w < W) y = Tl[x * W + w] + (x == 0 ? tb[w] : 0.f);
        acc = mfma4(xa[kk], y, acc);
      }
      if (w < W) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int aa = 4 * q + i;
          if (aa < A) gstore(aa * W + w, acc[i]);
        }
      }
    }
  } else {
    for (int e = lane; e < A * W; e += 64) {
      const int aa = e / W, w = e % W;
      const float tww = Tl[w] + tb[w], twa = Tl[aa] + tb[aa];
      float v = tww * twa;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) v += Tl[(1 + cc) * W + w] * Tl[(1 + cc) * W + aa];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float t3 = Tl[(1 + k) * W + A + aa];
        const float g2 = t3 * t3 + Tl[(4 + k) * W + A + aa];
        v += g2 * Tl[(4 + k) * W + w];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float g2 = r2 * Tl[(1 + k) * W + A + aa] * Tl[(1 + (k + 2) % 3) * W + A + aa] +
                         Tl[(7 + k) * W + A + aa];
        v += g2 * Tl[(7 + k) * W + w];
      }
      gstore(e, v);
    }
  }
  if (lane == 0) a.paircnt[gc] = pairs;
}

// ---------------------------------------------------------------------------------
// Backward: forces (scatter) and virial partials.  One wave per centre.

template <int T1, int T2, int T3, int Z2, int Z3, int PREC, int V1, bool TAB, bool FOLD>
__global__ void __launch_bounds__(256, FOLD ? 4 : 2) embed_bwd_kernel(const EmbedArgs a) {
  constexpr int V2 = Z2 * V1;
  static_assert(V1 <= 4 * T1 && V2 <= 4 * T2, "valid registers exceed the tiles");
  // fold form (FP32 MLP only): the thin tiles -- the last layer-2 tile (D2 rows of lane group 0),
  // forward and backward, and the last layer-3 tile (D3 rows) of the forward recompute -- on
  // 4x4x1 blocks; the registers of the last layer-2 and layer-3 tiles enter the products whose
  // K runs over them as one folded K step each
  static_assert(!FOLD || (PREC == 0 && !TAB && Z2 * Z3 <= 4 && T2 >= 2 && T3 >= 2), "fold form");
  constexpr int D2 = Z2, D3 = Z2 * Z3;
  extern __shared__ float smem[];
  __shared__ double vir_s[4][9];
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const rsrc_t wrs = __builtin_amdgcn_make_buffer_rsrc((void*)a.net, (short)0, a.net_bytes, 0x00020000);
  const rsrc_t hrs = __builtin_amdgcn_make_buffer_rsrc((void*)(PREC ? a.hnet : a.net), (short)0,
                                                       PREC ? a.hnet_bytes : a.net_bytes, 0x00020000);
  constexpr int S21 = (4 * T1 + 7) / 8, S22 = (4 * T2 + 7) / 8, S23 = (4 * T3 + 7) / 8;
  constexpr int H_W1A = 0, H_W2A = H_W1A + T2 * S21 * 1024, H_W2B = H_W2A + T3 * S22 * 1024,
                H_W1B = H_W2B + T2 * S23 * 1024;
  static_assert(kGroup * 4 == 8, "16-bit backward assumes two layer-3 tiles per group");
  const int q = lane >> 4, n = lane & 15;
  constexpr int HS = 16 * T3 + 1;
  // the fold form's layout fixes the output width (compile-time addresses of T and dE/dG)
  const int W = FOLD ? fold_width(V1, Z2, Z3) : a.W, A = a.A;
  float* small = smem;
  float* wbase = load_small_tables<T1, T2, T3>(a, smem) + wv * a.lds_wave_floats;
  // chunk-phase layout
  // 16-bit MLP forms: the dR and dh3 products on FP16 MFMAs (dE/dT, h3, R: bounded values;
  // FP16's 11-bit significand), from FP16 copies of dE/dT laid out as their operands (dTa, dTb)
  constexpr bool H16 = PREC != 0 && !TAB;
  typedef typename H8<2>::T HT16;
  typedef typename H8<2>::E HE16;
  constexpr int NG2 = (T3 + 1) / 2;
  constexpr int DTA = 11 * NG2 * 32;
  constexpr int DTB = 16 * T3 * 16;
  // the copies occupy the start of the wave's LDS in the chunk phase (dTp, which they are
  // made from, lives there only during the prologue)
  uint16_t* dTa = reinterpret_cast<uint16_t*>(wbase);
  uint16_t* dTb = dTa + DTA;
  // dE/dT rows 0..9.  MFMA operand rows beyond the ten environment rows read row 9 instead:
  // in the dR products those rows give output rows x >= 10, which nothing reads; in the dh3
  // products they meet the zero columns x >= 10 of the R tile (finite x 0 = 0).
  constexpr int ODT = embed_dt_floats(T3);
  float* dTp = wbase;                    // 10 * HS
  // ETAB (the FP32 MLP form): an environment table per super-chunk of SCH = 2 chunks, the pair
  // gradients formed once per neighbour after it; the 16-bit and tabulated forms stage each
  // chunk's R tile from every lane's own environment row (the table measured slower there)
  constexpr bool ETAB = PREC == 0 && !TAB;
  constexpr int SCH = ETAB ? 2 : 1, SN = 16 * SCH, RS = ETAB ? kES : kRS;
  float* Et = wbase + (H16 ? embed_dt16_floats(T3) : ODT);  // [SN neighbours][RS]
  float4* list = reinterpret_cast<float4*>(Et + bwd_env_floats(PREC, TAB));  // 16-byte aligned
  // tabulated nets: h3 and p' of one tile group, [neighbour][slot] with row stride kTS
  float* Tb = reinterpret_cast<float*>(list + a.list_cap);

  const SegInfo& S = a.seg[a.blk_seg[blockIdx.x]];
  const int c = (blockIdx.x - S.blk0) * 4 + wv;
  bool active = c < S.natoms;
  if (active) {
    int tc = 0;
    while (tc + 1 < a.ntypes && c >= S.off[tc + 1]) ++tc;
    active = (a.centre_mask >> tc & 1) != 0;
  }
  float vir[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) vir[k] = 0.f;

  if (active) {
    const int gc = S.base + c;
    int ti = 0;
    while (ti + 1 < a.ntypes && c >= S.off[ti + 1]) ++ti;
    const float4 pc = a.pos[gc];
    const int grow = a.rowbase[ti] + S.row[ti] + (c - S.off[ti]);
    // ---- prologue: dE/dT = (dG/dT)^T dE/dG / Nnbrs into the physical slot layout dTp.
    // G[a][w] = TW[w]TW[a] + sum_c T3[c][w]T3[c][a] + sum_k G2[k][a]T6[k][w] (dpmodel.py:74-78)
    //   V[x][w] = sum_a dG[a][w] X[x][a],  X = (TW, T3 rows)[:, :A] for x<4, G2 for x>=4
    //   U[x][a] = sum_w dG[a][w] Tx[x][w]   (row 0 of Tx is TW)
    // plus the U terms of the factors T[:, :A] and of G2's arguments T[:, A:2A].
    {
      const float* Tg = a.T + (size_t)gc * 10 * W;
      // the centre's dE/dG rows: FP32, or FP16 after a plain 16-bit fitting
      const size_t dgo = (size_t)grow * A * W;
      const void* dGv = a.dg16 ? (const void*)(reinterpret_cast<const uint16_t*>(a.dG) + dgo) : (const void*)(a.dG + dgo);
      auto dG_at = [&](int e) -> float {
        return a.dg16 ? (float)__builtin_bit_cast(_Float16, reinterpret_cast<const uint16_t*>(dGv)[e])
                      : reinterpret_cast<const float*>(dGv)[e];
      };
      float* Tl = dTp + ODT;       // 10*W   (aliases Rl and list)
      float* Xl = Tl + 10 * W;     // 10*A
      float* Ul = Xl + 10 * A;     // 10*A
      // compile-time trip count (10 W <= 160 T3): all loads of T issue before the stores.  Buffer
      // loads over T and the bias vector: entries past their ends read as zero, so no lane
      // branches around its load (a branch per load had each one wait for the previous)
      constexpr int TN = (160 * T3 + 63) / 64;
      const rsrc_t Trs = __builtin_amdgcn_make_buffer_rsrc((void*)Tg, (short)0, 10 * W * 4, 0x00020000);
      const rsrc_t Brs = __builtin_amdgcn_make_buffer_rsrc((void*)a.tbias, (short)0, W * 4, 0x00020000);
      float tv[TN];
#pragma unroll
      for (int k = 0; k < TN; ++k) {
        const int e = lane + 64 * k;
        // the bias has W <= 16 T3 entries: only the first passes can hold one
        tv[k] = bld(Trs, 4 * e, 0) + (64 * k < 16 * T3 ? bld(Brs, 4 * e, 0) : 0.f);
      }
      // the dG operands of U and V (A <= 16), loaded with T so that the two reads overlap.  W <=
      // 16 T3 (the layer-3 tiles hold every output feature): compile-time trip counts; masked
      // entries are buffer loads past the end of the centre's dG rows, which read as zero
      const rsrc_t Grs = __builtin_amdgcn_make_buffer_rsrc((void*)dGv, (short)0, A * W * (a.dg16 ? 2 : 4), 0x00020000);
      float dgu[4 * T3];  // dG[a = n][w = 4 kk + q]
      float dgv[T3][4];   // dG[a = 4 kk + q][w = 16 t + n]
      // entry e of the centre's rows (e = A W: past the end, reads zero), in either format
      auto load_dg = [&](auto g16c) {
        constexpr bool G16 = decltype(g16c)::value;
        auto ld = [&](int e) -> float {
          if constexpr (G16)
            return (float)__builtin_bit_cast(_Float16, (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(Grs, 2 * e, 0, 0));
          else
            return bld(Grs, 4 * e, 0);
        };
        // rows a >= A start at or past entry A W and read zero like the entries past the end; with a
        // compile-time W (fold form) the offsets are a per-lane base plus constants
#pragma unroll
        for (int kk = 0; kk < 4 * T3; ++kk) {
          const int w = 4 * kk + q;
          dgu[kk] = ld(w < W ? n * W + w : A * W);
        }
#pragma unroll
        for (int t = 0; t < T3; ++t)
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const int aa = 4 * kk + q, w = 16 * t + n;
            dgv[t][kk] = ld(w < W ? aa * W + w : A * W);
          }
      };
      if (A <= 16) {
        if (a.dg16) load_dg(std::true_type{});
        else load_dg(std::false_type{});
      }
      // the region holds TN * 64 floats: the zeros past 10 W are stored as well (no branch)
#pragma unroll
      for (int k = 0; k < TN; ++k) Tl[lane + 64 * k] = tv[k];
      // with A <= 16 the X and U tables use dTp's space, which is cleared after them
      if (A > 16)
        for (int e = lane; e < 10 * HS; e += 64) dTp[e] = 0.f;
      wave_sync();
      const float r2 = 1.4142135623730951f;
      if (A <= 16) {
        // On MFMA: U = Tx dG^T (M = x, K = w, N = a) and V = X dG (M = x, K = a, N = w); the U
        // terms are added to V in its output registers, lane (q, n) holding V[4q + i][16t + n].
        float* Xm = dTp;           // [12][16]: X[x][a], rows 10, 11 and columns a >= A zero
        float* Um = Xm + 12 * 16;  // [10][16]: U[x][a]
        static_assert(12 * 16 + 10 * 16 <= 10 * HS, "X and U fit dTp's space");
        // X[x][a], x = 4k + q (lane group q, pass k), a = n: every candidate term read from an
        // in-range row and selected per lane group, no branches.  x < 4: T[x][a]; 4 <= x < 7:
        // T3[x-3][A+a]^2 + T[x][A+a]; 7 <= x < 10: sqrt2 T3[kp][A+a] T3[(kp+2)%3][A+a] + T[x][A+a]
        // (kp = x - 7; rows of T3 are rows 1..3 of T)
        {
          const int aa = n;
          const float* TA = Tl + A + aa;
          const float x0 = Tl[q * W + aa];
          const float t31 = TA[(1 + (q < 3 ? q : 0)) * W];
          const float x1 = q < 3 ? t31 * t31 + TA[(4 + q) * W] : r2 * TA[W] * TA[3 * W] + TA[7 * W];
          const int q2 = q < 2 ? q : 1;
          const float x2 = q < 2 ? r2 * TA[(2 + q2) * W] * TA[(1 + q2) * W] + TA[(8 + q2) * W] : 0.f;
          const bool in = aa < A;
          Xm[lane] = in ? x0 : 0.f;
          Xm[lane + 64] = in ? x1 : 0.f;
          Xm[lane + 128] = in ? x2 : 0.f;
        }
        {
          floatx4 u0 = floatx4{0.f, 0.f, 0.f, 0.f}, u1 = u0;
          // rows n >= 10 read row 9 (their outputs are not kept); columns w >= W read past
          // the row, into the next one or the zeros after T, against a zero dG entry
#pragma unroll
          for (int kk = 0; kk < 4 * T3; ++kk) {
            const float av = Tl[(n < 10 ? n : 9) * W + 4 * kk + q];
            if (kk & 1) u1 = mfma4(av, dgu[kk], u1);
            else u0 = mfma4(av, dgu[kk], u0);
          }
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (4 * q + i < 10) Um[(4 * q + i) * 16 + n] = u0[i] + u1[i];
        }
        wave_sync();
        float xa[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) xa[kk] = n < 12 ? Xm[n * 16 + 4 * kk + q] : 0.f;
        // centroid-model hook: the output out_norm sum_w f_w T3[c][w] depends on T3 directly
        float gv[3] = {0.f, 0.f, 0.f};
        if (a.wc_g) {
          const float* gg = a.wc_g + (size_t)gc * 3;
          gv[0] = gg[0], gv[1] = gg[1], gv[2] = gg[2];
        }
        // results held until X and U have been read, then written over the cleared dTp.  Lane
        // (q, n) holds V[x = 4q + i][w = 16t + n]; the U terms reach w < 2A <= 32 only (tiles 0,
        // 1).  Lane group 0 (x = i < 4) adds U[x][w] for w < A and, for x >= 1 and A <= w < 2A,
        // the terms of the second-order factors; groups 1-3 (x >= 4) add U[x][w - A] for A <= w
        // < 2A.  Every read is in range (clamped index) and selected, no branches.
        float res[T3][4];
        int rslot[T3];
#pragma unroll
        for (int t = 0; t < T3; ++t) {
          const int w = 16 * t + n;
          floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) acc = mfma4(xa[kk], dgv[t][kk], acc);
          int slot = -1;
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) {
            const int rr = w - a.base3[qq];
            slot = (rr >= 0 && rr < a.size3[qq]) ? 16 * (rr >> 2) + 4 * qq + (rr & 3) : slot;
          }
          rslot[t] = w < W ? slot : -1;
          float y = 0.f;
          if (a.wc_g && w < W) y = a.wc_norm * a.wc_y3[(size_t)grow * a.wc_ldy + w];
          const bool lo = w < A, mid = w >= A && w < 2 * A;
          const int wl = lo ? w : 0, wm = mid ? w : 0, am = mid ? w - A : 0;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            float v = acc[i];
            if (t < 2) {
              float c0 = lo ? Um[i * 16 + wl] : 0.f;
              if (i >= 1) {
                const int cc = i - 1, cm = (cc + 2) % 3, cp = (cc + 1) % 3;
                const float m2 = 2.f * Tl[i * W + wm] * Um[(4 + cc) * 16 + am] +
                                 r2 * Tl[(1 + cm) * W + wm] * Um[(7 + cc) * 16 + am] +
                                 r2 * Tl[(1 + cp) * W + wm] * Um[(7 + cp) * 16 + am];
                c0 += mid ? m2 : 0.f;
              }
              const int xr = 4 * q + i < 10 ? 4 * q + i : 9;
              const float c1 = mid ? Um[xr * 16 + am] : 0.f;
              v += q == 0 ? c0 : c1;
            }
            if (i >= 1) v += q == 0 ? gv[i - 1] * y : 0.f;
            res[t][i] = v * a.inv_nnbrs;
          }
        }
        wave_sync();
        if constexpr (H16) {
          // the FP16 operand copies of the 16-bit forms, written from the registers (layouts
          // below); entries no result reaches (padding slots, row x = 10) stay zero.  Stores
          // without a target go to a scratch word past the copies (the chunk tiles, written
          // before they are read)
          constexpr int D16 = embed_dt16_floats(T3);
          // whole clearing passes: the last one may run into the chunk tiles (written before read)
#pragma unroll
          for (int k = 0; k < (D16 + 63) / 64; ++k) wbase[lane + 64 * k] = 0.f;
          wave_sync();
          uint16_t* h16 = reinterpret_cast<uint16_t*>(wbase);
#pragma unroll
          for (int t = 0; t < T3; ++t) {
            const int sl = rslot[t];
            const int tt = sl >> 4, qq = (sl >> 2) & 3, jj = 4 * (tt & 1) + (sl & 3);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int x = 4 * q + i;
              const bool ok = sl >= 0 && x < 10;
              const uint16_t h = h16_bits<2>(res[t][i]);
              h16[ok ? ((((tt >> 1) * 4 + qq) * 11 + x) << 3) + jj : 2 * D16 + lane] = h;
              h16[ok ? DTA + sl * 16 + x : 2 * D16 + 64 + lane] = h;
            }
          }
        } else {
          // whole clearing passes (the last one runs into T's space, no longer read); stores
          // without a target go to a scratch word in T's space
          constexpr int ZP = (10 * HS + 63) / 64;
          static_assert(64 * ZP <= ODT + 160 * T3, "clearing passes stay in the wave's space");
#pragma unroll
          for (int k = 0; k < ZP; ++k) dTp[lane + 64 * k] = 0.f;
          wave_sync();
#pragma unroll
          for (int t = 0; t < T3; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const bool ok = rslot[t] >= 0 && 4 * q + i < 10;
              dTp[ok ? (4 * q + i) * HS + rslot[t] : 64 * ZP + lane] = res[t][i];
            }
        }
      } else {
        for (int e = lane; e < 10 * A; e += 64) {
          const int x = e / A, aa = e % A;
          float v;
          if (x < 4) {
            v = Tl[x * W + aa];
          } else if (x < 7) {
            const int k = x - 4;
            const float t3 = Tl[(1 + k) * W + A + aa];
            v = t3 * t3 + Tl[x * W + A + aa];
          } else {
            const int kp = x - 7;
            v = r2 * Tl[(1 + kp) * W + A + aa] * Tl[(1 + (kp + 2) % 3) * W + A + aa] + Tl[x * W + A + aa];
          }
          Xl[e] = v;
        }
        {
          // U: lane (a = lane%16, wq = lane/16) sums w = wq mod 4, then the 4 lanes of a combine
          const int aa = lane & 15, wq = lane >> 4;
          float u[10];
#pragma unroll
          for (int x = 0; x < 10; ++x) u[x] = 0.f;
          if (aa < A) {
            for (int w = wq; w < W; w += 4) {
              const float d = dG_at(aa * W + w);
#pragma unroll
              for (int x = 0; x < 10; ++x) u[x] += d * Tl[x * W + w];
            }
          }
#pragma unroll
          for (int x = 0; x < 10; ++x) {
            u[x] += __shfl_xor(u[x], 16);
            u[x] += __shfl_xor(u[x], 32);
          }
          if (wq == 0 && aa < A) {
#pragma unroll
            for (int x = 0; x < 10; ++x) Ul[x * A + aa] = u[x];
          }
        }
        wave_sync();
        for (int w = lane; w < W; w += 64) {
          float v[10];
#pragma unroll
          for (int x = 0; x < 10; ++x) v[x] = 0.f;
          for (int aa = 0; aa < A; ++aa) {
            const float d = dG_at(aa * W + w);
#pragma unroll
            for (int x = 0; x < 10; ++x) v[x] += d * Xl[x * A + aa];
          }
          if (w < A) {
#pragma unroll
            for (int x = 0; x < 4; ++x) v[x] += Ul[x * A + w];
          } else if (w < 2 * A) {
            const int aa = w - A;
#pragma unroll
            for (int x = 4; x < 10; ++x) v[x] += Ul[x * A + aa];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
              v[1 + cc] += 2.f * Tl[(1 + cc) * W + w] * Ul[(4 + cc) * A + aa] +
                           r2 * Tl[(1 + (cc + 2) % 3) * W + w] * Ul[(7 + cc) * A + aa] +
                           r2 * Tl[(1 + (cc + 1) % 3) * W + w] * Ul[(7 + (cc + 1) % 3) * A + aa];
            }
          }
          int slot = -1;
          for (int qq = 0; qq < 4; ++qq) {
            const int rr = w - a.base3[qq];
            if (rr >= 0 && rr < a.size3[qq]) slot = 16 * (rr / 4) + 4 * qq + (rr % 4);
          }
          if (a.wc_g) {
            // the centroid output out_norm sum_w f_w T3[c][w] depends on T3 directly
            const float y = a.wc_norm * a.wc_y3[(size_t)grow * a.wc_ldy + w];
            const float* gg = a.wc_g + (size_t)gc * 3;
            v[1] += gg[0] * y;
            v[2] += gg[1] * y;
            v[3] += gg[2] * y;
          }
          if (slot >= 0) {
#pragma unroll
            for (int x = 0; x < 10; ++x) dTp[x * HS + slot] = v[x] * a.inv_nnbrs;
          }
        }
      }
      wave_sync();
      if (H16 && A > 16) {
        // dTa[pair g][lane group][x][j] = dT[x][16 (2g + j/4) + 4 lane group + j%4] (x = 10
        // zero): the 16 lanes of a group read 16 consecutive byte-16 pieces.  dTb[slot][x] =
        // dT[x][slot] (x >= 10 zero).  Both overwrite dTp: read into registers first.  (With
        // A <= 16 the copies were written from the registers above.)
        static_assert(DTA % 64 == 0 && DTB % 64 == 0, "conversion loops assume whole passes");
        float va[DTA / 64], vb[DTB / 64];
#pragma unroll
        for (int it = 0; it < DTA / 64; ++it) {
          const int e = lane + 64 * it;
          const int j = e & 7, rest = e >> 3, x = rest % 11, gq = rest / 11;
          const int qq = gq & 3, t = 2 * (gq >> 2) + j / 4;
          va[it] = x < 10 && t < T3 ? dTp[x * HS + 16 * t + 4 * qq + (j & 3)] : 0.f;
        }
#pragma unroll
        for (int it = 0; it < DTB / 64; ++it) {
          const int e = lane + 64 * it;
          const int x = e & 15, slot = e >> 4;
          vb[it] = x < 10 ? dTp[x * HS + slot] : 0.f;
        }
        wave_sync();
#pragma unroll
        for (int it = 0; it < DTA / 64; ++it) dTa[lane + 64 * it] = h16_bits<2>(va[it]);
#pragma unroll
        for (int it = 0; it < DTB / 64; ++it) dTb[lane + 64 * it] = h16_bits<2>(vb[it]);
        wave_sync();
      }
    }
    // ---- chunks
    // per-lane dE/dT operand rows: dR products use row n, dh3 products rows 4 kk + q
    // (only kk = 2 can pass row 9); rows past 9 read the zero row
    const float* dTr = dTp + (n < 10 ? n : 9) * HS + 4 * q;
    const float* dTc2 = dTp + (8 + q < 10 ? 8 + q : 9) * HS + n;
    // fold form: the folded step of dR reads slot (group 0, register 4 (T3 - 1) + q) of row n
    const float* dTf = dTp + (n < 10 ? n : 9) * HS + q;
    (void)dTf;
    float fc[3] = {0.f, 0.f, 0.f};
    for (int j = 0; j < a.ntypes; ++j) {
      float* prow = a.pgrad ? a.pgrad + ((size_t)gc * a.ksum + a.koff[j]) * 3 : nullptr;
      const int m = compact_row(a, S, c, j, pc, list, lane, prow);
      wave_sync();
      const int netb = __builtin_amdgcn_readfirstlane((ti * a.ntypes + j) * a.net_stride * 4);
      constexpr int O_K0 = 0, O_B0 = 16 * T1, O_B1 = 32 * T1, O_B2 = 32 * T1 + 16 * T2;
      constexpr int P0 = ((16 * T1 * 2 + 16 * T2 + 16 * T3) + 63) / 64 * 64;
      constexpr int O_W1A = P0, O_W2A = O_W1A + T2 * 4 * T1 * 64;
      constexpr int O_W2B = O_W2A + T3 * 4 * T2 * 64, O_W1B = O_W2B + T2 * 4 * T3 * 64;
      constexpr int SF2 = 4 * (T2 - 1) + 1, SF3 = 4 * (T3 - 1) + 1;  // K steps of the fold-form copies
      constexpr int O_F2A = O_W1B + T1 * 4 * T2 * 64, O_F2B = O_F2A + T3 * SF2 * 64;
      constexpr int O_F1B = O_F2B + T2 * SF3 * 64, O_F1A = O_F1B + T1 * SF2 * 64;
      constexpr int O_S3A = O_F1A + T2 * 4 * T1 * 64;
      const int vl4 = lane * 4, vl16 = lane * 16;
      const int hnetb = PREC ? __builtin_amdgcn_readfirstlane((ti * a.ntypes + j) * a.hnet_stride) : 0;
      for (int c0 = 0; c0 < m; c0 += SN) {
      if constexpr (ETAB) {
        // the super-chunk's environment rows, one neighbour per lane (lane l % SN)
        env_table(a, list, Et, c0, m, SN, lane, ti);
        wave_sync();
      }
      for (int cs = 0; cs < SCH && c0 + 16 * cs < m; ++cs) {
        const int ch = c0 + 16 * cs;
        const int nb = netb + opaque_zero();
        const int hb = hnetb + opaque_zero();
        const float* sm = small + (ti * a.ntypes + j) * (16 * (2 * T1 + T2 + T3)) + 4 * q + (nb - netb);
        auto k0 = [&](int r) { return sm[O_K0 + 16 * (r / 4) + (r % 4)]; };
        auto b0 = [&](int r) { return sm[O_B0 + 16 * (r / 4) + (r % 4)]; };
        auto b1 = [&](int r) { return sm[O_B1 + 16 * (r / 4) + (r % 4)]; };
        auto b2 = [&](int r) { return sm[O_B2 + 16 * (r / 4) + (r % 4)]; };
        // this chunk's R tile (rows n, the FP32 dh3 products' B operand); dE/dR (and with ETAB
        // dE/ds~) is written over it: the wave's last reads of R (the dh3 products) and s~ are
        // issued before its first writes, and one wave's LDS operations complete in order
        float* Rl = Et + 16 * cs * RS;
        float* dRl = Rl;
        const bool valid = ch + n < m;
        const float4 e = (!ETAB && valid) ? list[ch + n] : make_float4(0.f, 0.f, 0.f, 0.f);
        Geo g;
        float st;
        // 16-bit MLP forms: the dh3 products' B operand R[nbr n][x = 8q + j] from this lane's R
        HT16 bR{};
        if constexpr (ETAB) {
          st = Rl[n * RS + kESc];
        } else {
          float R[10];
          env_forward(e, valid, a.inv_rcut, a.mean[ti], a.inv_std[ti], g, R);
          st = g.st;
          // the R tile feeds the FP32 dh3 products; the 16-bit ones take R from registers
          if constexpr (!H16) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int x = 4 * q + i;
              float v = 0.f;
#pragma unroll
              for (int y = 0; y < 10; ++y) v = x == y ? R[y] : v;
              Rl[n * RS + x] = v;
            }
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) bR[j] = (HE16)(q == 0 ? R[j] : (q == 1 && j < 2) ? R[8 + j] : 0.f);
          }
        }
        float dst;
        if constexpr (TAB) {
          // tabulated net: h3 = p(x0) and dh3/ds~ = p'(x0) from one read of the interval's
          // coefficients; dE/ds~ = sum_slot dh3[slot] p'[slot] (no gradient through idx)
          // The gathers run with lane 4m + p on neighbour m (coalesced, as in the forward);
          // h3 and p' pass through LDS ([neighbour][slot], Tb) into the (q, n) layout the MFMA
          // operands need.
          const int net = ti * a.ntypes + j;
          float x0;
          const int row = tab_interval(st, a.tab_srmin[net], a.tab_delta[net], a.tab_n, x0);
          const int gm = lane >> 2, gp = lane & 3;
          const int rowm = __shfl(row, gm);
          const float x0m = __shfl(x0, gm);
          const int voff = rowm * (6 * 64 * T3) + 16 * gp;
          const int tb = __builtin_amdgcn_readfirstlane(net * a.tab_stride * 4) + opaque_zero();
          const rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc((void*)a.tab, (short)0, a.tab_bytes, 0x00020000);
          wave_sync();  // Rl visible to the dh3 products
          floatx4 dr = floatx4{0.f, 0.f, 0.f, 0.f};
          float ds = 0.f;
          for_groups<0, T3, kGroup>([&](auto g0c, auto szc) {
            constexpr int g0 = decltype(g0c)::value, SZ = decltype(szc)::value;
            float h3[4 * SZ], dp[4 * SZ];
#pragma unroll
            for (int k = 0; k < SZ; ++k) {
              floatx4 cf[6];
#pragma unroll
              for (int cc = 0; cc < 6; ++cc) cf[cc] = tab_load(trs, voff + 64 * (g0 + k), tb + cc * 64 * T3);
              floatx4 hv, dv;
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                float v = cf[0][i], d = 5.f * cf[0][i];
#pragma unroll
                for (int cc = 1; cc < 5; ++cc) {
                  v = __builtin_fmaf(v, x0m, cf[cc][i]);
                  d = __builtin_fmaf(d, x0m, (float)(5 - cc) * cf[cc][i]);
                }
                hv[i] = __builtin_fmaf(v, x0m, cf[5][i]);
                dv[i] = d;
              }
              *reinterpret_cast<floatx4*>(Tb + gm * kTS + 16 * k + 4 * gp) = hv;
              *reinterpret_cast<floatx4*>(Tb + (16 + gm) * kTS + 16 * k + 4 * gp) = dv;
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < SZ; ++k) {
              const floatx4 hv = *reinterpret_cast<const floatx4*>(Tb + n * kTS + 16 * k + 4 * q);
              const floatx4 dv = *reinterpret_cast<const floatx4*>(Tb + (16 + n) * kTS + 16 * k + 4 * q);
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                h3[4 * k + i] = hv[i];
                dp[4 * k + i] = dv[i];
              }
            }
            // dR^T[x][nbr] += sum_slot dT[x][slot] h3[slot][nbr]
#pragma unroll
            for (int k = 0; k < SZ; ++k)
#pragma unroll
              for (int i = 0; i < 4; ++i)
                dr = mfma4(n < 10 ? dTp[n * HS + 16 * (g0 + k) + 4 * q + i] : 0.f, h3[4 * k + i], dr);
            // dh3[slot][nbr] = sum_x dT[x][slot] R[nbr][x]
#pragma unroll
            for (int k = 0; k < SZ; ++k) {
              floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
              for (int kk = 0; kk < 3; ++kk)
                acc = mfma4(4 * kk + q < 10 ? dTp[(4 * kk + q) * HS + 16 * (g0 + k) + n] : 0.f, Rl[n * RS + 4 * kk + q], acc);
#pragma unroll
              for (int i = 0; i < 4; ++i) ds = __builtin_fmaf(acc[i], dp[4 * k + i], ds);
            }
          });
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (!ETAB || 4 * q + i < 10) dRl[n * RS + 4 * q + i] = dr[i];
          ds += __shfl_xor(ds, 16);
          ds += __shfl_xor(ds, 32);
          dst = ds;
        } else {
          float h1[4 * T1];
#pragma unroll
          for (int r = 0; r < 4 * T1; ++r) h1[r] = r < V1 ? tanh_s(st * k0(r) + b0(r)) : 0.f;
          // h2 = t2 + h1[r/Z2] is formed once; h3 = t3 + h2[r/Z3] is re-formed where used.  The
          // FP32 form re-forms t2 = h2 - h1[r/Z2] for tanh' (within an ulp of h2), which frees the
          // registers of a fourth wave per SIMD; the 16-bit forms are VALU bound and keep t2.
          constexpr bool KEEP_T2 = PREC != 0;
          float t2k[KEEP_T2 ? 4 * T2 : 1];
          float h2[4 * T2];
          auto H2 = [&](int r) -> float { return h2[r]; };
          floatx4 acc2[T2];
#pragma unroll
          for (int t = 0; t < T2; ++t) acc2[t] = floatx4{b1(4 * t), b1(4 * t + 1), b1(4 * t + 2), b1(4 * t + 3)};
          if constexpr (PREC == 0) {
            if constexpr (FOLD) {
              // full tiles and the thin last tile (4x4 blocks) in one chain
              mfma_chain<T2, 4 * T1, 0, V1, 1>(wrs, nb + 4 * O_F1A, vl4, [&](int s_) { return h1[s_]; }, acc2);
              thin_rows<D2>(acc2[T2 - 1], q);
            } else {
              mfma_chain<T2, 4 * T1, 0, V1>(wrs, nb + 4 * O_W1A, vl4, [&](int s_) { return h1[s_]; }, acc2);
            }
          } else {
            mfma_chain_h<PREC, T2, S21, 0, S21>(hrs, hb + H_W1A, vl16, [&](int s2) {
              return pack8<PREC, 4 * T1>([&](int r) { return h1[r]; }, s2); }, acc2);
          }
#pragma unroll
          for (int r = 0; r < 4 * T2; ++r) {
            const float t = r < V2 ? tanh_s(acc2[r / 4][r % 4]) : 0.f;
            if constexpr (KEEP_T2) t2k[r] = t;
            h2[r] = t + ((r / Z2 < 4 * T1) ? h1[r / Z2] : 0.f);
          }
          // folded K step: group 0's last-tile layer-2 registers spread over the lane groups
          const float x2 = !FOLD ? 0.f : (D2 == 2 ? fold2(h2[4 * (T2 - 1)], h2[4 * (T2 - 1) + 1]) : h2[4 * (T2 - 1)]);
          if constexpr (!H16) wave_sync();  // Rl visible to the dh3 products
          floatx4 dr = floatx4{0.f, 0.f, 0.f, 0.f};
          floatx4 dh2[T2];
#pragma unroll
          for (int t = 0; t < T2; ++t) dh2[t] = floatx4{0.f, 0.f, 0.f, 0.f};
          for_groups<0, T3, kGroup>([&](auto g0c, auto szc) {
            constexpr int g0 = decltype(g0c)::value, SZ = decltype(szc)::value;
            // fold form: this group holds the last layer-3 tile, whose four K steps fold to one
            constexpr bool LAST = FOLD && g0 + SZ == T3;
            floatx4 acc3[SZ];
#pragma unroll
            for (int t = 0; t < SZ; ++t) acc3[t] = floatx4{b2(4 * (g0 + t)), b2(4 * (g0 + t) + 1), b2(4 * (g0 + t) + 2), b2(4 * (g0 + t) + 3)};
            if constexpr (PREC == 0) {
              if constexpr (FOLD) {
                // the group's full tiles on the folded K steps; the last tile (last group) in 4x4 blocks
                if constexpr (!LAST || SZ > 1)
                  mfma_chain<LAST ? SZ - 1 : SZ, SF2, 0, SF2>(wrs, nb + 4 * (O_F2A + g0 * SF2 * 64), vl4,
                                                             [&](int s_) { return s_ < SF2 - 1 ? H2(s_) : x2; }, acc3);
                if constexpr (LAST) {
                  thin_chain<V2>(wrs, nb + 4 * O_S3A, vl4, [&](int s_) { return H2(s_); }, acc3[SZ - 1]);
                  thin_rows<D3>(acc3[SZ - 1], q);
                }
              } else
                mfma_chain<SZ, 4 * T2, 0, V2>(wrs, nb + 4 * (O_W2A + g0 * 4 * T2 * 64), vl4,
                                              [&](int s_) { return H2(s_); }, acc3);
            } else
              mfma_chain_h<PREC, SZ, S22, 0, S22>(hrs, hb + H_W2A + g0 * S22 * 1024, vl16, [&](int s2) {
                return pack8<PREC, 4 * T2>([&](int r) { return H2(r); }, s2); }, acc3);
            float t3[4 * SZ];
#pragma unroll
            for (int k = 0; k < SZ; ++k)
#pragma unroll
              for (int i = 0; i < 4; ++i) t3[4 * k + i] = tanh_s(acc3[k][i]);
            floatx4 dh3[SZ];
            if constexpr (H16) {
              // dR^T[x][nbr] += sum over the group's 32 slots in one 16x16x32 step: A = dT rows
              // (dTa, row n, zero row past 9), B = this lane's h3 of the group
              static_assert(g0 % 2 == 0, "tile groups start on a tile pair");
              HT16 bh;
#pragma unroll
              for (int j = 0; j < 8; ++j) {
                const int k = j / 4, r = 4 * (g0 + k) + (j & 3);
                const float h3v = k < SZ ? t3[4 * (k < SZ ? k : 0) + (j & 3)] + ((r / Z3 < 4 * T2) ? H2(r / Z3) : 0.f)
                                         : 0.f;
                bh[j] = (HE16)h3v;
              }
              const HT16 adt = *reinterpret_cast<const HT16*>(dTa + (((g0 / 2) * 4 + q) * 11 + (n < 10 ? n : 10)) * 8);
              dr = H8<2>::mfma(adt, bh, dr);
              // dh3[slot][nbr] = sum_x dT[x][slot] R[nbr][x]: K = x (10 of 32); lane groups 2,
              // 3 re-read the finite values of groups 0, 1 against their zero B operand
#pragma unroll
              for (int k = 0; k < SZ; ++k) {
                const HT16 at = *reinterpret_cast<const HT16*>(dTb + (16 * (g0 + k) + n) * 16 + 8 * (q & 1));
                dh3[k] = H8<2>::mfma(at, bR, floatx4{0.f, 0.f, 0.f, 0.f});
              }
            } else {
              // dR^T[x][nbr] += sum_slot dT[x][slot] h3[slot][nbr] over this group's slots
#pragma unroll
              for (int k = 0; k < SZ; ++k) {
                float h3v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                  const int r = 4 * (g0 + k) + i;
                  h3v[i] = t3[4 * k + i] + ((r / Z3 < 4 * T2) ? H2(r / Z3) : 0.f);
                }
                if (LAST && k == SZ - 1) {
                  const float x3 = D3 >= 3 ? fold4(h3v[0], h3v[1], h3v[2], h3v[3]) : D3 == 2 ? fold2(h3v[0], h3v[1]) : h3v[0];
                  dr = mfma4(dTf[16 * (T3 - 1)], x3, dr);
                } else {
#pragma unroll
                  for (int i = 0; i < 4; ++i) dr = mfma4(dTr[16 * (g0 + k) + i], h3v[i], dr);
                }
              }
              // dh3[slot][nbr] = sum_x dT[x][slot] R[nbr][x]
#pragma unroll
              for (int k = 0; k < SZ; ++k) {
                floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kk = 0; kk < 3; ++kk)
                  acc = mfma4(kk < 2 ? dTp[(4 * kk + q) * HS + 16 * (g0 + k) + n] : dTc2[16 * (g0 + k)], Rl[n * RS + 4 * kk + q], acc);
                dh3[k] = acc;
              }
            }
            // repeat-residual: dh2[r/Z3] += dh3[r]; then dz3 = dh3 (1 - t3^2)
#pragma unroll
            for (int k = 0; k < SZ; ++k)
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const int r = 4 * (g0 + k) + i;
                if (r / Z3 < 4 * T2) dh2[(r / Z3) / 4][(r / Z3) % 4] += dh3[k][i];
                dh3[k][i] *= (1.f - t3[4 * k + i] * t3[4 * k + i]);
              }
            if constexpr (PREC == 0) {
              if constexpr (FOLD) {
                // K steps of the group's full tiles, then (last group) the folded step
                constexpr int SZM = LAST ? SZ - 1 : SZ;
                float x3 = 0.f;
                if constexpr (LAST) {
                  const floatx4& d = dh3[SZ - 1];
                  x3 = D3 >= 3 ? fold4(d[0], d[1], d[2], d[3]) : D3 == 2 ? fold2(d[0], d[1]) : d[0];
                }
                // full tiles and the thin last layer-2 tile (lane group partials in 4x4 blocks)
                mfma_chain<T2, SF3, 4 * g0, 4 * SZM + (LAST ? 1 : 0), 1>(wrs, nb + 4 * O_F2B, vl4, [&](int s_) {
                  return s_ == SF3 - 1 ? x3 : dh3[(s_ - 4 * g0) / 4][(s_ - 4 * g0) % 4]; }, dh2);
              } else {
                mfma_chain<T2, 4 * T3, 4 * g0, 4 * SZ>(wrs, nb + 4 * O_W2B, vl4,
                                                      [&](int s_) { return dh3[(s_ - 4 * g0) / 4][(s_ - 4 * g0) % 4]; }, dh2);
              }
            } else
              mfma_chain_h<PREC, T2, S23, g0 / 2, 1>(hrs, hb + H_W2B, vl16, [&](int s2) {
                return pack8<PREC, 4 * SZ>([&](int r) { return dh3[r / 4][r % 4]; }, 0); }, dh2);
          });
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (!ETAB || 4 * q + i < 10) dRl[n * RS + 4 * q + i] = dr[i];
          // fold form: the last layer-2 tile's gradient (residual terms on lane group 0 included)
          if constexpr (FOLD) thin_rows<D2>(dh2[T2 - 1], q);
          // layer 2 backward
          floatx4 dh1[T1];
#pragma unroll
          for (int at = 0; at < T1; ++at) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int r1 = 4 * at + i;
              float v = 0.f;
              if (Z2 == 2) {
                if (2 * r1 < 4 * T2) v += dh2[(2 * r1) / 4][(2 * r1) % 4];
                if (2 * r1 + 1 < 4 * T2) v += dh2[(2 * r1 + 1) / 4][(2 * r1 + 1) % 4];
              } else {
                if (r1 < 4 * T2) v += dh2[r1 / 4][r1 % 4];
              }
              dh1[at][i] = v;
            }
          }
#pragma unroll
          for (int r = 0; r < 4 * T2; ++r) {
            float t;
            if constexpr (KEEP_T2) t = t2k[r];
            else t = h2[r] - ((r / Z2 < 4 * T1) ? h1[r / Z2] : 0.f);
            dh2[r / 4][r % 4] *= (1.f - t * t);
          }
          if constexpr (PREC == 0) {
            if constexpr (FOLD) {
              const float x = D2 == 2 ? fold2(dh2[T2 - 1][0], dh2[T2 - 1][1]) : dh2[T2 - 1][0];
              mfma_chain<T1, SF2, 0, SF2>(wrs, nb + 4 * O_F1B, vl4,
                                          [&](int s_) { return s_ < SF2 - 1 ? dh2[s_ / 4][s_ % 4] : x; }, dh1);
            } else {
              mfma_chain<T1, 4 * T2, 0, V2>(wrs, nb + 4 * O_W1B, vl4, [&](int s_) { return dh2[s_ / 4][s_ % 4]; }, dh1);
            }
          } else
            mfma_chain_h<PREC, T1, S22, 0, S22>(hrs, hb + H_W1B, vl16, [&](int s2) {
              return pack8<PREC, 4 * T2>([&](int r) { return dh2[r / 4][r % 4]; }, s2); }, dh1);
          dst = 0.f;
#pragma unroll
          for (int r1 = 0; r1 < 4 * T1; ++r1) dst += dh1[r1 / 4][r1 % 4] * (1.f - h1[r1] * h1[r1]) * k0(r1);
          dst *= 1.f / kTanhScale;  // k0 is stored scaled
          dst += __shfl_xor(dst, 16);
          dst += __shfl_xor(dst, 32);
        }
        if constexpr (ETAB) {
          if (q == 0) dRl[n * RS + kESc] = dst;
        } else {
          wave_sync();
          float dR[10];
#pragma unroll
          for (int x = 0; x < 10; ++x) dR[x] = dRl[n * RS + x];
          float gx[3];
          env_backward(g, dR, dst, a.inv_std[ti], gx);
          pair_out(a, prow, valid && q == 0, e, g, gx, fc, vir);
          wave_sync();
        }
      }
      if constexpr (ETAB) {
        // pair gradients of the super-chunk, one neighbour per lane: the geometry re-formed
        // from the pair vector, dE/dR and dE/ds~ from the table
        wave_sync();
        if (lane < SN) {
          const bool vl = c0 + lane < m;
          const float4 e = vl ? list[c0 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
          Geo g;
          float R[10];
          env_forward(e, vl, a.inv_rcut, a.mean[ti], a.inv_std[ti], g, R);
          float dR[10];
#pragma unroll
          for (int x = 0; x < 10; ++x) dR[x] = Et[lane * RS + x];
          float gx[3];
          env_backward(g, dR, Et[lane * RS + kESc], a.inv_std[ti], gx);
          pair_out(a, prow, vl, e, g, gx, fc, vir);
        }
        wave_sync();  // the table is rewritten by the next super-chunk
      }
      }
    }
    // centre force: reduce over the wave (the lanes that summed pair gradients: group 0, or with
    // ETAB the first SN lanes; the others hold zeros)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = fc[k];
#pragma unroll
      for (int sft = 1; sft < 64; sft <<= 1) v += __shfl_xor(v, sft);
      fc[k] = v;
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (a.pgrad) a.fcen[(size_t)gc * 3 + k] = fc[k];
        else atomicAdd(a.force + (size_t)gc * 3 + k, fc[k]);
      }
    }
  }
  // virial: wave reduce (see the centre force), block reduce in fp64, one partial per block
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    float v = vir[k];
#pragma unroll
    for (int sft = 1; sft < 64; sft <<= 1) v += __shfl_xor(v, sft);
    vir[k] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) vir_s[wv][k] = (double)vir[k];
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = vir_s[0][threadIdx.x] + vir_s[1][threadIdx.x] + vir_s[2][threadIdx.x] + vir_s[3][threadIdx.x];
    a.partial[(size_t)blockIdx.x * 9 + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------------
// Launchers

static void fill_args(System* s, EmbedArgs& a) {
  Model* M = s->model;
  memset(&a, 0, sizeof(a));
  a.pos = s->d_pos32;
  a.seg = s->d_seg;
  a.blk_seg = s->d_blk_seg;
  for (int j = 0; j < s->ntypes; ++j) {
    a.nbr[j] = s->d_nbr[j];
    a.nbr_cnt[j] = s->d_nbr_cnt[j];
    a.cap[j] = s->cap[j];
    a.rowbase[j] = s->rowbase[j];
    a.mean[j] = (float)M->mean[j];
    a.inv_std[j] = (float)M->istd[j];
  }
  a.ntypes = s->ntypes;
  a.rcut = (float)M->rcut;
  a.inv_rcut = (float)(1.0 / M->rcut);
  a.net = M->d_emb;
  a.net_stride = M->off.total;
  a.net_bytes = M->off.total * 4 * M->ntypes * M->ntypes;
  for (int q = 0; q < 4; ++q) {
    a.base3[q] = M->lay.base[2][q];
    a.size3[q] = M->lay.size[2][q];
  }
  a.W = M->wout;
  a.A = M->axis;
  a.inv_nnbrs = (float)(1.0 / M->nnbrs);
  a.tbias = M->d_tbias;
  a.T = s->d_T;
  a.G = (float*)s->d_G;
  a.dG = s->d_dG;
  a.dg16 = s->dg16;
  a.force = s->d_force;
  a.partial = s->partial_out ? s->partial_out : s->d_partial;
  a.wc_g = s->wc_g;
  a.wc_y3 = s->wc_y3;
  a.wc_ldy = s->wc_ldy;
  a.wc_norm = s->wc_norm;
  a.centre_mask = M->atomic ? M->nsel_mask : ~0;
  if (s->deterministic && !s->wc_model && !M->atomic && s->d_pgrad) {
    a.pgrad = s->d_pgrad;
    a.fcen = s->d_fcen;
    int k = 0;
    for (int j = 0; j < s->ntypes; ++j) {
      a.koff[j] = k;
      k += s->cap[j];
    }
    a.ksum = k;
  }
  a.paircnt = reinterpret_cast<int*>(s->d_pairs);
  int cmax = 1;
  for (int j = 0; j < s->ntypes; ++j) cmax = std::max(cmax, s->cap[j]);
  a.list_cap = (cmax + 3) / 4 * 4;
  if (M->tab_n > 0) {
    a.tab = M->d_tab;
    a.tab_n = M->tab_n;
    a.tab_stride = M->tab_stride;
    a.tab_bytes = (int)((size_t)M->tab_stride * 4 * M->ntypes * M->ntypes);
    for (int k = 0; k < M->ntypes * M->ntypes; ++k) {
      a.tab_srmin[k] = M->tab_srmin[k];
      a.tab_delta[k] = M->tab_delta[k];
    }
  }
}

template <int T1, int T2, int T3, int Z2, int Z3, int PREC, int V1, bool TAB, bool FOLD = false>
static int32_t launch_embed(System* s, bool backward, int precision, int g16) {
  EmbedArgs a;
  fill_args(s, a);
  a.g16 = backward ? 0 : (g16 >= 0 ? g16 : precision);
  if (PREC) {
    Model* M = s->model;
    a.hnet = M->d_emb_h[PREC - 1];
    a.hnet_stride = M->offh.total * 2;
    a.hnet_bytes = M->offh.total * 2 * M->ntypes * M->ntypes;
    if (!a.hnet) {
      set_error("16-bit embedding tables missing");
      return DPMD_E_STATE;
    }
  }
  constexpr int HS = 16 * T3 + 1;
  const int W = a.W, A = a.A;
  if (FOLD && W != fold_width(V1, Z2, Z3)) {
    set_error("fold-form embedding kernel launched for output width %d", W);
    return DPMD_E_STATE;
  }
  const int nblocks = s->nblocks;
  size_t wave_floats;
  if (!backward) {
    wave_floats = 4 * a.list_cap + fwd_env_floats(PREC, TAB) + 16 * (16 * kGroup + 1);
    wave_floats = std::max(wave_floats, (size_t)(10 * W + 12 * 16));
  } else {
    // dE/dT, then the chunk tiles (R / dR, pair list, tabulated h3 and p') or, in the prologue,
    // T in whole 64-float passes (and with A > 16 the X and U tables; with A <= 16 they sit in
    // dE/dT's space)
    constexpr int TPASS = 64 * ((160 * T3 + 63) / 64);
    if (10 * W > TPASS) {
      set_error("embedding output width %d exceeds the kernel's tiles", W);
      return DPMD_E_UNSUPPORTED;
    }
    wave_floats = (size_t)embed_dt_floats(T3) +
                  std::max((size_t)(bwd_env_floats(PREC, TAB) + 4 * a.list_cap + (TAB ? 2 * 16 * kTS : 0)),
                           (size_t)(A <= 16 ? TPASS : std::max(TPASS, 10 * W + 20 * A)));
    // 16-bit MLP forms: the FP16 dE/dT copies replace dTp in the chunk phase
    if (PREC != 0 && !TAB)
      wave_floats = std::max(wave_floats, (size_t)embed_dt16_floats(T3) + bwd_env_floats(PREC, TAB) + 4 * a.list_cap);
  }
  wave_floats = (wave_floats + 3) / 4 * 4;
  a.lds_wave_floats = (int)wave_floats;
  const size_t small = (size_t)(s->ntypes * s->ntypes * 16 * (2 * T1 + T2 + T3) + 3) / 4 * 4;
  const size_t lds = (4 * wave_floats + small) * sizeof(float);
  if (lds > 160 * 1024) {
    set_error("embedding kernel needs %zu bytes of LDS (neighbour capacity too large)", lds);
    return DPMD_E_UNSUPPORTED;
  }
  if (!backward) {
    hipLaunchKernelGGL((embed_fwd_kernel<T1, T2, T3, Z2, Z3, PREC, V1, TAB, FOLD>), dim3(nblocks), dim3(256), lds, s->stream, a);
  } else {
    hipLaunchKernelGGL((embed_bwd_kernel<T1, T2, T3, Z2, Z3, PREC, V1, TAB, FOLD>), dim3(nblocks), dim3(256), lds, s->stream, a);
  }
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// A tabulated (compressed) model has no embedding GEMMs: its kernel is FP32 whatever the
// precision of the fitting GEMMs.
// DPMD_EMBED_FOLD=0 runs the unfolded FP32 kernels on a folding layout (comparison runs)
static bool fold_disabled() {
  static const bool off = [] {
    const char* e = getenv("DPMD_EMBED_FOLD");
    return e && e[0] == '0';
  }();
  return off;
}

// FOLDS: the instantiation has a fold-form variant (its layouts that fold take it)
template <int T1, int T2, int T3, int Z2, int Z3, int V1, bool FOLDS = false>
static int32_t launch_prec(System* s, bool backward, int precision, int g16) {
  if (s->model->tab_n > 0) return launch_embed<T1, T2, T3, Z2, Z3, 0, V1, true>(s, backward, precision, g16);
  if (precision == DPMD_PREC_BF16) return launch_embed<T1, T2, T3, Z2, Z3, 1, V1, false>(s, backward, precision, g16);
  if (precision == DPMD_PREC_FP16) return launch_embed<T1, T2, T3, Z2, Z3, 2, V1, false>(s, backward, precision, g16);
  if constexpr (FOLDS) {
    if (s->model->lay.fold && !fold_disabled()) return launch_embed<T1, T2, T3, Z2, Z3, 0, V1, false, true>(s, backward, precision, g16);
  }
  return launch_embed<T1, T2, T3, Z2, Z3, 0, V1, false>(s, backward, precision, g16);
}

int32_t embed_launch(System* s, bool backward, int precision, int g16) {
  const EmbedLayout& L = s->model->lay;
  if (L.tiles[0] == 2 && L.tiles[1] == 4 && L.tiles[2] == 7 && L.z[1] == 2 && L.z[2] == 2)
    return launch_prec<2, 4, 7, 2, 2, 7, true>(s, backward, precision, g16);  // 7 layer-3 tiles imply v1 <= 7
  if (L.tiles[0] == 2 && L.tiles[1] == 2 && L.tiles[2] == 4 && L.z[1] == 1 && L.z[2] == 2)
    return launch_prec<2, 2, 4, 1, 2, 8>(s, backward, precision, g16);
  set_error("embedding widths (%d,%d,%d) have no kernel instantiation", L.w[0], L.w[1], L.w[2]);
  return DPMD_E_UNSUPPORTED;
}

// Deterministic forces: F_g = (centre part of g) - sum over g's row entries (neighbour i,
// slot rev in i's row of g's type) of the pair gradient dE/dx_ig that i's backward stored,
// summed in row order and combined across lanes by a fixed butterfly, so two evaluations of
// the same positions give bit-identical forces.  One wave per atom.
__global__ void __launch_bounds__(256) force_gather_kernel(const SegInfo* seg, const int* atom_seg, int NA,
                                                           int ntypes, NbrRef nr, const float* pgrad,
                                                           const float* fcen, int ksum, float* force) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= NA) return;
  const SegInfo& S = seg[atom_seg[g]];
  const int i0 = g - S.base;
  int tg = 0;
  while (tg + 1 < ntypes && i0 >= S.off[tg + 1]) ++tg;
  int koff = 0;
  for (int t = 0; t < tg; ++t) koff += nr.cap[t];
  float acc[3] = {0.f, 0.f, 0.f};
  for (int t = 0; t < ntypes; ++t) {
    const int capt = nr.cap[t];
    const int cnt = nr.cnt[t][g];
    const int* row = nr.nbr[t] + (size_t)g * capt;
    const int* rev = nr.rev[t] + (size_t)g * capt;
    for (int p = lane; p < cnt; p += 64) {
      const int e = rev[p];
      if (e < 0) continue;
      const size_t gi = (size_t)S.base + S.off[t] + row[p];
      const float* pg = pgrad + (gi * ksum + koff + e) * 3;
      acc[0] += pg[0];
      acc[1] += pg[1];
      acc[2] += pg[2];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float v = acc[k];
#pragma unroll
    for (int sft = 1; sft < 64; sft <<= 1) v += __shfl_xor(v, sft);
    acc[k] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) force[(size_t)g * 3 + k] = fcen[(size_t)g * 3 + k] - acc[k];
  }
}

int32_t force_gather(System* s) {
  NbrRef nr;
  for (int t = 0; t < s->ntypes; ++t) {
    nr.nbr[t] = s->d_nbr[t];
    nr.cnt[t] = s->d_nbr_cnt[t];
    nr.rev[t] = s->d_rev[t];
    nr.cap[t] = s->cap[t];
  }
  int ksum = 0;
  for (int t = 0; t < s->ntypes; ++t) ksum += s->cap[t];
  hipLaunchKernelGGL(force_gather_kernel, dim3((s->NA + 3) / 4), dim3(256), 0, s->stream, s->d_seg, s->d_atom_seg,
                     s->NA, s->ntypes, nr, s->d_pgrad, s->d_fcen, ksum, s->d_force);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// 16-bit operand tables derived from the packed fp32 tables: [ot][s2][lane][8] with
// element j of step s2 taken from fp32 register step 8*s2 + j (zero past the end).
EmbedNetOffsetsH embed_net_offsets_h(const EmbedLayout& L) {
  const int t1 = L.tiles[0], t2 = L.tiles[1], t3 = L.tiles[2];
  const int s21 = (4 * t1 + 7) / 8, s22 = (4 * t2 + 7) / 8, s23 = (4 * t3 + 7) / 8;
  EmbedNetOffsetsH o;
  o.w1a = 0;
  o.w2a = o.w1a + t2 * s21 * 512;
  o.w2b = o.w2a + t3 * s22 * 512;
  o.w1b = o.w2b + t2 * s23 * 512;
  o.total = o.w1b + t1 * s22 * 512;
  return o;
}

// Tabulated net coefficients (ngrids, 6, W) -> [interval][power][16*tiles3] in the
// layer-3 physical slot order (padding slots zero).
void embed_pack_table(const EmbedLayout& L, int ngrids, const float* coeff, float* out) {
  const int S = 16 * L.tiles[2], W = L.w[2];
  for (int g = 0; g < ngrids; ++g)
    for (int c = 0; c < 6; ++c)
      for (int slot = 0; slot < S; ++slot) {
        const int w = L.logical(2, slot);
        out[((size_t)g * 6 + c) * S + slot] = w >= 0 ? coeff[((size_t)g * 6 + c) * W + w] : 0.f;
      }
}

static uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static uint16_t to_fp16(float f) {
  _Float16 h = (_Float16)f;
  uint16_t u;
  memcpy(&u, &h, 2);
  return u;
}

void embed_pack_net_h(const EmbedLayout& L, const float* f32, uint16_t* out, bool fp16) {
  const EmbedNetOffsets o = embed_net_offsets(L);
  const EmbedNetOffsetsH oh = embed_net_offsets_h(L);
  const int t1 = L.tiles[0], t2 = L.tiles[1], t3 = L.tiles[2];
  struct Tab { int src, dst, ot, sregs; } tabs[4] = {
      {o.w1a, oh.w1a, t2, 4 * t1}, {o.w2a, oh.w2a, t3, 4 * t2}, {o.w2b, oh.w2b, t2, 4 * t3}, {o.w1b, oh.w1b, t1, 4 * t2}};
  memset(out, 0, sizeof(uint16_t) * oh.total);
  for (const Tab& t : tabs) {
    const int s2n = (t.sregs + 7) / 8;
    for (int ot = 0; ot < t.ot; ++ot)
      for (int s2 = 0; s2 < s2n; ++s2)
        for (int lane = 0; lane < 64; ++lane)
          for (int j = 0; j < 8; ++j) {
            const int r = 8 * s2 + j;
            const float v = r < t.sregs ? f32[t.src + (ot * t.sregs + r) * 64 + lane] : 0.f;
            out[t.dst + ((ot * s2n + s2) * 64 + lane) * 8 + j] = fp16 ? to_fp16(v) : to_bf16(v);
          }
  }
}

}  // namespace dpmd

// Integrator and reduction kernels.  Positions and momenta are FP64 ((NA, 3), segments
// concatenated, type-sorted inside each segment); forces arrive in FP32 from the model
// kernels.  Per-segment quantities (energy, virial, kinetic energy, thermostat state) are
// reported per periodic image: a segment that holds nimg images of a small box divides
// its sums by nimg.
//
// Restated integrator steps:
//   velocity Verlet (jax_md simulate.nve, wrapped at deepmd_jax/routine.py:139-144):
//     p += dt/2 F; x += dt p/m; F = F(x); p += dt/2 F
//   BAOAB O step (deepmd_jax/routine.py:86-96): p = c1 p + c2 sqrt(m) xi,
//     c1 = exp(-gamma dt), c2 = sqrt(kT (1-c1^2)); thermostat_work += dKE
//   PIMD free ring + PILE-L (deepmd_jax/pimd.py:92-111,160-179,210-227)
// Positions are kept unwrapped, which is what the reference's wrap-then-
// min_image_unwrap (deepmd_jax/md_utils.py:162-169) produces.
#include "internal.h"
#include <algorithm>
#include <string.h>

namespace dpmd {

__device__ __forceinline__ float wrap32(double x, double L) {
  double r = fmod(x, L);
  if (r < 0) r += L;
  if (r >= L) r -= L;
  float v = (float)r;
  if (v >= (float)L) v = 0.f;
  return v;
}

// FP32 mirror of a position wrapped into the segment's cell: per axis in an orthorhombic
// box, through fractional coordinates in [0, 1) in a general cell
__device__ __forceinline__ float4 wrap_pos32(const SegInfo& S, const double x[3]) {
  if (S.tri) {
    double f[3], w[3];
    cell_mul(S.Hid, x, f);
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] -= floor(f[d]);
    cell_mul(S.Hd, f, w);
    return make_float4((float)w[0], (float)w[1], (float)w[2], 0.f);
  }
  return make_float4(wrap32(x[0], S.boxd[0]), wrap32(x[1], S.boxd[1]), wrap32(x[2], S.boxd[2]), 0.f);
}

__global__ void write_pos32_kernel(const double* pos, float4* p32, const SegInfo* seg, const int* atom_seg,
                                   int total) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const SegInfo& S = seg[atom_seg[g]];
  const double x[3] = {pos[(size_t)g * 3], pos[(size_t)g * 3 + 1], pos[(size_t)g * 3 + 2]};
  p32[g] = wrap_pos32(S, x);
}

__global__ void zero_kernel(float* x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = 0.f;
}

__global__ void kick_kernel(double* mom, const float* force, size_t n3, double h) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n3) mom[i] += h * (double)force[i];
}

// x += dt p / m, fp32 mirror, NaN/Inf test -> error bit 16
__global__ void drift_kernel(double* pos, const double* mom, const double* mass, float4* p32, const SegInfo* seg,
                             const int* atom_seg, int total, double dt, int* flags, const float* mobile) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const SegInfo& S = seg[atom_seg[g]];
  const double m = mass[g];
  if (mobile && mobile[g] == 0.f) dt = 0.0;  // fixed atoms keep their positions
  double v[3];
  bool bad = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double p = mom[(size_t)g * 3 + d];
    const double x = pos[(size_t)g * 3 + d] + dt * p / m;
    pos[(size_t)g * 3 + d] = x;
    bad |= !isfinite(x) || !isfinite(p);
    v[d] = x;
  }
  p32[g] = wrap_pos32(S, v);
  if (bad) atomicOr(&flags[2], 16);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// per-segment energy (sum of per-row energies) and virial (sum of the segment's block
// partials), FP64, per periodic image: out[16 sid + 0] energy, out[16 sid + 1..9] virial
__global__ void reduce_kernel(const float* eatom, int nt, RowBase rb, const SegInfo* seg, const double* partial,
                              double* out) {
  __shared__ double red[256];
  const SegInfo& S = seg[blockIdx.x];
  double s = 0.0;
  for (int t = 0; t < nt; ++t) {
    const int base = rb.r[t] + S.row[t];
    const int cnt = S.off[t + 1] - S.off[t];
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) s += (double)eatom[base + i];
  }
  const double inv = 1.0 / S.nimg;
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x * 16] = s * inv;
  for (int c = 0; c < 9; ++c) {
    double v = 0.0;
    for (int b = threadIdx.x; b < S.nblk; b += blockDim.x) v += partial[(size_t)(S.blk0 + b) * 9 + c];
    v = block_sum(v, red);
    if (threadIdx.x == 0) out[blockIdx.x * 16 + 1 + c] = v * inv;
  }
}

// Kinetic energy per segment (per image) in two passes: block (kb, segment) sums
// p_a^2 / m per axis over its strided share of the segment's atoms into part, then one
// thread per segment adds the KB partials in a fixed order (deterministic).  Optionally
// the per-axis sums (2 KE per axis) of segment 0 go to axis_out.
__global__ void ke_part_kernel(const double* mom, const double* mass, const SegInfo* seg, int KB, double* part) {
  __shared__ double red[256];
  const SegInfo& S = seg[blockIdx.y];
  double k[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.natoms; i += KB * blockDim.x) {
    const size_t g = (size_t)S.base + i;
    const double im = 1.0 / mass[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) k[d] += mom[g * 3 + d] * mom[g * 3 + d] * im;
  }
  for (int d = 0; d < 3; ++d) {
    const double v = block_sum(k[d], red);
    if (threadIdx.x == 0) part[((size_t)blockIdx.y * KB + blockIdx.x) * 3 + d] = v;
  }
}

__global__ void ke_final_kernel(const double* part, const SegInfo* seg, int nseg, int KB, double* out,
                                double* axis_out) {
  const int sid = blockIdx.x * blockDim.x + threadIdx.x;
  if (sid >= nseg) return;
  double k[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < KB; ++b)
    for (int d = 0; d < 3; ++d) k[d] += part[((size_t)sid * KB + b) * 3 + d];
  if (out) out[sid] = 0.5 * (k[0] + k[1] + k[2]) / seg[sid].nimg;
  if (axis_out && sid == 0)
    for (int d = 0; d < 3; ++d) axis_out[d] = k[d];
}

// per-segment sum of a per-atom scratch array, added to thermo[64 sid + slot] (per image)
__global__ void seg_accumulate_kernel(const double* x, const SegInfo* seg, double* thermo, int slot) {
  __shared__ double red[256];
  const SegInfo& S = seg[blockIdx.x];
  double s = 0.0;
  for (int i = threadIdx.x; i < S.natoms; i += blockDim.x) s += x[(size_t)S.base + i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) thermo[blockIdx.x * 64 + slot] += s / S.nimg;
}

// --- counter-based normal variates: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller
__device__ __forceinline__ void philox(uint32_t ctr[4], uint32_t key[2]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * ctr[0];
    const uint64_t p1 = (uint64_t)M1 * ctr[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ ctr[1] ^ key[0];
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ ctr[3] ^ key[1];
    const uint32_t n3 = (uint32_t)p0;
    ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
    key[0] += W0; key[1] += W1;
  }
}

__device__ __forceinline__ void normal4(uint64_t seed, uint64_t counter, uint32_t idx, uint32_t stream, double z[4]) {
  uint32_t c[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), idx, stream};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox(c, k);
  const double two32 = 4294967296.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const double u1 = ((double)c[2 * h] + 1.0) / (two32 + 1.0);
    const double u2 = (double)c[2 * h + 1] / two32;
    const double r = sqrt(-2.0 * log(u1));
    z[2 * h] = r * cos(6.283185307179586 * u2);
    z[2 * h + 1] = r * sin(6.283185307179586 * u2);
  }
}

// Index of an atom's primary copy inside its segment.  A segment holding nimg images of a
// small box lists its atoms image-major before the stable type sort, so inside type block
// j the copy of primary atom k of image m sits at off[j] + m n_j + k.
__device__ __forceinline__ int primary_index(const SegInfo& S, int t) {
  if (S.nimg == 1) return t;
  int j = 0;
  while (j + 1 < kMaxT && t >= S.off[j + 1]) ++j;
  const int nj = (S.off[j + 1] - S.off[j]) / S.nimg;
  return S.off[j] / S.nimg + (t - S.off[j]) % nj;
}

__device__ __forceinline__ uint64_t seg_key(uint64_t seed, uint64_t seg_seed) {
  return seed ^ (seg_seed * 0x9E3779B97F4A7C15ull);
}

// BAOAB O step.  The normals of an atom depend only on (seed, segment seed, segment step,
// primary index), so a segment draws the same stream alone or batched with others, and
// all images of a small box receive the same kicks.  Per-atom kinetic-energy change ->
// scratch for the per-segment thermostat work.
__global__ void langevin_o_kernel(double* mom, const double* mass, const SegInfo* seg, const int* atom_seg,
                                  int total, double c1, double c2, uint64_t seed, uint64_t step, double* dke_out,
                                  const float* mobile) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const SegInfo& S = seg[atom_seg[g]];
  const double m = mass[g];
  double z[4];
  normal4(seg_key(seed, S.seed), S.step + step, (uint32_t)primary_index(S, g - S.base), 0x4c414e47u, z);
  if (mobile && mobile[g] == 0.f) z[0] = z[1] = z[2] = 0.0;  // no noise on fixed atoms (routine.py:93-94)
  double dke = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double p0 = mom[(size_t)g * 3 + d];
    const double p1 = c1 * p0 + c2 * sqrt(m) * z[d];
    mom[(size_t)g * 3 + d] = p1;
    dke += 0.5 * (p1 * p1 - p0 * p0) / m;
  }
  dke_out[g] = dke;
}

// Beads of a general cell onto bead 0's image through the fractional minimum image
// (deepmd_jax/pimd.py:210-227, general branch); one (bead >= 1, atom) per thread.
__global__ void pimd_unwrap_cell_kernel(double* pos, int natoms, int P, const SegInfo* seg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms * (P - 1)) return;
  const int b = 1 + t / natoms, i = t % natoms;
  const SegInfo& S = seg[0];
  double* x = pos + ((size_t)b * natoms + i) * 3;
  const double* x0 = pos + (size_t)i * 3;
  double d[3] = {x[0] - x0[0], x[1] - x0[1], x[2] - x0[2]}, f[3];
  cell_mul(S.Hid, d, f);
#pragma unroll
  for (int k = 0; k < 3; ++k) f[k] -= rint(f[k]);
  cell_mul(S.Hd, f, d);
#pragma unroll
  for (int k = 0; k < 3; ++k) x[k] = x0[k] + d[k];
}

// PIMD exact free ring-polymer step for one (atom, dim) per thread: unwrap beads onto
// bead 0's image, rotate to normal modes, propagate, rotate back.  A position or momentum
// that is not finite sets error bit 16 (deepmd_jax/md.py:637-640).
__global__ void pimd_free_ring_kernel(double* pos, double* mom, const double* mass, int natoms, int P,
                                      const double* C, const double* freqs, double h, double Lx, double Ly,
                                      double Lz, int* flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms * 3) return;
  const int i = t / 3, d = t % 3;
  const double L = d == 0 ? Lx : (d == 1 ? Ly : Lz);
  const double m = mass[i];
  double q[64], p[64], qn[64], pn[64];
  const double x0 = pos[(size_t)i * 3 + d];
  for (int b = 0; b < P; ++b) {
    const double x = pos[((size_t)b * natoms + i) * 3 + d];
    double dl = x - x0;
    if (L > 0) dl -= L * rint(dl / L);  // L = 0: beads already unwrapped (general cell)
    q[b] = x0 + dl;
    p[b] = mom[((size_t)b * natoms + i) * 3 + d];
  }
  for (int k = 0; k < P; ++k) {
    double sq = 0, sp = 0;
    for (int b = 0; b < P; ++b) {
      sq += C[b * P + k] * q[b];
      sp += C[b * P + k] * p[b];
    }
    qn[k] = sq;
    pn[k] = sp;
  }
  qn[0] += h * pn[0] / m;
  for (int k = 1; k < P; ++k) {
    const double w = freqs[k];
    const double cs = cos(w * h), sn = sin(w * h);
    const double qq = qn[k] * cs + pn[k] * sn / (m * w);
    const double pp = pn[k] * cs - m * w * qn[k] * sn;
    qn[k] = qq;
    pn[k] = pp;
  }
  bool bad = false;
  for (int b = 0; b < P; ++b) {
    double sq = 0, sp = 0;
    for (int k = 0; k < P; ++k) {
      sq += C[b * P + k] * qn[k];
      sp += C[b * P + k] * pn[k];
    }
    pos[((size_t)b * natoms + i) * 3 + d] = sq;
    mom[((size_t)b * natoms + i) * 3 + d] = sp;
    bad |= !isfinite(sq) || !isfinite(sp);
  }
  if (bad) atomicOr(&flags[2], 16);
}

// PILE-L O step in normal modes; work[0] accumulates the thermostat work
__global__ void pimd_o_kernel(double* mom, const double* mass, int natoms, int P, const double* C,
                              const double* c1, const double* c2, uint64_t seed, uint64_t counter, double* work) {
  __shared__ double red[256];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double dke = 0.0;
  if (t < natoms * 3) {
    const int i = t / 3, d = t % 3;
    const double m = mass[i];
    double p[64], pn[64];
    for (int b = 0; b < P; ++b) p[b] = mom[((size_t)b * natoms + i) * 3 + d];
    for (int k = 0; k < P; ++k) {
      double s = 0;
      for (int b = 0; b < P; ++b) s += C[b * P + k] * p[b];
      pn[k] = s;
    }
    for (int k = 0; k < P; k += 4) {
      double z[4];
      normal4(seed, counter, (uint32_t)t, 0x50494d00u + (uint32_t)k, z);
      for (int u = 0; u < 4 && k + u < P; ++u) pn[k + u] = c1[k + u] * pn[k + u] + c2[k + u] * sqrt(m) * z[u];
    }
    for (int b = 0; b < P; ++b) {
      double s = 0;
      for (int k = 0; k < P; ++k) s += C[b * P + k] * pn[k];
      dke += 0.5 * (s * s - p[b] * p[b]) / m;
      mom[((size_t)b * natoms + i) * 3 + d] = s;
    }
  }
  red[threadIdx.x] = dke;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&work[0], red[0]);
}

// ---------------------------------------------------------------------------------

int32_t md_write_pos32(System* s) {
  DPMD_TRY(sys_upload_segments(s));
  const int NA = s->NA;
  hipLaunchKernelGGL(write_pos32_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, s->d_pos, s->d_pos32,
                     s->d_seg, s->d_atom_seg, NA);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// One recorded trajectory frame in the caller's atom order: positions and velocities
// (momentum / mass), as Simulation.run returns them (deepmd_jax/md.py:806-964).
__global__ void traj_record_kernel(const double* pos, const double* mom, const double* mass, const int* perm,
                                   double* tpos, double* tvel, int n3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const int g = i / 3, d = i - 3 * g;
  const size_t o = (size_t)perm[g] * 3 + d;
  tpos[o] = pos[i];
  tvel[o] = mom[i] / mass[g];
}

int32_t md_traj_record(System* s, double* tpos, double* tvel) {
  const int n3 = s->NA * 3;
  hipLaunchKernelGGL(traj_record_kernel, dim3((n3 + 255) / 256), dim3(256), 0, s->stream, s->d_pos, s->d_mom,
                     s->d_mass, s->d_perm, tpos, tvel, n3);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_zero_force(System* s) {
  const size_t n = (size_t)s->NA * 3;
  hipLaunchKernelGGL(zero_kernel, dim3(std::min<size_t>(2048, (n + 255) / 256)), dim3(256), 0, s->stream,
                     s->d_force, n);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_kick(System* s, double h) {
  const size_t n3 = (size_t)s->NA * 3;
  hipLaunchKernelGGL(kick_kernel, dim3((n3 + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_force, n3, h);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_drift(System* s, double dt) {
  const int NA = s->NA;
  hipLaunchKernelGGL(drift_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, s->d_pos, s->d_mom, s->d_mass,
                     s->d_pos32, s->d_seg, s->d_atom_seg, NA, dt, s->d_flags, s->has_fixed ? s->d_mobile : nullptr);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_reduce_launch(System* s) {
  RowBase rb;
  for (int t = 0; t < kMaxT; ++t) rb.r[t] = s->rowbase[t];
  hipLaunchKernelGGL(reduce_kernel, dim3(s->nseg), dim3(256), 0, s->stream, s->d_eatom, s->ntypes, rb, s->d_seg,
                     s->d_partial, s->d_red);
  DPMD_HIP(hipGetLastError());
  // DPLR: the long-range energy of the last evaluation joins the segment energy
  if (s->wc_model) DPMD_TRY(dplr_add_energy(s));
  return DPMD_OK;
}

// blocks per segment for the kinetic-energy reductions: about 2,048 atoms per block
static int ke_blocks(System* s) {
  int mx = 1;
  for (const SegInfo& S : s->seg) mx = std::max(mx, S.natoms);
  return std::max(1, std::min(kKeBlocks, (mx + 2047) / 2048));
}

static int32_t ke_launch(System* s, double* d_out, double* d_axis) {
  const int KB = ke_blocks(s);
  hipLaunchKernelGGL(ke_part_kernel, dim3(KB, s->nseg), dim3(256), 0, s->stream, s->d_mom, s->d_mass, s->d_seg, KB,
                     s->d_kepart);
  hipLaunchKernelGGL(ke_final_kernel, dim3((s->nseg + 63) / 64), dim3(64), 0, s->stream, s->d_kepart, s->d_seg,
                     s->nseg, KB, d_out, d_axis);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_ke_launch(System* s, double* d_out) { return ke_launch(s, d_out, nullptr); }

// BAOAB O step for every segment; thermostat work of segment r is thermo[64 r]
int32_t md_langevin_o(System* s, double dt, double gamma, double kT, uint64_t seed, uint64_t step) {
  const double c1 = exp(-gamma * dt);
  const double c2 = sqrt(std::max(kT * (1.0 - c1 * c1), 0.0));
  const int NA = s->NA;
  hipLaunchKernelGGL(langevin_o_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_mass,
                     s->d_seg, s->d_atom_seg, NA, c1, c2, seed, step, s->d_scratch,
                     s->has_fixed ? s->d_mobile : nullptr);
  hipLaunchKernelGGL(seg_accumulate_kernel, dim3(s->nseg), dim3(256), 0, s->stream, s->d_scratch, s->d_seg,
                     s->d_thermo, 0);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pimd_free_ring(System* s, const double* d_C, const double* d_freqs, double h) {
  const SegInfo& S = s->seg[0];
  const int n = S.natoms * 3;
  if (S.tri) {
    const int nu = S.natoms * (s->nseg - 1);
    hipLaunchKernelGGL(pimd_unwrap_cell_kernel, dim3((nu + 255) / 256), dim3(256), 0, s->stream, s->d_pos, S.natoms,
                       s->nseg, s->d_seg);
  }
  const double Lx = S.tri ? 0.0 : S.boxd[0], Ly = S.tri ? 0.0 : S.boxd[1], Lz = S.tri ? 0.0 : S.boxd[2];
  hipLaunchKernelGGL(pimd_free_ring_kernel, dim3((n + 127) / 128), dim3(128), 0, s->stream, s->d_pos, s->d_mom,
                     s->d_mass, S.natoms, s->nseg, d_C, d_freqs, h, Lx, Ly, Lz, s->d_flags);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pimd_o(System* s, const double* d_C, const double* d_c1, const double* d_c2, uint64_t seed,
                  uint64_t counter) {
  const int n = s->seg[0].natoms * 3;
  hipLaunchKernelGGL(pimd_o_kernel, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_mass,
                     s->seg[0].natoms, s->nseg, d_C, d_c1, d_c2, seed, counter, s->d_thermo);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// Nose-Hoover chain half step (jax_md simulate.nose_hoover_chain with chain_steps = 1,
// sy_steps = 1, as wrapped by deepmd_jax/routine.py:147-158), one thread per segment.
// thermo layout per segment (64 doubles): [0] invariant term, [1] initialised,
// [8..16) xi, [16..24) p_xi, [24..32) Q, [40] momentum scale.  Kinetic energy and degrees
// of freedom are per periodic image.
// One Nose-Hoover chain half step (jax_md simulate.nose_hoover_chain with chain_steps = 1,
// sy_steps = 1): advances the chain (xi, p_xi) of masses Q that couples to kinetic energy KE
// of dof degrees of freedom at temperature kT over dt/2, and returns the factor that
// scales the coupled momenta.
__device__ double nhc_half_step(double* xi, double* px, const double* Q, int C, double KE, double dof, double kT,
                                double dt) {
  const double d2 = dt / 2, d4 = d2 / 2, d8 = d4 / 2;
  const int M = C - 1;
  double G;
  if (M > 0) {
    G = px[M - 1] * px[M - 1] / Q[M - 1] - kT;
    px[M] += d4 * G;
    double pnew = px[M];
    for (int m = M - 1; m > 0; --m) {
      G = px[m - 1] * px[m - 1] / Q[m - 1] - kT;
      const double sc = exp(-d8 * pnew / Q[m + 1]);
      pnew = sc * (sc * px[m] + d4 * G);
      px[m] = pnew;
    }
  }
  G = 2.0 * KE - dof * kT;
  {
    const double sc = (M > 0) ? exp(-d8 * px[1] / Q[1]) : 1.0;
    px[0] = sc * (sc * px[0] + d4 * G);
  }
  const double scale = exp(-d2 * px[0] / Q[0]);
  KE *= scale * scale;
  for (int k = 0; k < C; ++k) xi[k] += d2 * px[k] / Q[k];
  G = 2.0 * KE - dof * kT;
  for (int m = 0; m < M; ++m) {
    const double sc = exp(-d8 * px[m + 1] / Q[m + 1]);
    const double pu = sc * (sc * px[m] + d4 * G);
    G = pu * pu / Q[m] - kT;
    px[m] = pu;
  }
  if (M > 0) px[M] += d4 * G;
  return scale;
}

// chain_steps x sy_steps substeps of the chain half step with Suzuki-Yoshida weights, as
// jax_md's half_step_chain_fn composes them (delta = dt / chain_steps * w[i % sy_steps]); the
// coupled kinetic energy follows each substep's scale.  jax_md is not in the reference tree:
// the weights are the published Suzuki-Yoshida table (orders 3, 5, 7).
__device__ double nhc_half_steps(double* xi, double* px, const double* Q, int C, double KE, double dof, double kT,
                                 double dt, int cs, int sy) {
  if (cs <= 1 && sy <= 1) return nhc_half_step(xi, px, Q, C, KE, dof, kT, dt);
  const double w1[1] = {1.0};
  const double w3[3] = {0.828981543588751, -0.657963087177502, 0.828981543588751};
  const double w5[5] = {0.2967324292201065, 0.2967324292201065, -0.186929716880426, 0.2967324292201065,
                        0.2967324292201065};
  const double w7[7] = {0.784513610477560, 0.235573213359357, -1.17767998417887, 1.31518632068391,
                        -1.17767998417887, 0.235573213359357, 0.784513610477560};
  const double* w = sy == 3 ? w3 : (sy == 5 ? w5 : (sy == 7 ? w7 : w1));
  const int n = cs * sy;
  double scale = 1.0;
  for (int i = 0; i < n; ++i) {
    const double s = nhc_half_step(xi, px, Q, C, KE, dof, kT, dt / cs * w[i % sy]);
    scale *= s;
    KE *= s * s;
  }
  return scale;
}

// Particle thermostat half step, one thread per segment (deepmd_jax/routine.py:147-158).
// thermo layout per segment (64 doubles): [0] invariant term, [1] initialised,
// [8..16) xi, [16..24) p_xi, [24..32) Q, [40] momentum scale.  Kinetic energy and degrees
// of freedom are per periodic image.
__global__ void nh_half_kernel(double* thermo, const double* ke, const SegInfo* seg, int nseg, int C, double kT,
                               double dt, double tau, double dof_mobile, int cs, int sy) {
  for (int rep = threadIdx.x; rep < nseg; rep += blockDim.x) {
    // degrees of freedom of the mobile atoms (with_fixed_atoms, deepmd_jax/routine.py:110-136)
    const double dof = dof_mobile > 0 ? dof_mobile : 3.0 * (seg[rep].natoms / seg[rep].nimg);
    double* th = thermo + rep * 64;
    double* xi = th + 8;
    double* px = th + 16;
    double* Q = th + 24;
    if (th[1] == 0.0) {
      for (int k = 0; k < C; ++k) { xi[k] = 0.0; px[k] = 0.0; }
      th[1] = 1.0;
    }
    for (int k = 0; k < C; ++k) Q[k] = kT * tau * tau;
    Q[0] *= dof;
    th[40] = nhc_half_steps(xi, px, Q, C, ke[rep], dof, kT, dt, cs, sy);
    double e = px[0] * px[0] / (2.0 * Q[0]) + dof * kT * xi[0];
    for (int k = 1; k < C; ++k) e += px[k] * px[k] / (2.0 * Q[k]) + kT * xi[k];
    th[0] = e;
  }
}

__global__ void nh_scale_kernel(double* mom, const double* thermo, const int* atom_seg, int total3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total3) mom[i] *= thermo[atom_seg[i / 3] * 64 + 40];
}

int32_t md_nh_half(System* s, double kT, double dt, double tau, int chain, int cs, int sy) {
  if (chain < 1 || chain > 8) {
    set_error("Nose-Hoover chain length %d outside [1, 8]", chain);
    return DPMD_E_ARG;
  }
  double* d_ke = s->d_red + 16 * s->nseg;
  DPMD_TRY(md_ke_launch(s, d_ke));
  hipLaunchKernelGGL(nh_half_kernel, dim3(1), dim3(256), 0, s->stream, s->d_thermo, d_ke, s->d_seg, s->nseg, chain,
                     kT, dt, tau, s->has_fixed ? 3.0 * s->n_mobile / s->seg[0].nimg : 0.0, cs, sy);
  const int n3 = s->NA * 3;
  hipLaunchKernelGGL(nh_scale_kernel, dim3((n3 + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_thermo,
                     s->d_atom_seg, n3);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// ------------------------------------------------------------------------------
// NPT: MTK Nose-Hoover barostat over coupled axis groups (deepmd_jax/routine.py:161-423,
// adapted from jax_md simulate.npt_nose_hoover), one segment.  Axis a belongs to group
// gid[a] (or is held fixed, gid = -1); a group's axes share one log-scale R_b.  Device
// state (64 doubles): [0..3) R_b, [3..6) P_b, [6..9) M_b, [9..12) dU/deps per group,
// [12..15) reference box, [15] initialised, [16..24) barostat chain xi, [24..32) p_xi,
// [32..40) Q, [40..43) per-axis 2 KE, [43] barostat terms of the invariant,
// [45..48) momentum rates, [48..51) position exponents.

__device__ __forceinline__ double sinhx_x(double x) {
  const double x2 = x * x;
  return 1 + x2 / 6 + x2 * x2 / 120 + x2 * x2 * x2 / 5040 + x2 * x2 * x2 * x2 / 362880 +
         x2 * x2 * x2 * x2 * x2 / 39916800;
}

__device__ __forceinline__ void npt_set_box(const NptArgs& a, double* npt, SegInfo& S) {
  for (int d = 0; d < 3; ++d) {
    const int g = a.gid[d];
    const double L = g >= 0 ? npt[12 + d] * exp(npt[g]) : npt[12 + d];
    S.boxd[d] = L;
    S.box[d] = (float)L;
    S.scale[d] = S.L[d] / (float)L;
  }
  // drift threshold with the NPT safety factor (deepmd_jax/md.py:724-728)
  seg_set_threshold(S, a.rc_skin, a.rc, true);
}

// dU/deps of each group from the virial diagonal (virial = -dU/dstrain): red[1 + 4 a] = W_aa
__device__ __forceinline__ void npt_dudv(const NptArgs& a, const double* red, double* npt) {
  for (int g = 0; g < 3; ++g) npt[9 + g] = 0.0;
  for (int d = 0; d < 3; ++d)
    if (a.gid[d] >= 0) npt[9 + a.gid[d]] -= red[1 + 4 * d];
}

// set-up (init_fn): zero box position / momentum, reference box, chain at rest, dU/deps
// of the current forces.  refresh only: dU/deps of freshly computed forces.
__global__ void npt_init_kernel(NptArgs a, double* npt, SegInfo* seg, const double* red, int refresh_only) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (!refresh_only || npt[15] == 0.0) {
    for (int k = 0; k < 64; ++k) npt[k] = 0.0;
    for (int d = 0; d < 3; ++d) npt[12 + d] = seg[0].boxd[d];
    npt[15] = 1.0;
    npt_set_box(a, npt, seg[0]);
  }
  npt_dudv(a, red, npt);
}

// Barostat chain half step on the box momenta, with the masses of update_mass /
// update_box_mass: Q = kT tau_p^2 (G, 1, 1, ...), M_b = n_g (N + 1) kT tau_p^2.  Also
// forms the barostat part of the invariant: chain terms + P V + KE_box.
__global__ void npt_baro_half_kernel(NptArgs a, double* npt, const SegInfo* seg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double tp2 = a.kT * a.tau_p * a.tau_p;
  double* xi = npt + 16;
  double* px = npt + 24;
  double* Q = npt + 32;
  for (int k = 0; k < a.C; ++k) Q[k] = tp2;
  Q[0] *= a.G;
  double ke = 0.0;
  for (int g = 0; g < a.G; ++g) {
    npt[6 + g] = a.nper[g] * (a.natoms + 1.0) * tp2;
    ke += npt[3 + g] * npt[3 + g] / (2.0 * npt[6 + g]);
  }
  const double sc = nhc_half_steps(xi, px, Q, a.C, ke, (double)a.G, a.kT, a.dt, a.cs, a.sy);
  double e = 0.0;
  for (int g = 0; g < a.G; ++g) {
    npt[3 + g] *= sc;
    e += npt[3 + g] * npt[3 + g] / (2.0 * npt[6 + g]);
  }
  for (int k = 0; k < a.C; ++k) e += px[k] * px[k] / (2.0 * Q[k]) + a.kT * xi[k];
  const SegInfo& S = seg[0];
  e += a.pressure * S.boxd[0] * S.boxd[1] * S.boxd[2];
  npt[43] = e;
}

// Box force G_g = 2KE_g + (n_g / N_f) 2KE - dU/deps_g - P V n_g and the half kick of P_b.
// phase 0 (first kick) then sets the per-axis rates of exp(iL2) and exponents of exp(iL1)
// and moves the box: R_b += dt P_b / M_b.  phase 1 takes dU/deps from the new virial first.
__global__ void npt_box_kernel(NptArgs a, double* npt, SegInfo* seg, const double* red, int phase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  SegInfo& S = seg[0];
  if (phase == 1) npt_dudv(a, red, npt);
  const double Nf = 3.0 * a.natoms;
  const double V = S.boxd[0] * S.boxd[1] * S.boxd[2];
  const double k2tot = npt[40] + npt[41] + npt[42];
  double k2g[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < 3; ++d)
    if (a.gid[d] >= 0) k2g[a.gid[d]] += npt[40 + d];
  for (int g = 0; g < a.G; ++g) {
    const double Ge = k2g[g] + (a.nper[g] / Nf) * k2tot - npt[9 + g] - a.pressure * V * a.nper[g];
    npt[3 + g] += 0.5 * a.dt * Ge;
  }
  if (phase == 0) {
    double vb[3], tr = 0.0;
    for (int g = 0; g < a.G; ++g) {
      vb[g] = npt[3 + g] / npt[6 + g];
      tr += a.nper[g] * vb[g];
    }
    for (int d = 0; d < 3; ++d) {
      const double v = a.gid[d] >= 0 ? vb[a.gid[d]] : 0.0;
      npt[45 + d] = v + tr / Nf;
      npt[48 + d] = v * a.dt;
    }
    for (int g = 0; g < a.G; ++g) npt[g] += vb[g] * a.dt;
    npt_set_box(a, npt, S);
  }
}

// exp(iL2): p = p e^{-x} + dt/2 F sinhx_x(x/2) e^{-x/2}, x = rate dt/2; with move, then
// exp(iL1): r = r e^{y} + dt (p/m) e^{y/2} sinhx_x(y/2), y = exponent, and the fp32
// mirror in the new box (NaN/Inf -> error bit 16).
__global__ void npt_particle_kernel(double* pos, double* mom, const double* mass, const float* force, float4* p32,
                                    const double* npt, const SegInfo* seg, int natoms, double dt, int move,
                                    int* flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  const SegInfo& S = seg[0];
  const double m = mass[i];
  float v[3];
  bool bad = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double x = npt[45 + d] * dt * 0.5;
    double p = mom[(size_t)i * 3 + d];
    p = p * exp(-x) + 0.5 * dt * (double)force[(size_t)i * 3 + d] * sinhx_x(0.5 * x) * exp(-0.5 * x);
    mom[(size_t)i * 3 + d] = p;
    if (move) {
      const double y = npt[48 + d];
      const double r = pos[(size_t)i * 3 + d] * exp(y) + dt * (p / m) * exp(0.5 * y) * sinhx_x(0.5 * y);
      pos[(size_t)i * 3 + d] = r;
      bad |= !isfinite(r) || !isfinite(p);
      v[d] = wrap32(r, S.boxd[d]);
    }
  }
  if (move) {
    p32[i] = make_float4(v[0], v[1], v[2], 0.f);
    if (bad) atomicOr(&flags[2], 16);
  }
}

// ------------------------------------------------------------------------------
// Path-integral NPT (deepmd_jax/pimd.py:297-518, npt_langevin_pimd): the centroid follows
// the MTK barostat step (exp(iL2) on the centroid momentum, exp(iL1) on the centroid
// position), the internal ring modes propagate exactly, PILE-L acts on the bead momenta and
// a Langevin O step on the box momenta.  The beads are the P segments (identical, no
// images) and share one box.  Same 64-double block as NPT, with [40..43) the per-axis
// centroid 2KE sum_i p_c^2 / m, [43] KE_box + P V - box work (the barostat part of the
// reference's invariant), [44] box thermostat work, [51..54) the internal virial
// sum_{b,i} (x_b - x_c) F_b per axis.

__global__ void pnpt_init_kernel(NptArgs a, double* npt, SegInfo* seg, int nseg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int k = 0; k < 64; ++k) npt[k] = 0.0;
  for (int d = 0; d < 3; ++d) npt[12 + d] = seg[0].boxd[d];
  for (int g = 0; g < a.G; ++g) npt[6 + g] = a.nper[g] * (a.natoms + 1.0) * a.kT * a.tau_p * a.tau_p;
  npt[15] = 1.0;
  for (int r = 0; r < nseg; ++r) npt_set_box(a, npt, seg[r]);
}

// Pass 1 of a per-axis sum over the (atom, axis) pairs of the ring polymer, a strided share
// per block: mode 0 the centroid 2KE (p_c = mean over beads), mode 1 the internal virial
// with the beads unwrapped onto bead 0's image (pimd.py:210-227).  part[3 block + axis].
__global__ void pnpt_part_kernel(const double* pos, const double* mom, const double* mass, const float* force,
                                 const SegInfo* seg, int natoms, int P, int mode, int KB, double* part) {
  __shared__ double red[256];
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  const double invP = 1.0 / P;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < natoms * 3; t += KB * blockDim.x) {
    const int i = t / 3, d = t - 3 * i;
    double v = 0.0;
    if (mode == 0) {
      double pc = 0.0;
      for (int b = 0; b < P; ++b) pc += mom[((size_t)b * natoms + i) * 3 + d];
      pc *= invP;
      v = pc * pc / mass[i];
    } else {
      const double L = seg[0].boxd[d];
      const double x0 = pos[(size_t)i * 3 + d];
      double sq = 0.0;
      for (int b = 0; b < P; ++b) {
        double dl = pos[((size_t)b * natoms + i) * 3 + d] - x0;
        dl -= L * rint(dl / L);
        sq += x0 + dl;
      }
      const double xc = sq * invP;
      for (int b = 0; b < P; ++b) {
        double dl = pos[((size_t)b * natoms + i) * 3 + d] - x0;
        dl -= L * rint(dl / L);
        v += (x0 + dl - xc) * (double)force[((size_t)b * natoms + i) * 3 + d];
      }
    }
    acc0 += d == 0 ? v : 0.0;
    acc1 += d == 1 ? v : 0.0;
    acc2 += d == 2 ? v : 0.0;
  }
  const double s0 = block_sum(acc0, red);
  const double s1 = block_sum(acc1, red);
  const double s2 = block_sum(acc2, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 3 + 0] = s0;
    part[blockIdx.x * 3 + 1] = s1;
    part[blockIdx.x * 3 + 2] = s2;
  }
}

// pass 2: the KB partials of each axis added in a fixed order
__global__ void pnpt_sum3_kernel(const double* part, int KB, double* out) {
  const int d = threadIdx.x;
  if (d >= 3 || blockIdx.x != 0) return;
  double v = 0.0;
  for (int b = 0; b < KB; ++b) v += part[b * 3 + d];
  out[d] = v;
}

// dU/deps per group, the centroid-virial form of force_stress_fn (pimd.py:336-353):
// (sum over beads of -W_aa + internal virial_a) / P over the group's axes
__global__ void pnpt_dudv_kernel(NptArgs a, double* npt, const double* red, int P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int g = 0; g < 3; ++g) npt[9 + g] = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (a.gid[d] < 0) continue;
    double v = npt[51 + d];
    for (int r = 0; r < P; ++r) v -= red[16 * r + 1 + 4 * d];
    npt[9 + a.gid[d]] += v / P;
  }
}

// half kick of the box momenta, G_g = 2KE_c,g + (n_g / N_f) 2KE_c - dU/deps_g - P V n_g
// (box_force_cv), then the barostat part of the invariant
__global__ void pnpt_kick_kernel(NptArgs a, double* npt, const SegInfo* seg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const SegInfo& S = seg[0];
  const double Nf = 3.0 * a.natoms;
  const double V = S.boxd[0] * S.boxd[1] * S.boxd[2];
  const double k2tot = npt[40] + npt[41] + npt[42];
  double k2g[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < 3; ++d)
    if (a.gid[d] >= 0) k2g[a.gid[d]] += npt[40 + d];
  double e = a.pressure * V - npt[44];
  for (int g = 0; g < a.G; ++g) {
    const double Ge = k2g[g] + (a.nper[g] / Nf) * k2tot - npt[9 + g] - a.pressure * V * a.nper[g];
    npt[3 + g] += 0.5 * a.dt * Ge;
    e += npt[3 + g] * npt[3 + g] / (2.0 * npt[6 + g]);
  }
  npt[43] = e;
}

// R_b += h P_b / M_b and the new box on every bead
__global__ void pnpt_move_kernel(NptArgs a, double* npt, SegInfo* seg, int nseg, double h) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int g = 0; g < a.G; ++g) npt[g] += h * npt[3 + g] / npt[6 + g];
  for (int r = 0; r < nseg; ++r) npt_set_box(a, npt, seg[r]);
}

// Langevin O step of the box momenta, c1 = exp(-gamma dt), c2 = sqrt(kT (1 - c1^2)); the
// change of box kinetic energy accumulates in the box work
__global__ void pnpt_box_o_kernel(NptArgs a, double* npt, double c1, double c2, uint64_t seed, uint64_t counter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double z[4];
  normal4(seed, counter, 0u, 0x424f5800u, z);
  for (int g = 0; g < a.G; ++g) {
    const double Mb = npt[6 + g], p0 = npt[3 + g];
    const double p1 = c1 * p0 + c2 * sqrt(Mb) * z[g];
    npt[44] += 0.5 * (p1 * p1 - p0 * p0) / Mb;
    npt[3 + g] = p1;
  }
}

// exp(iL2) on the centroid momentum with the per-axis rate v_b,a + tr(v_b) / N_f, a plain
// kick of the internal part (exp_iL2_pimd)
__global__ void pnpt_il2_kernel(double* mom, const float* force, const double* npt, NptArgs a, int natoms, int P,
                                double dt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms * 3) return;
  const int i = t / 3, d = t - 3 * i;
  double tr = 0.0, vd = 0.0;
  for (int g = 0; g < a.G; ++g) {
    const double vb = npt[3 + g] / npt[6 + g];
    tr += a.nper[g] * vb;
    if (g == a.gid[d]) vd = vb;
  }
  const double x = (vd + tr / (3.0 * a.natoms)) * dt * 0.5;
  double pc = 0.0, fc = 0.0;
  for (int b = 0; b < P; ++b) {
    pc += mom[((size_t)b * natoms + i) * 3 + d];
    fc += (double)force[((size_t)b * natoms + i) * 3 + d];
  }
  pc /= P;
  fc /= P;
  const double pcn = pc * exp(-x) + 0.5 * dt * fc * sinhx_x(0.5 * x) * exp(-0.5 * x);
  for (int b = 0; b < P; ++b) {
    const size_t o = ((size_t)b * natoms + i) * 3 + d;
    mom[o] = pcn + (mom[o] - pc) + 0.5 * dt * ((double)force[o] - fc);
  }
}

// exp(iL1) on the centroid (every bead shifted by dx_c = x_c (e^y - 1) + h v_c e^{y/2}
// sinhx_x(y/2), y = v_b,a h) and the exact propagation of the internal normal modes over h
// (exp_iL1_centroid_pimd, exact_internal_ring_step), beads unwrapped onto bead 0's image.
__global__ void pnpt_drift_ring_kernel(double* pos, double* mom, const double* mass, const double* npt, NptArgs a,
                                       const SegInfo* seg, int natoms, int P, const double* C, const double* freqs,
                                       double h, int* flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= natoms * 3) return;
  const int i = t / 3, d = t - 3 * i;
  const double L = seg[0].boxd[d];
  const double m = mass[i];
  const int g = a.gid[d];
  double vd = 0.0;
  for (int k = 0; k < a.G; ++k)
    if (k == g) vd = npt[3 + k] / npt[6 + k];
  const double y = vd * h;
  double q[64], p[64], qn[64], pn[64];
  const double x0 = pos[(size_t)i * 3 + d];
  double xc = 0.0, vc = 0.0;
  for (int b = 0; b < P; ++b) {
    double dl = pos[((size_t)b * natoms + i) * 3 + d] - x0;
    dl -= L * rint(dl / L);
    q[b] = x0 + dl;
    p[b] = mom[((size_t)b * natoms + i) * 3 + d];
    xc += q[b];
    vc += p[b];
  }
  xc /= P;
  vc /= P * m;
  const double dxc = xc * (exp(y) - 1.0) + h * vc * exp(0.5 * y) * sinhx_x(0.5 * y);
  for (int b = 0; b < P; ++b) q[b] += dxc;
  for (int k = 0; k < P; ++k) {
    double sq = 0, sp = 0;
    for (int b = 0; b < P; ++b) {
      sq += C[b * P + k] * q[b];
      sp += C[b * P + k] * p[b];
    }
    qn[k] = sq;
    pn[k] = sp;
  }
  for (int k = 1; k < P; ++k) {
    const double w = freqs[k];
    const double cs = cos(w * h), sn = sin(w * h);
    const double qq = qn[k] * cs + pn[k] * sn / (m * w);
    const double pp = pn[k] * cs - m * w * qn[k] * sn;
    qn[k] = qq;
    pn[k] = pp;
  }
  bool bad = false;
  for (int b = 0; b < P; ++b) {
    double sq = 0, sp = 0;
    for (int k = 0; k < P; ++k) {
      sq += C[b * P + k] * qn[k];
      sp += C[b * P + k] * pn[k];
    }
    pos[((size_t)b * natoms + i) * 3 + d] = sq;
    mom[((size_t)b * natoms + i) * 3 + d] = sp;
    bad |= !isfinite(sq) || !isfinite(sp);
  }
  if (bad) atomicOr(&flags[2], 16);
}

__global__ void box_record_kernel(const SegInfo* seg, int nseg, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * nseg) out[i] = seg[i / 3].boxd[i % 3];
}

int32_t md_npt_init(System* s, const NptArgs& a, bool refresh_only) {
  DPMD_TRY(md_reduce_launch(s));
  hipLaunchKernelGGL(npt_init_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_seg, s->d_red,
                     refresh_only ? 1 : 0);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_npt_baro_half(System* s, const NptArgs& a) {
  hipLaunchKernelGGL(npt_baro_half_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_seg);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_npt_box(System* s, const NptArgs& a, int phase) {
  // per-axis sum of p^2 / m -> npt[40..43)
  DPMD_TRY(ke_launch(s, nullptr, s->d_npt + 40));
  if (phase == 1) DPMD_TRY(md_reduce_launch(s));
  hipLaunchKernelGGL(npt_box_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_seg, s->d_red, phase);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_npt_particles(System* s, double dt, bool move) {
  const int NA = s->NA;
  hipLaunchKernelGGL(npt_particle_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, s->d_pos, s->d_mom,
                     s->d_mass, s->d_force, s->d_pos32, s->d_npt, s->d_seg, NA, dt, move ? 1 : 0, s->d_flags);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// blocks for the per-axis ring-polymer sums: about 2,048 (atom, axis) pairs per block
static int pnpt_blocks(int n3) { return std::max(1, std::min(kKeBlocks, (n3 + 2047) / 2048)); }

int32_t md_pnpt_init(System* s, const NptArgs& a) {
  hipLaunchKernelGGL(pnpt_init_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_seg, s->nseg);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

static int32_t pnpt_sum(System* s, int mode, double* out) {
  const int N = s->seg[0].natoms, KB = pnpt_blocks(N * 3);
  hipLaunchKernelGGL(pnpt_part_kernel, dim3(KB), dim3(256), 0, s->stream, s->d_pos, s->d_mom, s->d_mass, s->d_force,
                     s->d_seg, N, s->nseg, mode, KB, s->d_kepart);
  hipLaunchKernelGGL(pnpt_sum3_kernel, dim3(1), dim3(64), 0, s->stream, s->d_kepart, KB, out);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pnpt_dudv(System* s, const NptArgs& a) {
  DPMD_TRY(md_reduce_launch(s));
  DPMD_TRY(pnpt_sum(s, 1, s->d_npt + 51));
  hipLaunchKernelGGL(pnpt_dudv_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_red, s->nseg);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pnpt_kick(System* s, const NptArgs& a) {
  DPMD_TRY(pnpt_sum(s, 0, s->d_npt + 40));
  hipLaunchKernelGGL(pnpt_kick_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_seg);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pnpt_move(System* s, const NptArgs& a, double h) {
  hipLaunchKernelGGL(pnpt_move_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, s->d_seg, s->nseg, h);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pnpt_box_o(System* s, const NptArgs& a, double gamma, uint64_t seed, uint64_t counter) {
  const double c1 = exp(-gamma * a.dt);
  const double c2 = sqrt(std::max(a.kT * (1.0 - c1 * c1), 0.0));
  hipLaunchKernelGGL(pnpt_box_o_kernel, dim3(1), dim3(64), 0, s->stream, a, s->d_npt, c1, c2, seed, counter);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pnpt_il2(System* s, const NptArgs& a) {
  const int N = s->seg[0].natoms;
  hipLaunchKernelGGL(pnpt_il2_kernel, dim3((N * 3 + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_force,
                     s->d_npt, a, N, s->nseg, a.dt);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_pnpt_drift_ring(System* s, const NptArgs& a, const double* C, const double* freqs, double h) {
  const int N = s->seg[0].natoms;
  hipLaunchKernelGGL(pnpt_drift_ring_kernel, dim3((N * 3 + 127) / 128), dim3(128), 0, s->stream, s->d_pos,
                     s->d_mom, s->d_mass, s->d_npt, a, s->d_seg, N, s->nseg, C, freqs, h, s->d_flags);
  // the fp32 mirror in the box the device just set (no upload of the host segments)
  const int NA = s->NA;
  hipLaunchKernelGGL(write_pos32_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, s->d_pos, s->d_pos32,
                     s->d_seg, s->d_atom_seg, NA);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_box_record(System* s, double* out) {
  hipLaunchKernelGGL(box_record_kernel, dim3((3 * s->nseg + 63) / 64), dim3(64), 0, s->stream, s->d_seg, s->nseg,
                     out);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// ------------------------------------------------------------------------------
// Fixed atoms and centre-of-mass motion (deepmd_jax/md.py:221-234,428-430,744-757,786-790)

// rows of fixed atoms set to zero: forces (float) after each evaluation, momenta (double)
// after each NVE / Nose-Hoover step (with_fixed_atoms, deepmd_jax/routine.py:130-134)
__global__ void mask_force_kernel(float* f, const float* mobile, int n3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3 && mobile[i / 3] == 0.f) f[i] = 0.f;
}

__global__ void mask_mom_kernel(double* p, const float* mobile, int n3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3 && mobile[i / 3] == 0.f) p[i] = 0.0;
}

// pass 1: per block, sum of mobile momenta (3) and mobile mass over a strided share of all atoms
__global__ void com_part_kernel(const double* mom, const double* mass, const float* mobile, int natoms, int KB,
                                double* part) {
  __shared__ double red[256];
  double k[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < natoms; i += KB * blockDim.x) {
    const double w = mobile ? (double)mobile[i] : 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) k[d] += w * mom[(size_t)i * 3 + d];
    k[3] += w * mass[i];
  }
  for (int d = 0; d < 4; ++d) {
    const double v = block_sum(k[d], red);
    if (threadIdx.x == 0) part[blockIdx.x * 4 + d] = v;
  }
}

// pass 2: every thread adds the partials in the same order and removes m v_com from its atom
__global__ void com_apply_kernel(double* mom, const double* mass, const double* part, int KB, int natoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natoms) return;
  double k[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < KB; ++b)
    for (int d = 0; d < 4; ++d) k[d] += part[b * 4 + d];
  const double m = mass[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) mom[(size_t)i * 3 + d] -= m * k[d] / k[3];
}

int32_t md_mask_forces(System* s) {
  const int n3 = s->NA * 3;
  hipLaunchKernelGGL(mask_force_kernel, dim3((n3 + 255) / 256), dim3(256), 0, s->stream, s->d_force, s->d_mobile, n3);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t md_mask_momenta(System* s) {
  const int n3 = s->NA * 3;
  hipLaunchKernelGGL(mask_mom_kernel, dim3((n3 + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_mobile, n3);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// v -= sum(v m_mobile) / sum(m_mobile) for every atom of the system (all beads together)
int32_t md_remove_com(System* s) {
  const int NA = s->NA;
  const int KB = std::max(1, std::min(kKeBlocks, (NA + 2047) / 2048));
  hipLaunchKernelGGL(com_part_kernel, dim3(KB), dim3(256), 0, s->stream, s->d_mom, s->d_mass,
                     s->has_fixed ? s->d_mobile : nullptr, NA, KB, s->d_kepart);
  hipLaunchKernelGGL(com_apply_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, s->d_mom, s->d_mass,
                     s->d_kepart, KB, NA);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

}  // namespace dpmd

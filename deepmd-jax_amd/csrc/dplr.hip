// DPLR: the long-range electrostatics of deepmd_jax/md.py:404-415 on the device.
//
// A Wannier-centroid model (an atomic vector model, deepmd_jax/dpmodel.py:83-98) predicts
// for each atom of its selected types a displacement dx = out_norm sum_w f_w T3[:, w]; the
// PPPM energy (pppm.hip) of charge q_atom at every atom and q_wc at every centroid x + dx is
// added to the energy.  Forces: the PPPM force on an atom's charge and on its centroid's
// charge act on the atom directly (d(x + dx)/dx = 1 + d(dx)/dx), and the remaining term
// -(d dx / dx)^T g with g = dE/d(centroid) is the backward of the centroid model for the
// upstream gradient g: through the hidden fitting layers into dE/dG, plus the direct
// dE/dT3 = out_norm g f, into the embedding backward (which accumulates the forces).
//
// Virial (-dE/dstrain with atoms and box scaled, as the reference's perturbation does): the
// centroids x + dx do not scale with the box, dx follows the centroid model, so
//   W_lr = W_recip (PPPM with every charge scaling) + W' + sum_k g_k (x) dx_k,
// W' = -d(g . dx)/dstrain, the virial the embedding backward accumulates for the centroid
// model's pseudo-energy g . dx.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "internal.h"

namespace dpmd {

namespace {

struct RowTypes {
  int rowbase[kMaxT + 1];
  int ntypes;
  const float* dt2[kMaxT];
};

// Charge index maps: v >= 0 for an atom or centroid that owns its charge (the primary cell of
// a supercell segment), -(q + 1) for an image, which reads its primary's charge q.
__host__ __device__ __forceinline__ int qdec(int v) { return v >= 0 ? v : -v - 1; }

// charge positions: the atoms and the centroids x + out_norm sum_w f_w T3[c][w], each at
// its charge index (segment blocks; images write nothing)
__global__ void dplr_charges_kernel(const double* pos, const float* T, const float* Y3, int ldy, int W,
                                    float out_norm, const int* wc_atom, const int* wc_row, const int* atom_q,
                                    const int* wc_q, int NA, int nwc, double* qxyz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < NA * 3) {
    const int g = t / 3;
    if (atom_q[g] >= 0) qxyz[(size_t)atom_q[g] * 3 + (t - 3 * g)] = pos[t];
    return;
  }
  const int u = t - NA * 3;
  if (u >= nwc * 3) return;
  const int k = u / 3, c = u - 3 * k;
  if (wc_q[k] < 0) return;
  const int g = wc_atom[k];
  const float* y = Y3 + (size_t)wc_row[k] * ldy;
  const float* t3 = T + ((size_t)g * 10 + 1 + c) * W;
  float s = 0.f;
  for (int w = 0; w < W; ++w) s += y[w] * t3[w];
  qxyz[(size_t)wc_q[k] * 3 + c] = pos[(size_t)g * 3 + c] + (double)(out_norm * s);
}

// forces of the atom and centroid charges onto the atoms; g = dE/d(centroid) per atom
__global__ void dplr_scatter_kernel(const double* qf, const int* atom_wc, const int* atom_q, const int* wc_q, int NA,
                                    float* force, float* wc_g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NA * 3) return;
  const int g = t / 3, c = t - 3 * g;
  const int k = atom_wc[g];
  double f = qf[(size_t)qdec(atom_q[g]) * 3 + c];
  float gw = 0.f;
  if (k >= 0) {
    const double fw = qf[(size_t)qdec(wc_q[k]) * 3 + c];
    f += fw;
    gw = (float)(-fw);
  }
  force[t] += (float)f;
  wc_g[t] = gw;
}

// dE/df[row][w] = out_norm sum_c g[c] T3[c][w] (DY3) and DZ3 = DY3 dt2 (1 - t3^2) for the
// last (residual) hidden layer
__global__ void dplr_dy3_kernel(const float* wc_g, const float* T, const float* T3l, int ldy, int W, float out_norm,
                                const int* wc_atom, const int* wc_row, int nwc, RowTypes rt, float* DY3, float* DZ3) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwc * W) return;
  const int k = t / W, w = t - W * k;
  const int g = wc_atom[k], row = wc_row[k];
  int ty = 0;
  while (ty + 1 < rt.ntypes && row >= rt.rowbase[ty + 1]) ++ty;
  const float* gg = wc_g + (size_t)g * 3;
  const float* Tg = T + (size_t)g * 10 * W;
  const float d = out_norm * (gg[0] * Tg[W + w] + gg[1] * Tg[2 * W + w] + gg[2] * Tg[3 * W + w]);
  const size_t o = (size_t)row * ldy + w;
  const float th = T3l[o];
  DY3[o] = d;
  DZ3[o] = d * rt.dt2[ty][w] * (1.f - th * th);
}

// atomic model outputs per selected atom: mode 1 the vector out_norm sum_w f_w T3[c][w];
// mode 2 the tensor head Y = T6 + 3to6(T3) + TW (1,1,1,0,0,0), 6 components mapped to the 9 of a
// symmetric 3x3 (deepmd_jax/dpmodel.py:90-99)
__global__ void atomic_out_kernel(const float* T, const float* tbias, const float* Y3, int ldy, int W,
                                  float out_norm, const int* wc_atom, const int* wc_row, int nwc, int mode,
                                  double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nwc) return;
  const int g = wc_atom[k];
  const float* y = Y3 + (size_t)wc_row[k] * ldy;
  const float* Tg = T + (size_t)g * 10 * W;
  if (mode == 1) {
    float v[3] = {0.f, 0.f, 0.f};
    for (int w = 0; w < W; ++w)
      for (int c = 0; c < 3; ++c) v[c] += y[w] * Tg[(1 + c) * W + w];
    for (int c = 0; c < 3; ++c) out[(size_t)k * 3 + c] = (double)(out_norm * v[c]);
    return;
  }
  const float r2 = 1.4142135623730951f;
  float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int w = 0; w < W; ++w) {
    const float tw = Tg[w] + tbias[w];
    const float t3[3] = {Tg[W + w], Tg[2 * W + w], Tg[3 * W + w]};
    for (int c = 0; c < 3; ++c) {
      v[c] += y[w] * (Tg[(4 + c) * W + w] + t3[c] * t3[c] + tw);
      v[3 + c] += y[w] * (Tg[(7 + c) * W + w] + r2 * t3[c] * t3[(c + 2) % 3]);
    }
  }
  const float s = 0.7071067811865476f;
  const int src[9] = {0, 4, 3, 4, 1, 5, 3, 5, 2};
  const float sc[9] = {1.f, s, s, s, 1.f, s, s, s, 1.f};
  for (int e = 0; e < 9; ++e) out[(size_t)k * 9 + e] = (double)(out_norm * v[src[e]] * sc[e]);
}

// segment y of the grid: red[16 y + i] += lr[kLrStride y + i], i < 10
__global__ void dplr_add_energy_kernel(double* red, const double* lr) {
  if (threadIdx.x < 10) red[(size_t)blockIdx.x * 16 + threadIdx.x] += lr[(size_t)blockIdx.x * kLrStride + threadIdx.x];
}

// pass 1, per (block, segment) (strided share of the segment): W' (the centroid backward's
// block partials, per image) + sum_k g_k (x) dx_k over the primary centroids ->
// lr[kLrStride seg + 10 + 9 block + c]
__global__ void dplr_virial_kernel(const double* partial, const SegInfo* seg, const int* wc_seg, const float* wc_g,
                                   const double* qxyz, const double* pos, const int* wc_atom, const int* wc_q,
                                   double* lr) {
  __shared__ double red[256];
  const SegInfo& S = seg[blockIdx.y];
  double* part = lr + (size_t)blockIdx.y * kLrStride + 10;
  double v[9];
  for (int c = 0; c < 9; ++c) v[c] = 0.0;
  const int stride = gridDim.x * blockDim.x, t0 = blockIdx.x * blockDim.x + threadIdx.x;
  const double inv = 1.0 / S.nimg;
  for (int b = t0; b < S.nblk; b += stride)
    for (int c = 0; c < 9; ++c) v[c] += partial[(size_t)(S.blk0 + b) * 9 + c] * inv;
  const int k0 = wc_seg[blockIdx.y], k1 = wc_seg[blockIdx.y + 1];
  for (int k = k0 + t0; k < k1; k += stride) {
    if (wc_q[k] < 0) continue;
    const int g = wc_atom[k];
    double dx[3];
    for (int d = 0; d < 3; ++d) dx[d] = qxyz[(size_t)wc_q[k] * 3 + d] - pos[(size_t)g * 3 + d];
    for (int a = 0; a < 3; ++a)
      for (int bb = 0; bb < 3; ++bb) v[a * 3 + bb] += (double)wc_g[(size_t)g * 3 + a] * dx[bb];
  }
  for (int c = 0; c < 9; ++c) {
    red[threadIdx.x] = v[c];
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
      if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * 9 + c] = red[0];
    __syncthreads();
  }
}

// pass 2, one block per segment: lr[1..10) += the block partials, added in a fixed order
__global__ void dplr_virial_sum_kernel(double* lr) {
  const int c = threadIdx.x;
  if (c >= 9) return;
  double* l = lr + (size_t)blockIdx.x * kLrStride;
  double s = 0.0;
  for (int b = 0; b < kLrBlocks; ++b) s += l[10 + b * 9 + c];
  l[1 + c] += s;
}

}  // namespace

int32_t dplr_add_energy(System* s) {
  hipLaunchKernelGGL(dplr_add_energy_kernel, dim3(s->nseg), dim3(64), 0, s->stream, s->d_red, s->d_lr);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

static int32_t dplr_run(System* s, Model* M) {
  const int NA = s->NA, nwc = s->n_wc, W = M->wout;
  const hipStream_t st = s->stream;
  // centroid model: descriptor, hidden layers, centroids
  DPMD_TRY(embed_launch(s, false, DPMD_PREC_FP32));
  AtomicFitBufs b;
  DPMD_TRY(fit_atomic_buffers(s, M, b));
  DPMD_TRY(fit_atomic_fwd(s, b));
  const int nq = NA + nwc;
  hipLaunchKernelGGL(dplr_charges_kernel, dim3((nq * 3 + 255) / 256), dim3(256), 0, st, s->d_pos, s->d_T, b.Y3,
                     b.FK, W, (float)M->out_norm, s->d_wc_atom, s->d_wc_row, s->d_atom_q, s->d_wc_q, NA, nwc,
                     s->d_qxyz);
  // long-range energy and the forces on every charge, one PPPM per segment in the segment's
  // box as the device holds it (it moves on the device under NPT)
  DPMD_HIP(hipMemsetAsync(s->d_qf, 0, sizeof(double) * 3 * nq, st));
  for (int r = 0; r < s->nseg; ++r) {
    // a supercell segment's long-range term is that of its primary cell
    const size_t box_at = s->seg[r].nimg > 1 ? offsetof(SegInfo, cellbox) : offsetof(SegInfo, boxd);
    const double* d_box =
        reinterpret_cast<const double*>(reinterpret_cast<const char*>(s->d_seg + r) + box_at);
    const size_t q0 = (size_t)s->q_base[r];
    DPMD_TRY(s->pppm->compute(st, d_box, s->d_qxyz + 3 * q0, s->d_q + q0, s->q_n[r], s->d_lr + (size_t)r * kLrStride,
                              s->d_qf + 3 * q0));
  }
  hipLaunchKernelGGL(dplr_scatter_kernel, dim3((NA * 3 + 255) / 256), dim3(256), 0, st, s->d_qf, s->d_atom_wc,
                     s->d_atom_q, s->d_wc_q, NA, s->d_force, s->d_wc_g);
  // backward of the centroid model for the upstream gradient g
  RowTypes rt;
  memset(&rt, 0, sizeof(rt));
  rt.ntypes = s->ntypes;
  for (int t = 0; t <= s->ntypes; ++t) rt.rowbase[t] = s->rowbase[t];
  for (int t = 0; t < s->ntypes; ++t) rt.dt2[t] = M->d_fit + M->fit_off[t][10];  // FDT2 block
  if (nwc > 0)
    hipLaunchKernelGGL(dplr_dy3_kernel, dim3((nwc * W + 255) / 256), dim3(256), 0, st, s->d_wc_g, s->d_T, b.T3, b.FK,
                       W, (float)M->out_norm, s->d_wc_atom, s->d_wc_row, nwc, rt, b.DY3, b.DZ3);
  DPMD_HIP(hipGetLastError());
  DPMD_TRY(fit_atomic_bwd(s, b));
  s->wc_g = s->d_wc_g;
  s->wc_y3 = b.Y3;
  s->wc_ldy = b.FK;
  s->wc_norm = (float)M->out_norm;
  s->partial_out = s->d_partial_wc;
  const int32_t rc = embed_launch(s, true, DPMD_PREC_FP32);
  s->wc_g = nullptr;
  s->wc_y3 = nullptr;
  s->partial_out = nullptr;
  DPMD_TRY(rc);
  hipLaunchKernelGGL(dplr_virial_kernel, dim3(kLrBlocks, s->nseg), dim3(256), 0, st, s->d_partial_wc, s->d_seg,
                     s->d_wc_seg, s->d_wc_g, s->d_qxyz, s->d_pos, s->d_wc_atom, s->d_wc_q, s->d_lr);
  hipLaunchKernelGGL(dplr_virial_sum_kernel, dim3(s->nseg), dim3(64), 0, st, s->d_lr);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t dplr_forces(System* s) {
  for (const SegInfo& S : s->seg)
    if (S.tri) {
      set_error("DPLR runs orthorhombic boxes (deepmd_jax/md.py:391-392)");
      return DPMD_E_UNSUPPORTED;
    }
  Model* sr = s->model;
  s->model = s->wc_model;
  const int32_t rc = dplr_run(s, s->wc_model);
  s->model = sr;
  s->wc_g = nullptr;
  s->wc_y3 = nullptr;
  s->partial_out = nullptr;
  return rc;
}

}  // namespace dpmd

using namespace dpmd;

#define SYS(h)                                             \
  System* s = reinterpret_cast<System*>(h);                \
  if (!s) {                                                \
    set_error("null system");                              \
    return DPMD_E_ARG;                                     \
  }                                                        \
  DPMD_HIP(hipSetDevice(s->model->device));

extern "C" int32_t dpmd_system_set_dplr(dpmd_system_t sys, dpmd_model_t wc_model, const double* q_atoms,
                                        const double* q_wc, double beta, const int32_t* grid) {
  SYS(sys);
  s->force_valid = false;
  if (!wc_model) {
    s->wc_model = nullptr;
    return DPMD_OK;
  }
  Model* W = reinterpret_cast<Model*>(wc_model);
  Model* M = s->model;
  if (!q_atoms || !q_wc || !grid) { set_error("null argument"); return DPMD_E_ARG; }
  if (W->atomic != 1) { set_error("the DPLR centroid model must be an atomic vector model"); return DPMD_E_ARG; }
  if (W->device != M->device) { set_error("the centroid model lives on another device"); return DPMD_E_ARG; }
  if (W->ntypes != M->ntypes || W->wout != M->wout || W->axis != M->axis || W->lay.tiles[0] != M->lay.tiles[0] ||
      W->lay.tiles[1] != M->lay.tiles[1] || W->lay.tiles[2] != M->lay.tiles[2] || W->lay.z[1] != M->lay.z[1] ||
      W->lay.z[2] != M->lay.z[2]) {
    set_error("the centroid model needs the energy model's types, embedding widths and axis");
    return DPMD_E_UNSUPPORTED;
  }
  if (W->tab_n > 0) { set_error("tabulated centroid models are not supported"); return DPMD_E_UNSUPPORTED; }
  for (const SegInfo& S : s->seg)
    if (S.nimg > 1 && (S.kimg[0] < 1 || S.kimg[1] < 1 || S.kimg[2] < 1)) {
      set_error("DPLR on periodic images needs their per-axis counts (dpmd_system_set_image_grid)");
      return DPMD_E_STATE;
    }
  if (W->rcut > M->rcut) {
    set_error("the centroid model's cutoff %.3f exceeds the energy model's %.3f (shared neighbour list)", W->rcut,
              M->rcut);
    return DPMD_E_UNSUPPORTED;
  }
  for (const SegInfo& S : s->seg)
    if (S.tri) {
      set_error("DPLR runs orthorhombic boxes (deepmd_jax/md.py:391-392)");
      return DPMD_E_UNSUPPORTED;
    }
  if (!(beta > 0)) { set_error("DPLR needs beta > 0"); return DPMD_E_ARG; }
  for (int d = 0; d < 3; ++d)
    if (grid[d] < 2 || grid[d] > 4096) { set_error("PPPM grid size %d outside [2, 4096]", grid[d]); return DPMD_E_ARG; }
  DPMD_TRY(atomic_setup(s, W));
  const int NA = s->NA, nwc = s->n_wc, nq = NA + nwc;
  // charge blocks per segment: the primary cell's atoms, then its centroids, in the order
  // atomic_setup gives them (segment by segment, sorted atom order).  A segment holding nimg
  // images lists each type's atoms image-major, so the first n_t / nimg of type t are the
  // primary cell; every image reads its primary's charge (encoded -(q + 1)).
  std::vector<int> atom_q(NA), wc_q(std::max(nwc, 1)), wc_seg(s->nseg + 1);
  std::vector<double> q;
  s->q_base.assign(s->nseg, 0);
  s->q_n.assign(s->nseg, 0);
  int k = 0, qi = 0;
  for (int r = 0; r < s->nseg; ++r) {
    const SegInfo& S = s->seg[r];
    const int ncell = S.natoms / S.nimg;
    wc_seg[r] = k;
    s->q_base[r] = qi;
    // primary atom and centroid charges of the cell
    std::vector<int> prim_q(S.natoms), prim_wc(S.natoms, -1);
    int nw = 0;
    for (int c = 0; c < S.natoms; ++c) {
      const int t = s->type_sorted[S.base + c];
      const int nt_ = (S.off[t + 1] - S.off[t]) / S.nimg;
      const int local = c - S.off[t];
      if (local < nt_) prim_q[c] = qi + S.off[t] / S.nimg + local;
    }
    for (int c = 0; c < S.natoms; ++c) {
      const int t = s->type_sorted[S.base + c];
      const int nt_ = (S.off[t + 1] - S.off[t]) / S.nimg;
      if (c - S.off[t] < nt_ && (W->nsel_mask >> t & 1)) prim_wc[c] = qi + ncell + nw++;
    }
    q.resize((size_t)qi + ncell + nw);
    for (int c = 0; c < S.natoms; ++c) {
      const int g = S.base + c, t = s->type_sorted[g];
      const int nt_ = (S.off[t + 1] - S.off[t]) / S.nimg;
      const int local = c - S.off[t];
      const int pc = S.off[t] + local % nt_;  // primary copy
      const bool own = local < nt_;
      atom_q[g] = own ? prim_q[c] : -prim_q[pc] - 1;
      if (own) q[prim_q[c]] = q_atoms[t];
      if (!(W->nsel_mask >> t & 1)) continue;
      wc_q[k] = own ? prim_wc[c] : -prim_wc[pc] - 1;
      if (own) q[prim_wc[c]] = q_wc[t];
      ++k;
    }
    s->q_n[r] = ncell + nw;
    qi += s->q_n[r];
  }
  wc_seg[s->nseg] = k;
  const int nqc = qi;  // charges (primary cells only)
  if (k != nwc || nqc > nq) { set_error("DPLR charge layout mismatch"); return DPMD_E_STATE; }
  DPMD_TRY(sys_ensure(s, (void**)&s->d_qxyz, sizeof(double) * 3 * nq));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_q, sizeof(double) * nq));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_qf, sizeof(double) * 3 * nq));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_lr, sizeof(double) * kLrStride * (size_t)s->nseg));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_wc_g, sizeof(float) * 3 * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_partial_wc, sizeof(double) * 9 * (size_t)std::max(s->nblocks, 1)));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_atom_q, sizeof(int) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_wc_q, sizeof(int) * wc_q.size()));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_wc_seg, sizeof(int) * wc_seg.size()));
  const hipStream_t st = s->stream;
  DPMD_HIP(hipMemcpyAsync(s->d_q, q.data(), sizeof(double) * nqc, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemcpyAsync(s->d_atom_q, atom_q.data(), sizeof(int) * NA, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemcpyAsync(s->d_wc_q, wc_q.data(), sizeof(int) * wc_q.size(), hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemcpyAsync(s->d_wc_seg, wc_seg.data(), sizeof(int) * wc_seg.size(), hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemsetAsync(s->d_lr, 0, sizeof(double) * kLrStride * (size_t)s->nseg, st));
  if (!s->pppm) s->pppm = new Pppm();
  for (int d = 0; d < 3; ++d) s->dplr_grid[d] = grid[d];
  s->dplr_beta = beta;
  DPMD_TRY(s->pppm->setup(s->dplr_grid, beta));
  DPMD_HIP(hipStreamSynchronize(st));
  s->wc_model = W;
  return DPMD_OK;
}

int32_t dpmd::atomic_setup(System* s, Model* W) {
  const int NA = s->NA;
  std::vector<int> wc_atom, wc_row, atom_wc(NA, -1);
  for (int r = 0; r < s->nseg; ++r) {
    const SegInfo& S = s->seg[r];
    for (int c = 0; c < S.natoms; ++c) {
      const int g = S.base + c;
      const int t = s->type_sorted[g];
      if (!(W->nsel_mask >> t & 1)) continue;
      atom_wc[g] = (int)wc_atom.size();
      wc_atom.push_back(g);
      wc_row.push_back(s->rowbase[t] + S.row[t] + (c - S.off[t]));
    }
  }
  const int nwc = (int)wc_atom.size();
  // output k -> rank among the selected atoms in caller order (segment-major)
  std::vector<int> order(nwc);
  for (int k = 0; k < nwc; ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return s->perm[wc_atom[a]] < s->perm[wc_atom[b]]; });
  s->wc_perm.assign(nwc, 0);
  for (int r = 0; r < nwc; ++r) s->wc_perm[order[r]] = r;
  const size_t nn = std::max(nwc, 1);
  DPMD_TRY(sys_ensure(s, (void**)&s->d_wc_atom, sizeof(int) * nn));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_wc_row, sizeof(int) * nn));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_atom_wc, sizeof(int) * NA));
  const size_t fitbytes = sizeof(float) * 10 * (size_t)std::max(s->mrows, 1) * W->fit_k;
  DPMD_TRY(sys_ensure(s, (void**)&s->d_wc_fit, fitbytes));
  const hipStream_t st = s->stream;
  DPMD_HIP(hipMemsetAsync(s->d_wc_fit, 0, fitbytes, st));
  if (nwc > 0) {
    DPMD_HIP(hipMemcpyAsync(s->d_wc_atom, wc_atom.data(), sizeof(int) * nwc, hipMemcpyHostToDevice, st));
    DPMD_HIP(hipMemcpyAsync(s->d_wc_row, wc_row.data(), sizeof(int) * nwc, hipMemcpyHostToDevice, st));
  }
  DPMD_HIP(hipMemcpyAsync(s->d_atom_wc, atom_wc.data(), sizeof(int) * NA, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipStreamSynchronize(st));
  s->n_wc = nwc;
  s->atomic_maps_of = W;
  return DPMD_OK;
}

extern "C" int32_t dpmd_atomic_predict(dpmd_system_t sys, double* out) {
  SYS(sys);
  Model* M = s->model;
  if (!out) { set_error("null argument"); return DPMD_E_ARG; }
  if (!M->atomic) { set_error("the system's model is not an atomic model"); return DPMD_E_STATE; }
  if (!s->nl_ready) { set_error("no neighbour list: call dpmd_nlist_allocate first"); return DPMD_E_STATE; }
  if (s->wc_model) { set_error("an atomic model's system carries no DPLR term"); return DPMD_E_STATE; }
  if (s->atomic_maps_of != M) DPMD_TRY(atomic_setup(s, M));
  DPMD_TRY(sys_upload_segments(s));
  DPMD_TRY(md_write_pos32(s));
  DPMD_TRY(embed_launch(s, false, DPMD_PREC_FP32));
  AtomicFitBufs b;
  DPMD_TRY(fit_atomic_buffers(s, M, b));
  DPMD_TRY(fit_atomic_fwd(s, b));
  const int nwc = s->n_wc, ncomp = M->atomic == 1 ? 3 : 9;
  DPMD_TRY(sys_ensure(s, (void**)&s->d_qf, sizeof(double) * ncomp * (size_t)std::max(nwc, 1)));
  if (nwc > 0)
    hipLaunchKernelGGL(atomic_out_kernel, dim3((nwc + 127) / 128), dim3(128), 0, s->stream, s->d_T, M->d_tbias, b.Y3,
                       b.FK, M->wout, (float)M->out_norm, s->d_wc_atom, s->d_wc_row, nwc, M->atomic, s->d_qf);
  DPMD_HIP(hipGetLastError());
  std::vector<double> h((size_t)ncomp * std::max(nwc, 1));
  DPMD_HIP(hipMemcpyAsync(h.data(), s->d_qf, sizeof(double) * ncomp * nwc, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  for (int k = 0; k < nwc; ++k)
    for (int e = 0; e < ncomp; ++e) out[(size_t)s->wc_perm[k] * ncomp + e] = h[(size_t)k * ncomp + e];
  s->force_valid = false;
  return DPMD_OK;
}

extern "C" int32_t dpmd_dplr_get(dpmd_system_t sys, double* e_lr, double* centroids) {
  SYS(sys);
  if (!s->wc_model) { set_error("DPLR is not set on this system"); return DPMD_E_STATE; }
  const int NA = s->NA, nwc = s->n_wc, nq = NA + nwc;
  std::vector<double> lr((size_t)kLrStride * s->nseg);
  if (e_lr)
    DPMD_HIP(hipMemcpyAsync(lr.data(), s->d_lr, sizeof(double) * lr.size(), hipMemcpyDeviceToHost, s->stream));
  std::vector<double> c((size_t)3 * nq);
  std::vector<int> wq(std::max(nwc, 1));
  if (centroids && nwc > 0) {
    DPMD_HIP(hipMemcpyAsync(c.data(), s->d_qxyz, sizeof(double) * 3 * nq, hipMemcpyDeviceToHost, s->stream));
    DPMD_HIP(hipMemcpyAsync(wq.data(), s->d_wc_q, sizeof(int) * nwc, hipMemcpyDeviceToHost, s->stream));
  }
  DPMD_HIP(hipStreamSynchronize(s->stream));
  if (e_lr)
    for (int r = 0; r < s->nseg; ++r) e_lr[r] = lr[(size_t)r * kLrStride];
  if (centroids)
    for (int k = 0; k < nwc; ++k)
      for (int d = 0; d < 3; ++d) centroids[(size_t)s->wc_perm[k] * 3 + d] = c[(size_t)qdec(wq[k]) * 3 + d];
  return DPMD_OK;
}

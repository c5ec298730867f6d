// C ABI of libdpmd_hip.so (declared in include/dpmd.h): model upload, system state,
// neighbour-list protocol, energy/force/virial evaluation and the device MD loop.
#include "internal.h"
#include <cstdlib>
#include <math.h>
#include <cmath>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <numeric>
#include <string>

namespace dpmd {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}


void timer_begin(System* s, int which, hipEvent_t* ev) {
  *ev = nullptr;
  if (!s->timer.on) return;
  hipEventCreate(ev);
  hipEventRecord(*ev, s->stream);
}

void timer_end(System* s, int which, hipEvent_t ev) {
  if (!s->timer.on || !ev) return;
  hipEvent_t e2;
  hipEventCreate(&e2);
  hipEventRecord(e2, s->stream);
  s->timer.pending[which].push_back({ev, e2});
}

static void timer_harvest(System* s) {
  for (int w = 0; w < 8; ++w) {
    for (auto& pr : s->timer.pending[w]) {
      hipEventSynchronize(pr.second);
      float ms = 0.f;
      hipEventElapsedTime(&ms, pr.first, pr.second);
      s->timer.ms[w] += ms;
      s->timer.count[w] += 1;
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
    s->timer.pending[w].clear();
  }
}

System::~System() {
  for (auto& c : caps)
    if (*c.first) hipFree(*c.first);
  if (d_nm) hipFree(d_nm);
  for (hipEvent_t e : ev_frame) hipEventDestroy(e);
  for (hipEvent_t e : ev_copied) hipEventDestroy(e);
  if (h_traj) hipHostFree(h_traj);
  delete pppm;
  if (copy_stream) hipStreamDestroy(copy_stream);
  if (stream) hipStreamDestroy(stream);
}

// Pinned staging for nsteps recorded frames and one pair of events per frame (grow-only).
static int32_t traj_staging(System* s, int nsteps) {
  const size_t bytes = sizeof(double) * (size_t)nsteps * s->NA * 6;
  if (!s->copy_stream) DPMD_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
  if (s->h_traj_bytes < bytes) {
    if (s->h_traj) DPMD_HIP(hipHostFree(s->h_traj));
    s->h_traj = nullptr;
    s->h_traj_bytes = 0;
    DPMD_HIP(hipHostMalloc((void**)&s->h_traj, bytes, hipHostMallocDefault));
    s->h_traj_bytes = bytes;
  }
  while ((int)s->ev_frame.size() < nsteps) {
    hipEvent_t a, b;
    DPMD_HIP(hipEventCreateWithFlags(&a, hipEventDisableTiming));
    s->ev_frame.push_back(a);
    DPMD_HIP(hipEventCreateWithFlags(&b, hipEventDisableTiming));
    s->ev_copied.push_back(b);
  }
  return DPMD_OK;
}

// ------------------------------------------------------------------------------
// model

enum { FK0T = 0, FK1T, FK2T, FK0, FK1, FK2, FB0, FB1, FB2, FDT1, FDT2, FK3, FB3, FNB };

static uint16_t bf16_bits(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static uint16_t half_bits(float f) {
  _Float16 h = (_Float16)f;
  uint16_t u;
  memcpy(&u, &h, 2);
  return u;
}

static int32_t model_create(const dpmd_model_desc* d, int device, Model** out) {
  if (!d || !out) { set_error("null argument"); return DPMD_E_ARG; }
  if (d->ntypes < 1 || d->ntypes > kMaxT) { set_error("ntypes %d out of range [1,%d]", d->ntypes, kMaxT); return DPMD_E_ARG; }
  Model* M = new Model();
  M->device = device;
  M->ntypes = d->ntypes;
  if (!M->lay.init(d->embed_widths)) {
    set_error("embedding widths (%d,%d,%d) unsupported: need w0<=32 and each width equal to or twice the previous",
              d->embed_widths[0], d->embed_widths[1], d->embed_widths[2]);
    delete M;
    return DPMD_E_UNSUPPORTED;
  }
  for (int l = 0; l < 3; ++l) M->fitw[l] = d->fit_widths[l];
  if (!(d->fit_widths[0] == d->fit_widths[1] && d->fit_widths[1] == d->fit_widths[2])) {
    set_error("fitting widths must be three equal layers");
    delete M;
    return DPMD_E_UNSUPPORTED;
  }
  M->axis = d->axis;
  M->wout = d->embed_widths[2];
  if (2 * M->axis > M->wout) { set_error("axis %d needs 2*axis <= %d", M->axis, M->wout); delete M; return DPMD_E_ARG; }
  M->rcut = d->rcut;
  M->nnbrs = d->nnbrs;
  M->out_norm = d->out_norm;
  for (int t = 0; t < d->ntypes; ++t) {
    if (!(d->sr_std[t] > 0)) { set_error("sr_std[%d] must be positive", t); delete M; return DPMD_E_ARG; }
    M->mean[t] = d->sr_mean[t];
    M->istd[t] = 1.0 / d->sr_std[t];
    M->ebias[t] = d->ebias[t];
  }
  DPMD_HIP(hipSetDevice(device));
  // embedding nets
  M->off = embed_net_offsets(M->lay);
  const int nn = d->ntypes * d->ntypes;
  std::vector<float> emb((size_t)nn * M->off.total, 0.f);
  for (int k = 0; k < nn; ++k) {
    bool have = true;
    for (int l = 0; l < 3; ++l) have = have && d->emb_kernel[k][l] && d->emb_bias[k][l];
    if (!have) {
      if (d->tab_ngrids > 0) continue;  // tabulated nets need no MLP weights
      set_error("embedding net %d weights missing", k);
      delete M;
      return DPMD_E_ARG;
    }
    embed_pack_net(M->lay, d->emb_kernel[k][0], d->emb_bias[k][0], d->emb_kernel[k][1], d->emb_bias[k][1],
                   d->emb_kernel[k][2], d->emb_bias[k][2], emb.data() + (size_t)k * M->off.total);
  }
  if (d->tab_ngrids < 0 || d->tab_ngrids > (1 << 20)) {
    set_error("tab_ngrids %d out of range [0, 2^20]", d->tab_ngrids);
    delete M;
    return DPMD_E_ARG;
  }
  if (d->tab_ngrids > 0) {
    // delta in float32 as the reference forms it: (srmax - srmin) / Ngrids
    M->tab_n = d->tab_ngrids;
    M->tab_stride = d->tab_ngrids * 6 * 16 * M->lay.tiles[2];
    if ((size_t)nn * M->tab_stride * sizeof(float) >= (size_t)1 << 31) {
      set_error("tabulated embedding tables must stay below 2 GiB (%d intervals)", d->tab_ngrids);
      delete M;
      return DPMD_E_UNSUPPORTED;
    }
    std::vector<float> tab((size_t)nn * M->tab_stride);
    for (int k = 0; k < nn; ++k) {
      if (!d->tab_coeff[k]) { set_error("tabulated embedding net %d missing", k); delete M; return DPMD_E_ARG; }
      const float lo = d->tab_srmin[k], hi = d->tab_srmax[k];
      const float delta = (hi - lo) / (float)d->tab_ngrids;
      if (!(delta > 0.f)) { set_error("tabulated net %d: srmax must exceed srmin", k); delete M; return DPMD_E_ARG; }
      M->tab_srmin[k] = lo;
      M->tab_delta[k] = delta;
      embed_pack_table(M->lay, d->tab_ngrids, d->tab_coeff[k], tab.data() + (size_t)k * M->tab_stride);
    }
    DPMD_HIP(hipMalloc(&M->d_tab, sizeof(float) * tab.size()));
    DPMD_HIP(hipMemcpy(M->d_tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
  }
  DPMD_HIP(hipMalloc(&M->d_emb, sizeof(float) * emb.size()));
  DPMD_HIP(hipMemcpy(M->d_emb, emb.data(), sizeof(float) * emb.size(), hipMemcpyHostToDevice));
  M->offh = embed_net_offsets_h(M->lay);
  for (int h = 0; h < 2; ++h) {
    std::vector<uint16_t> tab((size_t)nn * M->offh.total);
    for (int k = 0; k < nn; ++k)
      embed_pack_net_h(M->lay, emb.data() + (size_t)k * M->off.total, tab.data() + (size_t)k * M->offh.total, h == 1);
    DPMD_HIP(hipMalloc(&M->d_emb_h[h], sizeof(uint16_t) * tab.size()));
    DPMD_HIP(hipMemcpy(M->d_emb_h[h], tab.data(), sizeof(uint16_t) * tab.size(), hipMemcpyHostToDevice));
  }
  DPMD_HIP(hipMalloc(&M->d_tbias, sizeof(float) * M->wout));
  if (d->tbias) DPMD_HIP(hipMemcpy(M->d_tbias, d->tbias, sizeof(float) * M->wout, hipMemcpyHostToDevice));
  else DPMD_HIP(hipMemset(M->d_tbias, 0, sizeof(float) * M->wout));
  // fitting nets: output width FN (multiple of 16), K / row stride FKS (multiple of 32)
  const int fw = d->fit_widths[0];
  const int FN = (fw + 15) / 16 * 16;
  const int FKS = (fw + 31) / 32 * 32;
  const int D = M->axis * M->wout;
  if (D % 32) { set_error("axis*width %d must be a multiple of 32", D); delete M; return DPMD_E_UNSUPPORTED; }
  const size_t sizes[FNB] = {(size_t)FN * D, (size_t)FN * FKS, (size_t)FN * FKS, (size_t)D * FKS, (size_t)FN * FKS,
                             (size_t)FN * FKS, (size_t)FN, (size_t)FN, (size_t)FN, (size_t)FN, (size_t)FN, (size_t)FN, 1};
  size_t pos = 0;
  for (int t = 0; t < d->ntypes; ++t)
    for (int b = 0; b < FNB; ++b) {
      M->fit_off[t][b] = pos;
      pos += (sizes[b] + 63) / 64 * 64;
    }
  M->fit_total = pos;
  std::vector<float> fit(pos, 0.f);
  if (d->atomic < 0 || d->atomic > 2) { set_error("atomic must be 0, 1 or 2"); delete M; return DPMD_E_ARG; }
  M->atomic = d->atomic;
  M->nsel_mask = d->nsel_mask;
  if (M->atomic && (d->fit_widths[2] != M->wout || d->fit_widths[0] != d->fit_widths[2])) {
    set_error("atomic models need fit_widths[2] == embed_widths[2] (and equal hidden widths)");
    delete M;
    return DPMD_E_UNSUPPORTED;
  }
  const std::vector<float> no_head(fw, 0.f);  // atomic models have no energy head
  for (int t = 0; t < d->ntypes; ++t) {
    for (int l = 0; l < (M->atomic ? 3 : 4); ++l)
      if (!d->fit_kernel[t][l]) { set_error("fitting net %d layer %d missing", t, l); delete M; return DPMD_E_ARG; }
    const float* K0 = d->fit_kernel[t][0];
    const float* K1 = d->fit_kernel[t][1];
    const float* K2 = d->fit_kernel[t][2];
    const float* K3 = M->atomic ? no_head.data() : d->fit_kernel[t][3];
    float* o[FNB];
    for (int b = 0; b < FNB; ++b) o[b] = fit.data() + M->fit_off[t][b];
    for (int k = 0; k < D; ++k)
      for (int n = 0; n < fw; ++n) {
        o[FK0T][(size_t)n * D + k] = K0[(size_t)k * fw + n];   // Bt[n][k] = K0[k][n]
        o[FK0][(size_t)k * FKS + n] = K0[(size_t)k * fw + n];  // Bt[k_in][n_out] for dG = dz1 K0^T
      }
    for (int k = 0; k < fw; ++k)
      for (int n = 0; n < fw; ++n) {
        o[FK1T][n * FKS + k] = K1[k * fw + n];
        o[FK2T][n * FKS + k] = K2[k * fw + n];
        o[FK1][k * FKS + n] = K1[k * fw + n];
        o[FK2][k * FKS + n] = K2[k * fw + n];
      }
    for (int n = 0; n < fw; ++n) {
      o[FB0][n] = d->fit_bias[t][0] ? d->fit_bias[t][0][n] : 0.f;
      o[FB1][n] = d->fit_bias[t][1] ? d->fit_bias[t][1][n] : 0.f;
      o[FB2][n] = d->fit_bias[t][2] ? d->fit_bias[t][2][n] : 0.f;
      o[FDT1][n] = d->fit_dt[t][1] ? d->fit_dt[t][1][n] : 1.f;
      o[FDT2][n] = d->fit_dt[t][2] ? d->fit_dt[t][2][n] : 1.f;
      o[FK3][n] = (float)(K3[n] * d->out_norm);
    }
    const float b3 = !M->atomic && d->fit_bias[t][3] ? d->fit_bias[t][3][0] : 0.f;
    o[FB3][0] = (float)(b3 * d->out_norm);
    M->fit_b3[t] = (float)(b3 * d->out_norm);
  }
  M->fit_n = FN;
  M->fit_k = FKS;
  DPMD_HIP(hipMalloc(&M->d_fit, sizeof(float) * fit.size()));
  DPMD_HIP(hipMemcpy(M->d_fit, fit.data(), sizeof(float) * fit.size(), hipMemcpyHostToDevice));
  // BF16 and FP16 copies, and the BF16 low parts bf16(w - bf16(w)) of the split form
  for (int h = 0; h < 3; ++h) {
    std::vector<uint16_t> fh(fit.size());
    for (size_t i = 0; i < fit.size(); ++i) {
      if (h == 2) {
        const uint16_t hi = bf16_bits(fit[i]);
        const uint32_t u = (uint32_t)hi << 16;
        float fhi;
        memcpy(&fhi, &u, 4);
        fh[i] = bf16_bits(fit[i] - fhi);
      } else {
        fh[i] = h ? half_bits(fit[i]) : bf16_bits(fit[i]);
      }
    }
    DPMD_HIP(hipMalloc(&M->d_fit_h[h], sizeof(uint16_t) * fh.size()));
    DPMD_HIP(hipMemcpy(M->d_fit_h[h], fh.data(), sizeof(uint16_t) * fh.size(), hipMemcpyHostToDevice));
  }
  *out = M;
  return DPMD_OK;
}

// ------------------------------------------------------------------------------
// system: segments, layout, grow-only buffers

int32_t sys_ensure(System* s, void** p, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  for (auto& c : s->caps) {
    if (c.first != p) continue;
    if (*p && c.second >= bytes) return DPMD_OK;
    if (*p) DPMD_HIP(hipFree(*p));
    *p = nullptr;
    const size_t nb = bytes + bytes / 4;  // regrowth keeps headroom for the next layout
    DPMD_HIP(hipMalloc(p, nb));
    c.second = nb;
    return DPMD_OK;
  }
  DPMD_HIP(hipMalloc(p, bytes));
  s->caps.push_back({p, bytes});
  return DPMD_OK;
}

// The stream is idle whenever host code changes a segment (every entry point ends with a
// stream synchronisation), so a copy followed by a synchronisation is race-free.
int32_t sys_upload_segments(System* s) {
  if (!s->seg_dirty) return DPMD_OK;
  DPMD_HIP(hipMemcpyAsync(s->d_seg, s->seg.data(), sizeof(SegInfo) * s->nseg, hipMemcpyHostToDevice, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  s->seg_dirty = false;
  return DPMD_OK;
}

static bool same_types(const int32_t* a, const int32_t* b, int n) {
  for (int i = 0; i < n; ++i)
    if (a[i] != b[i]) return false;
  return true;
}

// Lay out nseg segments (seg_natoms atoms each, caller order concatenated in type_idx):
// stable type sort inside each segment, fitting rows grouped by type over all segments,
// embedding blocks never straddling segments.  Buffers are grown, never shrunk.
static int32_t sys_layout(System* s, int nseg, const int32_t* seg_natoms, const int32_t* type_idx) {
  Model* M = s->model;
  if (nseg < 1 || !seg_natoms || !type_idx) { set_error("need nseg >= 1, segment sizes and type indices"); return DPMD_E_ARG; }
  int64_t tot = 0;
  for (int r = 0; r < nseg; ++r) {
    if (seg_natoms[r] < 1) { set_error("segment %d has %d atoms", r, seg_natoms[r]); return DPMD_E_ARG; }
    tot += seg_natoms[r];
  }
  if (tot > (1 << 26)) { set_error("%lld atoms above the 2^26 limit of one system", (long long)tot); return DPMD_E_ARG; }
  const int NA = (int)tot;
  for (int i = 0; i < NA; ++i)
    if (type_idx[i] < 0 || type_idx[i] >= M->ntypes) {
      set_error("type index %d out of trained scope of model (ntypes=%d)", type_idx[i], M->ntypes);
      return DPMD_E_ARG;
    }
  DPMD_HIP(hipSetDevice(M->device));
  if (s->stream) DPMD_HIP(hipStreamSynchronize(s->stream));
  s->nseg = nseg;
  s->NA = NA;
  s->ntypes = M->ntypes;
  s->seg.assign(nseg, SegInfo());
  s->perm.resize(NA);
  s->type_sorted.resize(NA);
  s->uniform = true;
  for (int t = 0; t < kMaxT; ++t) s->count[t] = 0;
  int base = 0, blk = 0;
  std::vector<int> atom_seg(NA);
  s->blk_seg.clear();
  for (int r = 0; r < nseg; ++r) {
    SegInfo& S = s->seg[r];
    memset(&S, 0, sizeof(S));
    const int n = seg_natoms[r];
    const int32_t* ty = type_idx + base;
    if (r > 0 && !(n == seg_natoms[0] && same_types(ty, type_idx, n))) s->uniform = false;
    S.base = base;
    S.natoms = n;
    int cnt[kMaxT] = {0, 0, 0, 0};
    for (int i = 0; i < n; ++i) cnt[ty[i]]++;
    S.off[0] = 0;
    for (int t = 0; t < s->ntypes; ++t) S.off[t + 1] = S.off[t] + cnt[t];
    for (int t = s->ntypes; t < kMaxT; ++t) S.off[t + 1] = S.off[t];
    for (int t = 0; t < s->ntypes; ++t) {
      S.row[t] = s->count[t];
      s->count[t] += cnt[t];
    }
    std::vector<int> lp(n);
    std::iota(lp.begin(), lp.end(), 0);
    std::stable_sort(lp.begin(), lp.end(), [&](int x, int y) { return ty[x] < ty[y]; });
    for (int i = 0; i < n; ++i) {
      s->perm[base + i] = base + lp[i];
      s->type_sorted[base + i] = ty[lp[i]];
      atom_seg[base + i] = r;
    }
    S.blk0 = blk;
    S.nblk = (n + 3) / 4;
    for (int b = 0; b < S.nblk; ++b) s->blk_seg.push_back(r);
    blk += S.nblk;
    S.nimg = 1;
    S.kimg[0] = S.kimg[1] = S.kimg[2] = 1;
    S.seed = (unsigned long long)r;
    S.step = 0;
    base += n;
  }
  s->nblocks = blk;
  int rb = 0;
  for (int t = 0; t < s->ntypes; ++t) {
    s->rowbase[t] = rb;
    rb += (s->count[t] + 63) / 64 * 64;
  }
  for (int t = s->ntypes; t <= kMaxT; ++t) s->rowbase[t] = rb;
  s->mrows = rb;
  s->fpad = M->fit_n;
  s->fstride = M->fit_k;
  s->mass.assign(NA, 1.0);
  const size_t W = M->wout, D = (size_t)M->axis * M->wout;
  const size_t mr = std::max(1, s->mrows);
  DPMD_TRY(sys_ensure(s, (void**)&s->d_seg, sizeof(SegInfo) * nseg));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_atom_seg, sizeof(int) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_blk_seg, sizeof(int) * s->nblocks));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_pos, sizeof(double) * NA * 3));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_mom, sizeof(double) * NA * 3));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_pos_save, sizeof(double) * NA * 3));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_mom_save, sizeof(double) * NA * 3));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_mass, sizeof(double) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_pos32, sizeof(float4) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_force, sizeof(float) * NA * 3));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_scratch, sizeof(double) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_flags, sizeof(int) * 32));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_thermo, sizeof(double) * 64 * nseg));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_T, sizeof(float) * (size_t)NA * 10 * W));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_G, sizeof(float) * mr * D));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_dG, sizeof(float) * mr * D));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_fitbuf, sizeof(float) * 7 * mr * s->fstride));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_eatom, sizeof(float) * mr));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_partial, sizeof(double) * 9 * s->nblocks));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_red, sizeof(double) * (32 * (size_t)nseg + 64)));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_pairs, sizeof(int) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_perm, sizeof(int) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_npt, sizeof(double) * 128));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_kepart, sizeof(double) * kKeBlocks * std::max<size_t>(3 * (size_t)nseg, 4)));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_mobile, sizeof(float) * NA));
  hipStream_t st = s->stream;
  DPMD_HIP(hipMemsetAsync(s->d_pos, 0, sizeof(double) * NA * 3, st));
  DPMD_HIP(hipMemsetAsync(s->d_mom, 0, sizeof(double) * NA * 3, st));
  DPMD_HIP(hipMemsetAsync(s->d_force, 0, sizeof(float) * NA * 3, st));
  DPMD_HIP(hipMemsetAsync(s->d_flags, 0, sizeof(int) * 32, st));
  DPMD_HIP(hipMemsetAsync(s->d_thermo, 0, sizeof(double) * 64 * nseg, st));
  DPMD_HIP(hipMemsetAsync(s->d_pairs, 0, sizeof(int) * NA, st));
  // padding rows of G and the K-padding columns of the fitting activations must read as zero
  DPMD_HIP(hipMemsetAsync(s->d_G, 0, sizeof(float) * mr * D, st));
  DPMD_HIP(hipMemsetAsync(s->d_fitbuf, 0, sizeof(float) * 7 * mr * s->fstride, st));
  DPMD_HIP(hipMemcpyAsync(s->d_atom_seg, atom_seg.data(), sizeof(int) * NA, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemcpyAsync(s->d_blk_seg, s->blk_seg.data(), sizeof(int) * s->nblocks, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemcpyAsync(s->d_mass, s->mass.data(), sizeof(double) * NA, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipMemcpyAsync(s->d_perm, s->perm.data(), sizeof(int) * NA, hipMemcpyHostToDevice, st));
  DPMD_HIP(hipStreamSynchronize(st));
  s->seg_dirty = true;
  s->nl_ready = false;
  s->nl_overflow = false;
  s->force_valid = false;
  s->rng_counter = 0;
  s->npt_ready = false;
  s->has_fixed = false;
  s->n_mobile = NA;
  s->wc_model = nullptr;  // DPLR / atomic-output maps belong to the old layout
  s->atomic_maps_of = nullptr;
  return DPMD_OK;
}

static int32_t sys_create(Model* M, int nseg, const int32_t* seg_natoms, const int32_t* type_idx, System** out) {
  if (!M || !out) { set_error("null argument"); return DPMD_E_ARG; }
  DPMD_HIP(hipSetDevice(M->device));
  System* s = new System();
  s->model = M;
  for (int t = 0; t < kMaxT; ++t) { s->knbr[t] = 0; s->cap[t] = 0; }
  // DPMD_FIT_PER_GEMM=1 keeps the fitting net's hidden layers on one GEMM launch each (the
  // fused kernel is the default); for comparing the two paths
  const char* pg = std::getenv("DPMD_FIT_PER_GEMM");
  s->fit_fused = !(pg && pg[0] == '1');
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("HIP error '%s' creating a stream", hipGetErrorString(e));
    delete s;
    return DPMD_E_HIP;
  }
  int32_t rc = sys_layout(s, nseg, seg_natoms, type_idx);
  if (rc != DPMD_OK) { delete s; return rc; }
  *out = s;
  return DPMD_OK;
}

static int32_t model_forces_impl(System* s, int precision) {
  hipEvent_t ev;
  DPMD_TRY(sys_upload_segments(s));
  s->force_valid = true;
  s->force_prec = precision;
  // deterministic forces: the backward stores pair gradients, a gather sums them per atom
  // (no DPLR: its long-range and centroid terms still add forces by atomics)
  const bool det = s->deterministic && !s->wc_model && s->d_pgrad;
  if (!det) DPMD_TRY(md_zero_force(s));
  timer_begin(s, 1, &ev);
  // embedding and fitting precisions of the form; the descriptor is stored in the format the
  // fitting reads (16-bit only for the plain 16-bit GEMMs; the split and FP32 ones read FP32)
  const int pe = prec_embed(precision), pf = prec_fit(precision);
  DPMD_TRY(embed_launch(s, false, pe, pf == DPMD_PREC_BF16 || pf == DPMD_PREC_FP16 ? pf : 0));
  timer_end(s, 1, ev);
  timer_begin(s, 2, &ev);
  DPMD_TRY(fit_launch(s, pf));
  timer_end(s, 2, ev);
  timer_begin(s, 3, &ev);
  DPMD_TRY(embed_launch(s, true, pe));
  if (det) DPMD_TRY(force_gather(s));
  timer_end(s, 3, ev);
  // DPLR long-range term (centroid model FP32 whatever the energy model's precision)
  if (s->wc_model) DPMD_TRY(dplr_forces(s));
  // fixed atoms: their coordinates carry no gradient (deepmd_jax/md.py:428-430)
  if (s->has_fixed) DPMD_TRY(md_mask_forces(s));
  return DPMD_OK;
}

static int32_t nl_read_flags(System* s, int* h) {
  DPMD_HIP(hipMemcpyAsync(h, s->d_flags, sizeof(int) * 16, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

// buffers of the deterministic force gather for the current capacities
static int32_t det_buffers(System* s) {
  size_t ksum = 0;
  for (int j = 0; j < s->ntypes; ++j) {
    ksum += s->cap[j];
    DPMD_TRY(sys_ensure(s, (void**)&s->d_rev[j], sizeof(int) * (size_t)s->NA * s->cap[j]));
  }
  DPMD_TRY(sys_ensure(s, (void**)&s->d_pgrad, sizeof(float) * (size_t)s->NA * ksum * 3));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_fcen, sizeof(float) * (size_t)s->NA * 3));
  return DPMD_OK;
}

static int32_t nl_allocate(System* s, double rc_skin, double ratio, const int32_t* kover, int32_t* kout) {
  if (!(rc_skin > 0)) { set_error("rc_skin must be positive"); return DPMD_E_ARG; }
  for (int r = 0; r < s->nseg; ++r) {
    const SegInfo& S = s->seg[r];
    if (!(S.boxd[0] > 0)) { set_error("box not set"); return DPMD_E_STATE; }
    for (int d = 0; d < 3; ++d)
      if (S.boxd[d] < 2.0 * rc_skin) {
        set_error("box length %g below 2*(rcut+skin)=%g: neighbour list needs the minimum-image regime",
                  S.boxd[d], 2 * rc_skin);
        return DPMD_E_ARG;
      }
  }
  s->rc_skin = rc_skin;
  s->buffer_ratio = ratio;
  s->nl_rebuilds = 0;
  for (SegInfo& S : s->seg)
    for (int d = 0; d < 3; ++d) {
      S.L[d] = (float)S.boxd[d];
      S.scale[d] = 1.f;
    }
  s->seg_dirty = true;
  DPMD_TRY(nlist_alloc_cells(s));
  int h[16];
  int kn[kMaxT];
  if (kover) {
    for (int j = 0; j < s->ntypes; ++j) kn[j] = kover[j];
  } else {
    for (;;) {
      DPMD_TRY(nlist_clear_flags(s));
      DPMD_TRY(nlist_build(s, true, false));
      DPMD_TRY(nl_read_flags(s, h));
      if (h[8] == 0) break;
      if (s->nl_lds_list >= 4096) { set_error("more than 4096 neighbours within rc_skin"); return DPMD_E_UNSUPPORTED; }
      s->nl_lds_list *= 2;
    }
    for (int j = 0; j < s->ntypes; ++j) {
      const int k = h[4 + j];
      if (k == 0) kn[j] = 1;
      else kn[j] = (int)(1.0 + (k * ratio + std::max(20 - k, 0) * std::max(ratio - 1.2, 0.0)));
    }
  }
  int ksum = 0;
  for (int j = 0; j < s->ntypes; ++j) {
    if (kn[j] < 1) { set_error("knbr must be >= 1"); return DPMD_E_ARG; }
    s->knbr[j] = kn[j];
    s->cap[j] = std::max(1, kn[j] - 1);
    ksum += kn[j];
    if (kout) kout[j] = kn[j];
    DPMD_TRY(sys_ensure(s, (void**)&s->d_nbr[j], sizeof(int) * (size_t)s->NA * s->cap[j]));
    DPMD_TRY(sys_ensure(s, (void**)&s->d_nbr_cnt[j], sizeof(int) * (size_t)s->NA));
  }
  if (s->deterministic) DPMD_TRY(det_buffers(s));
  while (s->nl_lds_list < ksum + 64) s->nl_lds_list *= 2;
  DPMD_TRY(nlist_clear_flags(s));
  DPMD_TRY(nlist_build(s, false, false));
  DPMD_TRY(nl_read_flags(s, h));
  s->nl_ready = true;
  s->nl_overflow = (h[2] & 2) != 0;
  s->force_valid = false;
  return DPMD_OK;
}

static void to_sorted(System* s, const double* in, std::vector<double>& out) {
  out.resize((size_t)s->NA * 3);
  for (int g = 0; g < s->NA; ++g)
    for (int d = 0; d < 3; ++d) out[(size_t)g * 3 + d] = in[(size_t)s->perm[g] * 3 + d];
}

static void from_sorted(System* s, const double* in, double* out) {
  for (int g = 0; g < s->NA; ++g)
    for (int d = 0; d < 3; ++d) out[(size_t)s->perm[g] * 3 + d] = in[(size_t)g * 3 + d];
}

static double seg_ebias(System* s, int r) {
  const Model* M = s->model;
  const SegInfo& S = s->seg[r];
  double eb = 0;
  for (int t = 0; t < s->ntypes; ++t) eb += M->ebias[t] * (S.off[t + 1] - S.off[t]) * M->out_norm;
  return eb / S.nimg;
}

static int32_t compute_impl(System* s, int precision, double* energy, double* force, double* virial, double* eatom) {
  if (!s->nl_ready) { set_error("no neighbour list: call dpmd_nlist_allocate first"); return DPMD_E_STATE; }
  if (s->nl_overflow) {
    // rows were truncated at the capacity: the evaluation would miss neighbours
    set_error("neighbour list overflowed its capacity: call dpmd_nlist_allocate again");
    return DPMD_E_STATE;
  }
  Model* M = s->model;
  DPMD_TRY(md_write_pos32(s));
  DPMD_TRY(model_forces_impl(s, precision));
  DPMD_TRY(md_reduce_launch(s));
  const size_t NA = s->NA;
  std::vector<double> red((size_t)16 * s->nseg);
  DPMD_HIP(hipMemcpyAsync(red.data(), s->d_red, sizeof(double) * red.size(), hipMemcpyDeviceToHost, s->stream));
  std::vector<float> f;
  if (force) {
    f.resize(NA * 3);
    DPMD_HIP(hipMemcpyAsync(f.data(), s->d_force, sizeof(float) * NA * 3, hipMemcpyDeviceToHost, s->stream));
  }
  std::vector<float> ea;
  if (eatom) {
    ea.resize(s->mrows);
    DPMD_HIP(hipMemcpyAsync(ea.data(), s->d_eatom, sizeof(float) * s->mrows, hipMemcpyDeviceToHost, s->stream));
  }
  DPMD_HIP(hipStreamSynchronize(s->stream));
  for (int r = 0; r < s->nseg; ++r) {
    if (energy) energy[r] = red[r * 16] + seg_ebias(s, r);
    if (virial)
      for (int k = 0; k < 9; ++k) virial[r * 9 + k] = red[r * 16 + 1 + k];
  }
  if (force) {
    std::vector<double> fd(NA * 3);
    for (size_t i = 0; i < NA * 3; ++i) fd[i] = f[i];
    from_sorted(s, fd.data(), force);
  }
  if (eatom) {
    for (int r = 0; r < s->nseg; ++r) {
      const SegInfo& S = s->seg[r];
      for (int i = 0; i < S.natoms; ++i) {
        const int g = S.base + i;
        const int t = s->type_sorted[g];
        const int row = s->rowbase[t] + S.row[t] + (i - S.off[t]);
        eatom[s->perm[g]] = ea[row] + M->ebias[t] * M->out_norm;
      }
    }
  }
  return DPMD_OK;
}

static int32_t set_seg_box(System* s, int r, const double* b) {
  for (int d = 0; d < 3; ++d)
    if (!(b[d] > 0)) { set_error("box lengths must be positive"); return DPMD_E_ARG; }
  SegInfo& S = s->seg[r];
  if (S.tri) {
    // the list was binned in the general cell's fractional coordinates
    S.tri = 0;
    s->nl_ready = false;
  }
  for (int d = 0; d < 3; ++d) {
    S.boxd[d] = b[d];
    S.box[d] = (float)b[d];
    if (!s->nl_ready) S.L[d] = (float)b[d];
    S.scale[d] = S.L[d] / (float)b[d];
    S.cellbox[d] = S.kimg[d] > 0 ? b[d] / S.kimg[d] : 0.0;
  }
  // the drift test follows the box through scale = L_ref / box (deepmd_jax/md.py:724-732)
  if (s->nl_ready) seg_set_threshold(S, s->rc_skin, s->model->rcut, false);
  s->seg_dirty = true;
  s->force_valid = false;
  return DPMD_OK;
}

// General cell H (rows: lattice vectors, deepmd_jax/utils.py:24-27).  A diagonal H is the
// orthorhombic box.  Otherwise the segment keeps H and its inverse, and box/boxd hold the
// perpendicular widths 1/|column d of H^-1|, which the minimum-image checks read.  A change
// of a general cell drops the neighbour list (it bins fractional coordinates of the cell).
static int32_t set_seg_cell(System* s, int r, const double* H) {
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(H[k])) { set_error("cell entries must be finite"); return DPMD_E_ARG; }
  if (H[1] == 0 && H[2] == 0 && H[3] == 0 && H[5] == 0 && H[6] == 0 && H[7] == 0) {
    const double b[3] = {H[0], H[4], H[8]};
    return set_seg_box(s, r, b);
  }
  const double det = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                     H[2] * (H[3] * H[7] - H[4] * H[6]);
  const double scale = std::max(std::fabs(H[0]), std::max(std::fabs(H[4]), std::fabs(H[8])));
  if (!(std::fabs(det) > 1e-12 * scale * scale * scale)) { set_error("cell is singular"); return DPMD_E_ARG; }
  double Hi[9];
  Hi[0] = (H[4] * H[8] - H[5] * H[7]) / det;
  Hi[1] = (H[2] * H[7] - H[1] * H[8]) / det;
  Hi[2] = (H[1] * H[5] - H[2] * H[4]) / det;
  Hi[3] = (H[5] * H[6] - H[3] * H[8]) / det;
  Hi[4] = (H[0] * H[8] - H[2] * H[6]) / det;
  Hi[5] = (H[2] * H[3] - H[0] * H[5]) / det;
  Hi[6] = (H[3] * H[7] - H[4] * H[6]) / det;
  Hi[7] = (H[1] * H[6] - H[0] * H[7]) / det;
  Hi[8] = (H[0] * H[4] - H[1] * H[3]) / det;
  SegInfo& S = s->seg[r];
  bool same = S.tri != 0;
  for (int k = 0; k < 9 && same; ++k) same = S.Hd[k] == H[k];
  if (!same) s->nl_ready = false;
  S.tri = 1;
  for (int k = 0; k < 9; ++k) {
    S.Hd[k] = H[k];
    S.Hid[k] = Hi[k];
    S.H[k] = (float)H[k];
    S.Hi[k] = (float)Hi[k];
  }
  for (int d = 0; d < 3; ++d) {
    const double w = 1.0 / std::sqrt(Hi[d] * Hi[d] + Hi[3 + d] * Hi[3 + d] + Hi[6 + d] * Hi[6 + d]);
    S.boxd[d] = w;
    S.box[d] = (float)w;
    if (!s->nl_ready) S.L[d] = (float)w;
    S.scale[d] = 1.f;
  }
  s->seg_dirty = true;
  s->force_valid = false;
  return DPMD_OK;
}

}  // namespace dpmd

using namespace dpmd;

#define SYS(sys)                                                        \
  System* s = reinterpret_cast<System*>(sys);                           \
  if (!s) { set_error("null system handle"); return DPMD_E_ARG; }       \
  DPMD_HIP(hipSetDevice(s->model->device))

extern "C" {

const char* dpmd_last_error(void) { return g_err; }
int32_t dpmd_version(void) { return 2; }

int32_t dpmd_model_create(const dpmd_model_desc* desc, int32_t device, dpmd_model_t* out) {
  Model* m = nullptr;
  int32_t rc = model_create(desc, device, &m);
  if (rc == DPMD_OK) *out = reinterpret_cast<dpmd_model_t>(m);
  return rc;
}

int32_t dpmd_model_destroy(dpmd_model_t model) {
  Model* M = reinterpret_cast<Model*>(model);
  if (!M) return DPMD_OK;
  if (M->d_emb) hipFree(M->d_emb);
  if (M->d_tbias) hipFree(M->d_tbias);
  if (M->d_fit) hipFree(M->d_fit);
  if (M->d_tab) hipFree(M->d_tab);
  for (int k = 0; k < 2; ++k)
    if (M->d_emb_h[k]) hipFree(M->d_emb_h[k]);
  for (int k = 0; k < 3; ++k)
    if (M->d_fit_h[k]) hipFree(M->d_fit_h[k]);
  delete M;
  return DPMD_OK;
}

int32_t dpmd_system_create(dpmd_model_t model, int32_t natoms, int32_t nreplica, const int32_t* type_idx,
                           dpmd_system_t* out) {
  if (!type_idx || !out) { set_error("null argument"); return DPMD_E_ARG; }
  if (natoms < 1 || nreplica < 1) { set_error("natoms and nreplica must be positive"); return DPMD_E_ARG; }
  std::vector<int32_t> sn(nreplica, natoms), ty((size_t)natoms * nreplica);
  for (int r = 0; r < nreplica; ++r) std::copy(type_idx, type_idx + natoms, ty.begin() + (size_t)r * natoms);
  System* s = nullptr;
  int32_t rc = sys_create(reinterpret_cast<Model*>(model), nreplica, sn.data(), ty.data(), &s);
  if (rc == DPMD_OK) *out = reinterpret_cast<dpmd_system_t>(s);
  return rc;
}

int32_t dpmd_system_create_ragged(dpmd_model_t model, int32_t nseg, const int32_t* seg_natoms,
                                  const int32_t* type_idx, dpmd_system_t* out) {
  System* s = nullptr;
  int32_t rc = sys_create(reinterpret_cast<Model*>(model), nseg, seg_natoms, type_idx, &s);
  if (rc == DPMD_OK) *out = reinterpret_cast<dpmd_system_t>(s);
  return rc;
}

int32_t dpmd_system_reshape(dpmd_system_t sys, int32_t nseg, const int32_t* seg_natoms, const int32_t* type_idx) {
  SYS(sys);
  return sys_layout(s, nseg, seg_natoms, type_idx);
}

int32_t dpmd_system_destroy(dpmd_system_t sys) {
  System* s = reinterpret_cast<System*>(sys);
  if (s) {
    hipSetDevice(s->model->device);
    if (s->stream) hipStreamSynchronize(s->stream);
    delete s;
  }
  return DPMD_OK;
}

int32_t dpmd_system_set_box(dpmd_system_t sys, const double* box3) {
  SYS(sys);
  if (!box3) { set_error("null argument"); return DPMD_E_ARG; }
  for (int r = 0; r < s->nseg; ++r) DPMD_TRY(set_seg_box(s, r, box3));
  return DPMD_OK;
}

int32_t dpmd_system_set_boxes(dpmd_system_t sys, const double* boxes) {
  SYS(sys);
  if (!boxes) { set_error("null argument"); return DPMD_E_ARG; }
  for (int r = 0; r < s->nseg; ++r) DPMD_TRY(set_seg_box(s, r, boxes + 3 * r));
  return DPMD_OK;
}

int32_t dpmd_system_set_cells(dpmd_system_t sys, const double* cells) {
  SYS(sys);
  if (!cells) { set_error("null argument"); return DPMD_E_ARG; }
  for (int r = 0; r < s->nseg; ++r) DPMD_TRY(set_seg_cell(s, r, cells + 9 * r));
  return DPMD_OK;
}

int32_t dpmd_system_set_images(dpmd_system_t sys, const int32_t* nimg) {
  SYS(sys);
  if (!nimg) { set_error("null argument"); return DPMD_E_ARG; }
  for (int r = 0; r < s->nseg; ++r) {
    SegInfo& S = s->seg[r];
    if (nimg[r] < 1) { set_error("segment %d: image count %d", r, nimg[r]); return DPMD_E_ARG; }
    for (int t = 0; t < s->ntypes; ++t)
      if ((S.off[t + 1] - S.off[t]) % nimg[r]) {
        set_error("segment %d: %d atoms of type %d are not %d whole images", r, S.off[t + 1] - S.off[t], t, nimg[r]);
        return DPMD_E_ARG;
      }
    S.nimg = nimg[r];
    for (int d = 0; d < 3; ++d) {
      S.kimg[d] = S.nimg == 1 ? 1 : 0;  // per-axis factors unknown
      S.cellbox[d] = S.nimg == 1 ? S.boxd[d] : 0.0;
    }
  }
  s->seg_dirty = true;
  s->force_valid = false;
  return DPMD_OK;
}

int32_t dpmd_system_set_image_grid(dpmd_system_t sys, const int32_t* k3) {
  SYS(sys);
  if (!k3) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<int32_t> n(s->nseg);
  for (int r = 0; r < s->nseg; ++r) {
    if (k3[3 * r] < 1 || k3[3 * r + 1] < 1 || k3[3 * r + 2] < 1) {
      set_error("segment %d: image factors must be >= 1", r);
      return DPMD_E_ARG;
    }
    n[r] = k3[3 * r] * k3[3 * r + 1] * k3[3 * r + 2];
  }
  DPMD_TRY(dpmd_system_set_images(sys, n.data()));
  for (int r = 0; r < s->nseg; ++r) {
    SegInfo& S = s->seg[r];
    for (int d = 0; d < 3; ++d) {
      S.kimg[d] = k3[3 * r + d];
      S.cellbox[d] = S.boxd[d] / S.kimg[d];
    }
  }
  return DPMD_OK;
}

int32_t dpmd_system_set_rng(dpmd_system_t sys, const uint64_t* seeds, const uint64_t* steps) {
  SYS(sys);
  for (int r = 0; r < s->nseg; ++r) {
    if (seeds) s->seg[r].seed = seeds[r];
    if (steps) s->seg[r].step = steps[r];
  }
  s->seg_dirty = true;
  return DPMD_OK;
}

int32_t dpmd_system_get_rng(dpmd_system_t sys, uint64_t* seeds, uint64_t* steps) {
  SYS(sys);
  for (int r = 0; r < s->nseg; ++r) {
    if (seeds) seeds[r] = s->seg[r].seed;
    if (steps) steps[r] = s->seg[r].step;
  }
  return DPMD_OK;
}

int32_t dpmd_system_set_positions(dpmd_system_t sys, const double* pos) {
  SYS(sys);
  if (!pos) { set_error("null argument"); return DPMD_E_ARG; }
  if (!(s->seg[0].boxd[0] > 0)) { set_error("box not set"); return DPMD_E_STATE; }
  std::vector<double> v;
  to_sorted(s, pos, v);
  DPMD_HIP(hipMemcpyAsync(s->d_pos, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, s->stream));
  DPMD_TRY(md_write_pos32(s));
  s->force_valid = false;
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_system_get_positions(dpmd_system_t sys, double* pos) {
  SYS(sys);
  if (!pos) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> v((size_t)s->NA * 3);
  DPMD_HIP(hipMemcpyAsync(v.data(), s->d_pos, sizeof(double) * v.size(), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  from_sorted(s, v.data(), pos);
  return DPMD_OK;
}

int32_t dpmd_system_set_velocities(dpmd_system_t sys, const double* vel) {
  SYS(sys);
  if (!vel) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> v;
  to_sorted(s, vel, v);
  for (int g = 0; g < s->NA; ++g)
    for (int d = 0; d < 3; ++d) v[(size_t)g * 3 + d] *= s->mass[g];
  DPMD_HIP(hipMemcpyAsync(s->d_mom, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_system_get_velocities(dpmd_system_t sys, double* vel) {
  SYS(sys);
  if (!vel) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> v((size_t)s->NA * 3);
  DPMD_HIP(hipMemcpyAsync(v.data(), s->d_mom, sizeof(double) * v.size(), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  for (int g = 0; g < s->NA; ++g)
    for (int d = 0; d < 3; ++d) v[(size_t)g * 3 + d] /= s->mass[g];
  from_sorted(s, v.data(), vel);
  return DPMD_OK;
}

int32_t dpmd_system_set_masses(dpmd_system_t sys, const double* mass) {
  SYS(sys);
  if (!mass) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> m(s->NA);
  for (int g = 0; g < s->NA; ++g) {
    m[g] = mass[s->perm[g]];
    if (!(m[g] > 0)) { set_error("masses must be positive"); return DPMD_E_ARG; }
  }
  s->mass = m;
  DPMD_HIP(hipMemcpyAsync(s->d_mass, s->mass.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_nlist_allocate(dpmd_system_t sys, double rc_skin, double buffer_ratio, const int32_t* knbr_override,
                            int32_t* knbr_out) {
  SYS(sys);
  return nl_allocate(s, rc_skin, buffer_ratio, knbr_override, knbr_out);
}

int32_t dpmd_nlist_update(dpmd_system_t sys, int32_t* overflow) {
  SYS(sys);
  if (!s->nl_ready) { set_error("neighbour list not allocated"); return DPMD_E_STATE; }
  DPMD_TRY(nlist_clear_flags(s));
  DPMD_TRY(nlist_build(s, false, false));
  int h[16];
  DPMD_TRY(nl_read_flags(s, h));
  s->nl_overflow = (h[2] & 2) != 0;
  s->force_valid = false;
  if (overflow) *overflow = s->nl_overflow ? 1 : 0;
  return DPMD_OK;
}

int32_t dpmd_nlist_export(dpmd_system_t sys, int32_t j, int32_t* out) {
  SYS(sys);
  if (!out) { set_error("null argument"); return DPMD_E_ARG; }
  if (!s->nl_ready) { set_error("neighbour list not allocated"); return DPMD_E_STATE; }
  if (j < 0 || j >= s->ntypes) { set_error("type %d out of range", j); return DPMD_E_ARG; }
  DPMD_HIP(hipMemcpyAsync(out, s->d_nbr[j], sizeof(int) * (size_t)s->NA * s->cap[j], hipMemcpyDeviceToHost,
                          s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_nlist_overflow(dpmd_system_t sys, int32_t* overflow) {
  SYS(sys);
  if (!overflow) { set_error("null argument"); return DPMD_E_ARG; }
  *overflow = s->nl_overflow ? 1 : 0;
  return DPMD_OK;
}

int32_t dpmd_nlist_capacity(dpmd_system_t sys, int32_t* knbr) {
  SYS(sys);
  if (!knbr) { set_error("null argument"); return DPMD_E_ARG; }
  for (int j = 0; j < s->ntypes; ++j) knbr[j] = s->knbr[j];
  return DPMD_OK;
}

int32_t dpmd_system_set_deterministic(dpmd_system_t sys, int32_t on) {
  SYS(sys);
  const bool want = on != 0;
  if (want == s->deterministic) return DPMD_OK;
  s->deterministic = want;
  s->force_valid = false;
  if (want && s->nl_ready) {
    DPMD_TRY(det_buffers(s));
    DPMD_TRY(nlist_reverse(s, false));
    DPMD_HIP(hipStreamSynchronize(s->stream));
  }
  return DPMD_OK;
}

int32_t dpmd_compute(dpmd_system_t sys, int32_t precision, double* energy, double* force, double* virial,
                     double* atom_energy) {
  SYS(sys);
  if (!prec_valid(precision)) {
    set_error("unknown precision %d", precision);
    return DPMD_E_ARG;
  }
  return compute_impl(s, precision, energy, force, virial, atom_energy);
}

// NPT parameters from dpmd_md_params; couple_axes rules of deepmd_jax/routine.py:196-223
static int32_t npt_args(System* s, const dpmd_md_params* p, NptArgs& a) {
  const bool pimd = p->routine == DPMD_PIMD_NPT;
  for (const SegInfo& S : s->seg)
    if (S.tri) { set_error("NPT runs orthorhombic boxes (deepmd_jax/routine.py:161-170)"); return DPMD_E_UNSUPPORTED; }
  if (!pimd && (s->nseg != 1 || s->seg[0].nimg != 1)) {
    set_error("NPT runs one system held without periodic images");
    return DPMD_E_UNSUPPORTED;
  }
  if (pimd) {
    if (s->nseg < 2 || !s->uniform || !s->d_nm) {
      set_error("path-integral NPT needs identical beads and dpmd_pimd_setup");
      return DPMD_E_STATE;
    }
    for (const SegInfo& S : s->seg)
      if (S.nimg != 1) { set_error("path-integral NPT runs beads held without periodic images"); return DPMD_E_UNSUPPORTED; }
  }
  if (!pimd && (p->chain_length_p < 1 || p->chain_length_p > 8)) {
    set_error("barostat chain length %d outside [1, 8]", p->chain_length_p);
    return DPMD_E_ARG;
  }
  if (!(p->tau_p > 0) || !(p->kT > 0)) {
    set_error("NPT needs tau_p > 0 and a temperature");
    return DPMD_E_ARG;
  }
  int G = 0;
  for (int d = 0; d < 3; ++d) {
    const int g = p->couple[d];
    if (g < -1 || g > 2) { set_error("couple[%d] = %d outside [-1, 2]", d, g); return DPMD_E_ARG; }
    a.gid[d] = g;
    if (g >= 0) {
      a.nper[g] += 1;
      G = std::max(G, g + 1);
    }
  }
  if (G == 0) { set_error("couple_axes must contain at least one group; use NVT instead"); return DPMD_E_ARG; }
  for (int g = 0; g < G; ++g)
    if (a.nper[g] == 0) { set_error("coupled groups must be numbered 0..G-1 without gaps"); return DPMD_E_ARG; }
  a.G = G;
  a.C = p->chain_length_p;
  a.pressure = p->pressure;
  a.kT = p->kT;
  a.dt = p->dt;
  a.tau_p = p->tau_p;
  a.rc = s->model->rcut;
  a.rc_skin = s->rc_skin;
  a.natoms = s->seg[0].natoms;
  a.cs = std::max(1, p->chain_steps_p);
  a.sy = std::max(1, p->sy_steps_p);
  return DPMD_OK;
}

// drift threshold with the NPT safety factor on the host copy of the segments
static void npt_threshold(System* s, const NptArgs& a) {
  for (SegInfo& S : s->seg) seg_set_threshold(S, a.rc_skin, a.rc, true);
  s->seg_dirty = true;
}

int32_t dpmd_md_run(dpmd_system_t sys, const dpmd_md_params* p, int32_t nsteps, double* traj_pos, double* traj_vel,
                    uint32_t* error_code) {
  SYS(sys);
  if (!p) { set_error("null argument"); return DPMD_E_ARG; }
  if (!s->nl_ready) { set_error("no neighbour list: call dpmd_nlist_allocate first"); return DPMD_E_STATE; }
  if (nsteps < 0) { set_error("nsteps must be >= 0"); return DPMD_E_ARG; }
  if (p->routine != DPMD_NVE && p->routine != DPMD_NVT_LANGEVIN && p->routine != DPMD_NVT_NH &&
      p->routine != DPMD_PIMD_PILE_L && p->routine != DPMD_NPT_NH && p->routine != DPMD_PIMD_NPT) {
    set_error("unknown routine %d", p->routine);
    return DPMD_E_ARG;
  }
  if (s->has_fixed && p->routine != DPMD_NVE && p->routine != DPMD_NVT_NH && p->routine != DPMD_NVT_LANGEVIN) {
    set_error("fixed atoms are supported with NVE, NVT and NVT_langevin (deepmd_jax/md.py:228-229)");
    return DPMD_E_UNSUPPORTED;
  }
  const bool npt = p->routine == DPMD_NPT_NH;
  const bool pnpt = p->routine == DPMD_PIMD_NPT;
  // Nose-Hoover chain substeps (0 reads as 1); Suzuki-Yoshida orders 1, 3, 5, 7
  const int cs_t = std::max(1, p->chain_steps), sy_t = std::max(1, p->sy_steps);
  for (int v : {p->sy_steps, p->sy_steps_p})
    if (v != 0 && v != 1 && v != 3 && v != 5 && v != 7) {
      set_error("sy_steps must be 1, 3, 5 or 7 (got %d)", v);
      return DPMD_E_ARG;
    }
  if (p->chain_steps < 0 || p->chain_steps_p < 0 || cs_t > 64 || p->chain_steps_p > 64) {
    set_error("chain_steps must lie in [1, 64]");
    return DPMD_E_ARG;
  }

  NptArgs na;
  memset(&na, 0, sizeof(na));
  if (npt || pnpt) DPMD_TRY(npt_args(s, p, na));
  if (!prec_valid(p->precision)) {
    set_error("unknown precision %d", p->precision);
    return DPMD_E_ARG;
  }
  if (p->routine == DPMD_PIMD_PILE_L && !s->d_nm) { set_error("PIMD normal modes not set"); return DPMD_E_STATE; }
  if (p->remove_com && (p->routine == DPMD_PIMD_PILE_L || pnpt || s->nseg != 1)) {
    set_error("per-step centre-of-mass removal runs one system without beads");
    return DPMD_E_UNSUPPORTED;
  }
  if (p->routine == DPMD_PIMD_PILE_L && !s->uniform) { set_error("PIMD needs identical beads"); return DPMD_E_STATE; }
  if (p->routine == DPMD_PIMD_PILE_L && s->seg[0].tri) {
    // the free-ring step unwraps every bead through bead 0's cell
    for (const SegInfo& S : s->seg) {
      bool same = S.tri != 0;
      for (int k = 0; k < 9 && same; ++k) same = S.Hd[k] == s->seg[0].Hd[k];
      if (!same) { set_error("PIMD beads must share one cell"); return DPMD_E_ARG; }
    }
  }
  const size_t NA = s->NA;
  const bool rec = traj_pos || traj_vel;
  if (rec) {
    DPMD_TRY(sys_ensure(s, (void**)&s->d_traj, sizeof(double) * (size_t)nsteps * NA * 6));
    DPMD_TRY(traj_staging(s, nsteps));
  }
  DPMD_HIP(hipMemcpyAsync(s->d_pos_save, s->d_pos, sizeof(double) * NA * 3, hipMemcpyDeviceToDevice, s->stream));
  DPMD_HIP(hipMemcpyAsync(s->d_mom_save, s->d_mom, sizeof(double) * NA * 3, hipMemcpyDeviceToDevice, s->stream));
  if (s->nl_overflow) {
    // the current list already lacks neighbours (deepmd_jax/md.py:637-641): report
    // bit 2 before stepping so the caller reallocates
    if (error_code) *error_code |= 2u;
    DPMD_HIP(hipStreamSynchronize(s->stream));
    return DPMD_OK;
  }
  std::vector<double> thermo_save((size_t)64 * s->nseg);
  DPMD_HIP(hipMemcpyAsync(thermo_save.data(), s->d_thermo, sizeof(double) * thermo_save.size(),
                          hipMemcpyDeviceToHost, s->stream));
  if (npt || pnpt) {
    // the drift threshold carries the NPT safety factor from the first step on
    npt_threshold(s, na);
    DPMD_HIP(hipMemcpyAsync(s->d_npt + 64, s->d_npt, sizeof(double) * 64, hipMemcpyDeviceToDevice, s->stream));
  }
  DPMD_TRY(sys_upload_segments(s));
  DPMD_TRY(nlist_clear_flags(s));
  const double dt = p->dt;
  const int prec = p->precision;
  // forces held from an evaluation at another precision are recomputed at the run's
  const bool fresh_forces = !s->force_valid || s->force_prec != prec;
  if (fresh_forces) {
    DPMD_TRY(md_write_pos32(s));
    DPMD_TRY(model_forces_impl(s, prec));
  }
  if ((npt || pnpt) && s->npt_ready && s->npt_kind != p->routine) s->npt_ready = false;
  if (npt && (!s->npt_ready || fresh_forces)) {
    // init_fn (deepmd_jax/routine.py:344-378), or dU/deps of recomputed forces
    DPMD_TRY(md_npt_init(s, na, s->npt_ready));
    s->npt_ready = true;
    s->npt_kind = DPMD_NPT_NH;
  }
  if (pnpt && (!s->npt_ready || fresh_forces)) {
    // init_fn (deepmd_jax/pimd.py:402-428): box at rest, then the centroid-virial dU/deps
    if (!s->npt_ready) DPMD_TRY(md_pnpt_init(s, na));
    DPMD_TRY(md_pnpt_dudv(s, na));
    s->npt_ready = true;
    s->npt_kind = DPMD_PIMD_NPT;
  }
  if (rec) {
    DPMD_TRY(sys_ensure(s, (void**)&s->d_box_traj, sizeof(double) * 3 * (size_t)s->nseg * std::max(nsteps, 1)));
    s->box_traj_steps = 0;
  }
  const uint64_t counter0 = s->rng_counter;
  for (int step = 0; step < nsteps; ++step) {
    DPMD_TRY(nlist_drift_check(s, npt || pnpt));
    DPMD_TRY(nlist_build(s, false, true));
    hipEvent_t ev;
    if (p->routine == DPMD_NVE) {
      timer_begin(s, 4, &ev);
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_drift(s, dt));
      timer_end(s, 4, ev);
      DPMD_TRY(model_forces_impl(s, prec));
      timer_begin(s, 4, &ev);
      DPMD_TRY(md_kick(s, 0.5 * dt));
      if (s->has_fixed) DPMD_TRY(md_mask_momenta(s));
      timer_end(s, 4, ev);
    } else if (p->routine == DPMD_NVT_LANGEVIN) {
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_drift(s, 0.5 * dt));
      DPMD_TRY(md_langevin_o(s, dt, p->gamma, p->kT, p->seed, (uint64_t)step));
      DPMD_TRY(md_drift(s, 0.5 * dt));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_kick(s, 0.5 * dt));
    } else if (p->routine == DPMD_NVT_NH) {
      DPMD_TRY(md_nh_half(s, p->kT, dt, p->tau, p->chain_length, cs_t, sy_t));
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_drift(s, dt));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_nh_half(s, p->kT, dt, p->tau, p->chain_length, cs_t, sy_t));
      if (s->has_fixed) DPMD_TRY(md_mask_momenta(s));
    } else if (npt) {
      // apply_fn of deepmd_jax/routine.py:396-423 around inner_step :365-394
      DPMD_TRY(md_npt_baro_half(s, na));
      DPMD_TRY(md_nh_half(s, p->kT, dt, p->tau, p->chain_length, cs_t, sy_t));
      DPMD_TRY(md_npt_box(s, na, 0));
      DPMD_TRY(md_npt_particles(s, dt, true));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_npt_particles(s, dt, false));
      DPMD_TRY(md_npt_box(s, na, 1));
      DPMD_TRY(md_nh_half(s, p->kT, dt, p->tau, p->chain_length, cs_t, sy_t));
      DPMD_TRY(md_npt_baro_half(s, na));
    } else if (pnpt) {
      // step_fn of deepmd_jax/pimd.py:430-516
      const int P = s->nseg;
      const double* C = s->d_nm;
      const double* fr = s->d_nm + P * P;
      const double* c1 = fr + P;
      const double* c2 = c1 + P;
      DPMD_TRY(md_pnpt_kick(s, na));
      DPMD_TRY(md_pnpt_il2(s, na));
      DPMD_TRY(md_pnpt_move(s, na, 0.5 * dt));
      DPMD_TRY(md_pnpt_drift_ring(s, na, C, fr, 0.5 * dt));
      DPMD_TRY(md_pimd_o(s, C, c1, c2, p->seed, counter0 + step));
      DPMD_TRY(md_pnpt_box_o(s, na, p->gamma, p->seed, counter0 + step));
      DPMD_TRY(md_pnpt_move(s, na, 0.5 * dt));
      DPMD_TRY(md_pnpt_drift_ring(s, na, C, fr, 0.5 * dt));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_pnpt_dudv(s, na));
      DPMD_TRY(md_pnpt_il2(s, na));
      DPMD_TRY(md_pnpt_kick(s, na));
    } else {
      const int P = s->nseg;
      const double* C = s->d_nm;
      const double* fr = s->d_nm + P * P;
      const double* c1 = fr + P;
      const double* c2 = c1 + P;
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_pimd_free_ring(s, C, fr, 0.5 * dt));
      DPMD_TRY(md_pimd_o(s, C, c1, c2, p->seed, counter0 + step));
      DPMD_TRY(md_pimd_free_ring(s, C, fr, 0.5 * dt));
      DPMD_TRY(md_write_pos32(s));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_kick(s, 0.5 * dt));
    }
    // per-step centre-of-mass removal over the mobile atoms (deepmd_jax/md.py:786-790)
    if (p->remove_com) DPMD_TRY(md_remove_com(s));
    if (rec) {
      // frame -> caller order on the device, then to pinned memory on the copy stream
      double* frame = s->d_traj + (size_t)step * NA * 6;
      DPMD_TRY(md_traj_record(s, frame, frame + NA * 3));
      DPMD_HIP(hipEventRecord(s->ev_frame[step], s->stream));
      DPMD_HIP(hipStreamWaitEvent(s->copy_stream, s->ev_frame[step], 0));
      DPMD_HIP(hipMemcpyAsync(s->h_traj + (size_t)step * NA * 6, frame, sizeof(double) * NA * 6,
                              hipMemcpyDeviceToHost, s->copy_stream));
      DPMD_HIP(hipEventRecord(s->ev_copied[step], s->copy_stream));
      DPMD_TRY(md_box_record(s, s->d_box_traj + (size_t)step * 3 * s->nseg));
    }
  }
  if (rec) {
    // copy each frame out of pinned memory while the device computes the later steps;
    // on an overflow the frames are discarded by the caller's re-run of the chunk.  The
    // destination pages are touched first (a fresh trajectory array is mapped lazily), so
    // the page faults fall while the first steps compute.
    const size_t n3 = (size_t)nsteps * NA * 3, stride = 4096 / sizeof(double);
    for (double* dst : {traj_pos, traj_vel}) {
      if (!dst || n3 == 0) continue;
      for (size_t i = 0; i < n3; i += stride) dst[i] = 0.0;
      dst[n3 - 1] = 0.0;
    }
    for (int k = 0; k < nsteps; ++k) {
      DPMD_HIP(hipEventSynchronize(s->ev_copied[k]));
      const double* hp = s->h_traj + (size_t)k * NA * 6;
      if (traj_pos) memcpy(traj_pos + (size_t)k * NA * 3, hp, sizeof(double) * NA * 3);
      if (traj_vel) memcpy(traj_vel + (size_t)k * NA * 3, hp + NA * 3, sizeof(double) * NA * 3);
    }
  }
  int h[16];
  DPMD_TRY(nl_read_flags(s, h));
  s->nl_rebuilds += h[9];
  uint32_t code = (uint32_t)(h[2] & (1 | 2 | 16));
  if (h[8]) code |= 2;
  if (error_code) *error_code |= code;
  if (code & 3) {
    // neighbour capacity exceeded, or (NPT) the box left the list's reference box behind,
    // during the chunk: restore the entry state so the caller can reallocate and re-run
    // (deepmd_jax/md.py:914-916)
    DPMD_HIP(hipMemcpyAsync(s->d_pos, s->d_pos_save, sizeof(double) * NA * 3, hipMemcpyDeviceToDevice, s->stream));
    DPMD_HIP(hipMemcpyAsync(s->d_mom, s->d_mom_save, sizeof(double) * NA * 3, hipMemcpyDeviceToDevice, s->stream));
    DPMD_HIP(hipMemcpyAsync(s->d_thermo, thermo_save.data(), sizeof(double) * thermo_save.size(),
                            hipMemcpyHostToDevice, s->stream));
    if (npt || pnpt) {
      // barostat state and the box of the chunk entry (the host copy of the segments)
      DPMD_HIP(hipMemcpyAsync(s->d_npt, s->d_npt + 64, sizeof(double) * 64, hipMemcpyDeviceToDevice, s->stream));
      s->seg_dirty = true;
      DPMD_TRY(sys_upload_segments(s));
    }
    DPMD_TRY(md_write_pos32(s));
    DPMD_HIP(hipStreamSynchronize(s->stream));
    s->force_valid = false;
    s->nl_overflow = true;
    return DPMD_OK;
  }
  if (npt || pnpt) {
    // the box moved on the device: the host copy follows before any upload
    DPMD_HIP(hipMemcpyAsync(s->seg.data(), s->d_seg, sizeof(SegInfo) * s->nseg, hipMemcpyDeviceToHost, s->stream));
    DPMD_HIP(hipStreamSynchronize(s->stream));
  }
  if (rec) s->box_traj_steps = nsteps;
  s->rng_counter = counter0 + nsteps;
  for (SegInfo& S : s->seg) S.step += nsteps;
  s->seg_dirty = true;
  return DPMD_OK;
}

int32_t dpmd_pimd_setup(dpmd_system_t sys, const double* C, const double* freqs, const double* c1, const double* c2) {
  SYS(sys);
  if (!C || !freqs || !c1 || !c2) { set_error("null argument"); return DPMD_E_ARG; }
  const int P = s->nseg;
  if (P < 2) { set_error("PIMD needs nreplica >= 2"); return DPMD_E_ARG; }
  if (P > 64) { set_error("PIMD supports at most 64 beads"); return DPMD_E_ARG; }
  if (!s->uniform) { set_error("PIMD needs identical beads"); return DPMD_E_STATE; }
  for (int k = 1; k < P; ++k)
    if (!(freqs[k] > 0)) { set_error("non-centroid normal-mode frequencies must be positive"); return DPMD_E_ARG; }
  std::vector<double> buf((size_t)P * P + 3 * P);
  std::copy(C, C + P * P, buf.begin());
  std::copy(freqs, freqs + P, buf.begin() + P * P);
  std::copy(c1, c1 + P, buf.begin() + P * P + P);
  std::copy(c2, c2 + P, buf.begin() + P * P + 2 * P);
  if (s->d_nm) DPMD_HIP(hipFree(s->d_nm));
  s->d_nm = nullptr;
  DPMD_HIP(hipMalloc(&s->d_nm, sizeof(double) * buf.size()));
  DPMD_HIP(hipMemcpy(s->d_nm, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
  return DPMD_OK;
}

int32_t dpmd_kinetic_energy(dpmd_system_t sys, double* ke) {
  SYS(sys);
  if (!ke) { set_error("null argument"); return DPMD_E_ARG; }
  DPMD_TRY(sys_upload_segments(s));
  double* d_out = s->d_red + 16 * s->nseg;
  DPMD_TRY(md_ke_launch(s, d_out));
  DPMD_HIP(hipMemcpyAsync(ke, d_out, sizeof(double) * s->nseg, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_thermostat_energy(dpmd_system_t sys, double* e) {
  SYS(sys);
  if (!e) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> t((size_t)64 * s->nseg);
  DPMD_HIP(hipMemcpyAsync(t.data(), s->d_thermo, sizeof(double) * t.size(), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  for (int r = 0; r < s->nseg; ++r) e[r] = t[(size_t)r * 64];
  if (s->npt_ready && s->npt_kind == DPMD_NPT_NH) {
    // NPT: barostat chain, P V and box kinetic energy join the invariant
    double b[64];
    DPMD_HIP(hipMemcpy(b, s->d_npt, sizeof(b), hipMemcpyDeviceToHost));
    e[0] += b[43];
  }
  return DPMD_OK;
}

int32_t dpmd_system_get_boxes(dpmd_system_t sys, double* boxes) {
  SYS(sys);
  if (!boxes) { set_error("null argument"); return DPMD_E_ARG; }
  for (int r = 0; r < s->nseg; ++r)
    for (int d = 0; d < 3; ++d) boxes[r * 3 + d] = s->seg[r].boxd[d];
  return DPMD_OK;
}

int32_t dpmd_md_box_trajectory(dpmd_system_t sys, double* boxes) {
  SYS(sys);
  if (!boxes) { set_error("null argument"); return DPMD_E_ARG; }
  if (s->box_traj_steps == 0) return DPMD_OK;
  DPMD_HIP(hipMemcpyAsync(boxes, s->d_box_traj, sizeof(double) * 3 * (size_t)s->nseg * s->box_traj_steps,
                          hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_npt_state_get(dpmd_system_t sys, double* state) {
  SYS(sys);
  if (!state) { set_error("null argument"); return DPMD_E_ARG; }
  if (!s->npt_ready) { set_error("no NPT state: run an NPT chunk first"); return DPMD_E_STATE; }
  DPMD_HIP(hipMemcpyAsync(state, s->d_npt, sizeof(double) * 64, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_thermostat_state_get(dpmd_system_t sys, double* state) {
  SYS(sys);
  if (!state) { set_error("null argument"); return DPMD_E_ARG; }
  DPMD_HIP(hipMemcpyAsync(state, s->d_thermo, sizeof(double) * 64 * s->nseg, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_thermostat_state_set(dpmd_system_t sys, const double* state) {
  SYS(sys);
  if (!state) { set_error("null argument"); return DPMD_E_ARG; }
  DPMD_HIP(hipMemcpyAsync(s->d_thermo, state, sizeof(double) * 64 * s->nseg, hipMemcpyHostToDevice, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_timing_enable(dpmd_system_t sys, int32_t on) {
  SYS(sys);
  timer_harvest(s);
  s->timer.on = on != 0;
  for (int w = 0; w < 8; ++w) { s->timer.ms[w] = 0; s->timer.count[w] = 0; }
  return DPMD_OK;
}

int32_t dpmd_timing_get(dpmd_system_t sys, int32_t which, double* ms, int64_t* count) {
  SYS(sys);
  if (which < 0 || which >= 8) { set_error("bad argument"); return DPMD_E_ARG; }
  timer_harvest(s);
  if (ms) *ms = s->timer.ms[which];
  if (count) *count = s->timer.count[which];
  return DPMD_OK;
}

int32_t dpmd_system_set_fixed(dpmd_system_t sys, const int32_t* fixed) {
  SYS(sys);
  if (fixed && s->nseg != 1) { set_error("fixed atoms need a system of one segment"); return DPMD_E_UNSUPPORTED; }
  const int NA = s->NA;
  std::vector<float> mob(NA, 1.f);
  int nm = NA;
  if (fixed) {
    nm = 0;
    for (int g = 0; g < NA; ++g) {
      mob[g] = fixed[s->perm[g]] ? 0.f : 1.f;
      nm += mob[g] != 0.f;
    }
  }
  if (nm == 0) { set_error("every atom is fixed"); return DPMD_E_ARG; }
  s->has_fixed = nm < NA;
  s->n_mobile = nm;
  DPMD_HIP(hipMemcpyAsync(s->d_mobile, mob.data(), sizeof(float) * NA, hipMemcpyHostToDevice, s->stream));
  if (s->has_fixed) DPMD_TRY(md_mask_momenta(s));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  s->force_valid = false;
  return DPMD_OK;
}

int32_t dpmd_nlist_update_rate(dpmd_system_t sys, double* rate) {
  SYS(sys);
  if (!rate) { set_error("null argument"); return DPMD_E_ARG; }
  float f = 0.f;
  DPMD_HIP(hipMemcpyAsync(&f, s->d_flags + 16, sizeof(float), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  *rate = f;
  return DPMD_OK;
}

int32_t dpmd_nlist_rebuilds(dpmd_system_t sys, int64_t* count) {
  SYS(sys);
  if (!count) { set_error("null argument"); return DPMD_E_ARG; }
  *count = s->nl_rebuilds;
  return DPMD_OK;
}

int32_t dpmd_pair_count(dpmd_system_t sys, double* pairs) {
  SYS(sys);
  if (!pairs) { set_error("null argument"); return DPMD_E_ARG; }
  const size_t NA = s->NA;
  std::vector<int> pc(NA);
  DPMD_HIP(hipMemcpyAsync(pc.data(), s->d_pairs, sizeof(int) * NA, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  double t = 0;
  for (size_t i = 0; i < NA; ++i) t += pc[i];
  *pairs = t / NA;
  return DPMD_OK;
}

int32_t dpmd_synchronize(dpmd_system_t sys) {
  SYS(sys);
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Allocation-free evaluation (dpmd_compute_noalloc).  The per-evaluation work buffers --
// T, G, dE/dG, the seven fitting activations, per-row energies, per-block virial partials,
// per-segment sums and the caller-order outputs -- live in one file-scope device array
// instead of the system's own buffers, so the call makes no device allocation.  Every GPU
// the library runs on holds its own copy of the array until the process ends; its address
// on a device comes from hipGetSymbolAddress (the array's name is a host-side stand-in,
// never a device address).  The array keeps the last call's contents: the regions that
// must read as zero (padding rows of G, K-padding columns of the activations) are cleared
// by a fill kernel on every call.  Calls that overlapped on one GPU would race on the
// array; a mutex serialises the calls of one process.

namespace dpmd {

constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kScratchDevices];  // per device, guarded by g_scratch_mu

// device address of `scratch` on the current device (set up on the first call there)
static int32_t scratch_device_ptr(int device, unsigned char** out) {
  if (device < 0 || device >= kScratchDevices) {
    set_error("device %d outside the %d devices the scratch table covers", device, kScratchDevices);
    return DPMD_E_ARG;
  }
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
      return DPMD_E_HIP;
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return DPMD_OK;
}

__global__ void scratch_zero_kernel(float4* x, size_t n4) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// forces to caller order in FP64; per-segment energy and virial out of the 16-double sums
__global__ void scratch_out_kernel(const float* force, const int* perm, int NA, const double* red, int nseg,
                                   double* fout, double* eout, double* vout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NA * 3) {
    const int g = i / 3, d = i - 3 * g;
    fout[(size_t)perm[g] * 3 + d] = (double)force[i];
  }
  if (i < nseg) {
    eout[i] = red[(size_t)i * 16];
    for (int k = 0; k < 9; ++k) vout[(size_t)i * 9 + k] = red[(size_t)i * 16 + 1 + k];
  }
}

// The system's work-buffer pointers, pointed into the scratch array for one call and
// restored on every exit path.
struct WorkBuffers {
  System* s;
  float* T;
  void* G;
  float* dG;
  float* fit;
  float* eatom;
  double* partial;
  double* red;
  explicit WorkBuffers(System* sys)
      : s(sys), T(sys->d_T), G(sys->d_G), dG(sys->d_dG), fit(sys->d_fitbuf), eatom(sys->d_eatom),
        partial(sys->d_partial), red(sys->d_red) {}
  ~WorkBuffers() {
    s->d_T = T;
    s->d_G = G;
    s->d_dG = dG;
    s->d_fitbuf = fit;
    s->d_eatom = eatom;
    s->d_partial = partial;
    s->d_red = red;
  }
};

}  // namespace dpmd

extern "C" int32_t dpmd_compute_noalloc(dpmd_system_t sys, int32_t precision, double* energy, double* force,
                                        double* virial) {
  SYS(sys);
  if (!prec_valid(precision)) {
    set_error("unknown precision %d", precision);
    return DPMD_E_ARG;
  }
  if (!s->nl_ready) { set_error("no neighbour list: call dpmd_nlist_allocate first"); return DPMD_E_STATE; }
  if (s->nl_overflow) {
    set_error("neighbour list overflowed its capacity: call dpmd_nlist_allocate again");
    return DPMD_E_STATE;
  }
  if (s->wc_model) {
    set_error("dpmd_compute_noalloc does not evaluate DPLR systems (their PPPM plans allocate)");
    return DPMD_E_UNSUPPORTED;
  }
  const Model* M = s->model;
  const size_t NA = s->NA, W = M->wout, D = (size_t)M->axis * M->wout;
  const size_t mr = std::max(1, s->mrows);
  // scratch layout, every region 256-byte aligned
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oT = take(sizeof(float) * NA * 10 * W);
  const size_t oG = take(sizeof(float) * mr * D);
  const size_t odG = take(sizeof(float) * mr * D);
  const size_t oFit = take(sizeof(float) * 7 * mr * s->fstride);
  const size_t oE = take(sizeof(float) * mr);
  const size_t oP = take(sizeof(double) * 9 * (size_t)s->nblocks);
  const size_t oR = take(sizeof(double) * (32 * (size_t)s->nseg + 64));
  const size_t oF = take(sizeof(double) * NA * 3);
  const size_t oEo = take(sizeof(double) * s->nseg);
  const size_t oVo = take(sizeof(double) * 9 * s->nseg);
  if (off > kScratchBytes) {
    set_error("system needs %zu bytes of evaluation scratch; the allocation-free entry point holds at most %zu "
              "(about %zu atoms of this model)", off, kScratchBytes, kScratchBytes / std::max<size_t>(off / NA, 1));
    return DPMD_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  unsigned char* base = nullptr;
  DPMD_TRY(scratch_device_ptr(M->device, &base));
  WorkBuffers keep(s);
  s->d_T = reinterpret_cast<float*>(base + oT);
  s->d_G = base + oG;
  s->d_dG = reinterpret_cast<float*>(base + odG);
  s->d_fitbuf = reinterpret_cast<float*>(base + oFit);
  s->d_eatom = reinterpret_cast<float*>(base + oE);
  s->d_partial = reinterpret_cast<double*>(base + oP);
  s->d_red = reinterpret_cast<double*>(base + oR);
  // G and the activations must start at zero (padding rows and K-padding columns)
  const size_t n4 = (oE - oG) / 16;  // G, dE/dG and the activations: one contiguous span
  hipLaunchKernelGGL(scratch_zero_kernel, dim3((unsigned)std::min<size_t>(4096, (n4 + 255) / 256)), dim3(256), 0,
                     s->stream, reinterpret_cast<float4*>(base + oG), n4);
  DPMD_HIP(hipGetLastError());
  DPMD_TRY(md_write_pos32(s));
  DPMD_TRY(model_forces_impl(s, precision));
  DPMD_TRY(md_reduce_launch(s));
  double* fout = reinterpret_cast<double*>(base + oF);
  double* eout = reinterpret_cast<double*>(base + oEo);
  double* vout = reinterpret_cast<double*>(base + oVo);
  const int nthr = (int)std::max(NA * 3, (size_t)s->nseg);
  hipLaunchKernelGGL(scratch_out_kernel, dim3((nthr + 255) / 256), dim3(256), 0, s->stream, s->d_force, s->d_perm,
                     (int)NA, s->d_red, s->nseg, fout, eout, vout);
  DPMD_HIP(hipGetLastError());
  if (force) DPMD_HIP(hipMemcpyAsync(force, fout, sizeof(double) * NA * 3, hipMemcpyDeviceToHost, s->stream));
  if (energy) DPMD_HIP(hipMemcpyAsync(energy, eout, sizeof(double) * s->nseg, hipMemcpyDeviceToHost, s->stream));
  if (virial) DPMD_HIP(hipMemcpyAsync(virial, vout, sizeof(double) * 9 * s->nseg, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  if (energy)
    for (int r = 0; r < s->nseg; ++r) energy[r] += seg_ebias(s, r);
  return DPMD_OK;
}

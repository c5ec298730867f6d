// C ABI of libdpmd_hip.so (declared in include/dpmd.h): model upload, system state,
// neighbour-list protocol, energy/force/virial evaluation and the device MD loop.
#include "internal.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
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

int32_t embed_launch(System* s, bool backward, int precision);
int32_t fit_launch(System* s, int precision);
int32_t md_zero_force(System* s);
int32_t md_reduce_launch(System* s, int* d_rowbase, int* d_count);
int32_t md_ke_launch(System* s, double* d_out);
int32_t md_scale_mom(System* s, const double* d_scale);
int32_t nlist_clear_flags(System* s);
int32_t md_nh_half(System* s, double kT, double dt, double tau, int chain);

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
  void* ptrs[] = {d_pos, d_mom, d_mass, d_pos_save, d_mom_save, d_pos32, d_force, d_canon, d_cell_count,
                  d_cell_start, d_cell_fill, d_cell_atoms, d_atom_cell, d_flags, d_T, d_G, d_dG, d_fitbuf,
                  d_fitbuf_h, d_eatom, d_partial, d_red, d_pairs, d_thermo, d_traj, d_nm};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (int j = 0; j < kMaxT; ++j) {
    if (d_nbr[j]) hipFree(d_nbr[j]);
    if (d_nbr_cnt[j]) hipFree(d_nbr_cnt[j]);
  }
  if (stream) hipStreamDestroy(stream);
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
  std::vector<float> emb((size_t)nn * M->off.total);
  for (int k = 0; k < nn; ++k) {
    for (int l = 0; l < 3; ++l)
      if (!d->emb_kernel[k][l] || !d->emb_bias[k][l]) { set_error("embedding net %d layer %d missing", k, l); delete M; return DPMD_E_ARG; }
    embed_pack_net(M->lay, d->emb_kernel[k][0], d->emb_bias[k][0], d->emb_kernel[k][1], d->emb_bias[k][1],
                   d->emb_kernel[k][2], d->emb_bias[k][2], emb.data() + (size_t)k * M->off.total);
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
  for (int t = 0; t < d->ntypes; ++t) {
    for (int l = 0; l < 4; ++l)
      if (!d->fit_kernel[t][l]) { set_error("fitting net %d layer %d missing", t, l); delete M; return DPMD_E_ARG; }
    const float* K0 = d->fit_kernel[t][0];
    const float* K1 = d->fit_kernel[t][1];
    const float* K2 = d->fit_kernel[t][2];
    const float* K3 = d->fit_kernel[t][3];
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
    const float b3 = d->fit_bias[t][3] ? d->fit_bias[t][3][0] : 0.f;
    o[FB3][0] = (float)(b3 * d->out_norm);
    M->fit_b3[t] = (float)(b3 * d->out_norm);
  }
  M->fit_n = FN;
  M->fit_k = FKS;
  DPMD_HIP(hipMalloc(&M->d_fit, sizeof(float) * fit.size()));
  DPMD_HIP(hipMemcpy(M->d_fit, fit.data(), sizeof(float) * fit.size(), hipMemcpyHostToDevice));
  for (int h = 0; h < 2; ++h) {
    std::vector<uint16_t> fh(fit.size());
    for (size_t i = 0; i < fit.size(); ++i) fh[i] = h ? half_bits(fit[i]) : bf16_bits(fit[i]);
    DPMD_HIP(hipMalloc(&M->d_fit_h[h], sizeof(uint16_t) * fh.size()));
    DPMD_HIP(hipMemcpy(M->d_fit_h[h], fh.data(), sizeof(uint16_t) * fh.size(), hipMemcpyHostToDevice));
  }
  *out = M;
  return DPMD_OK;
}

// ------------------------------------------------------------------------------
// system

static int32_t sys_alloc_work(System* s) {
  Model* M = s->model;
  const size_t RN = (size_t)s->nrep * s->natoms;
  const int W = M->wout, D = M->axis * M->wout;
  int rb = 0;
  for (int t = 0; t < s->ntypes; ++t) {
    s->rowbase[t] = rb;
    rb += (s->nrep * s->count[t] + 63) / 64 * 64;
  }
  s->rowbase[s->ntypes] = rb;
  s->mrows = rb;
  s->fpad = M->fit_n;
  s->fstride = M->fit_k;
  DPMD_HIP(hipMalloc(&s->d_T, sizeof(float) * RN * 10 * W));
  DPMD_HIP(hipMalloc(&s->d_G, sizeof(float) * (size_t)s->mrows * D));
  DPMD_HIP(hipMemset(s->d_G, 0, sizeof(float) * (size_t)s->mrows * D));
  DPMD_HIP(hipMalloc(&s->d_dG, sizeof(float) * (size_t)s->mrows * D));
  DPMD_HIP(hipMalloc(&s->d_fitbuf, sizeof(float) * 7 * (size_t)s->mrows * s->fstride));
  DPMD_HIP(hipMemset(s->d_fitbuf, 0, sizeof(float) * 7 * (size_t)s->mrows * s->fstride));
  DPMD_HIP(hipMalloc(&s->d_eatom, sizeof(float) * std::max(1, s->mrows)));
  s->partial_blocks = s->nrep * ((s->natoms + 3) / 4);
  DPMD_HIP(hipMalloc(&s->d_partial, sizeof(double) * 9 * s->partial_blocks));
  DPMD_HIP(hipMalloc(&s->d_red, sizeof(double) * (16 * s->nrep + 64)));
  DPMD_HIP(hipMalloc(&s->d_pairs, sizeof(int) * RN));
  DPMD_HIP(hipMemset(s->d_pairs, 0, sizeof(int) * RN));
  return DPMD_OK;
}

static int32_t sys_create(Model* M, int natoms, int nrep, const int32_t* type_idx, System** out) {
  if (!M || !out || !type_idx) { set_error("null argument"); return DPMD_E_ARG; }
  if (natoms < 1 || nrep < 1) { set_error("natoms and nreplica must be positive"); return DPMD_E_ARG; }
  if (nrep > 64) { set_error("nreplica %d above 64", nrep); return DPMD_E_ARG; }
  for (int i = 0; i < natoms; ++i)
    if (type_idx[i] < 0 || type_idx[i] >= M->ntypes) {
      set_error("type index %d out of trained scope of model (ntypes=%d)", type_idx[i], M->ntypes);
      return DPMD_E_ARG;
    }
  DPMD_HIP(hipSetDevice(M->device));
  System* s = new System();
  s->model = M;
  s->natoms = natoms;
  s->nrep = nrep;
  s->ntypes = M->ntypes;
  for (int t = 0; t < kMaxT; ++t) { s->count[t] = 0; s->knbr[t] = 0; s->cap[t] = 0; }
  for (int i = 0; i < natoms; ++i) s->count[type_idx[i]]++;
  s->off[0] = 0;
  for (int t = 0; t < s->ntypes; ++t) s->off[t + 1] = s->off[t] + s->count[t];
  s->perm.resize(natoms);
  std::iota(s->perm.begin(), s->perm.end(), 0);
  std::stable_sort(s->perm.begin(), s->perm.end(), [&](int a, int b) { return type_idx[a] < type_idx[b]; });
  s->type_sorted.resize(natoms);
  for (int i = 0; i < natoms; ++i) s->type_sorted[i] = type_idx[s->perm[i]];
  DPMD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  const size_t RN = (size_t)natoms * nrep;
  DPMD_HIP(hipMalloc(&s->d_pos, sizeof(double) * RN * 3));
  DPMD_HIP(hipMalloc(&s->d_mom, sizeof(double) * RN * 3));
  DPMD_HIP(hipMalloc(&s->d_pos_save, sizeof(double) * RN * 3));
  DPMD_HIP(hipMalloc(&s->d_mom_save, sizeof(double) * RN * 3));
  DPMD_HIP(hipMalloc(&s->d_mass, sizeof(double) * natoms));
  DPMD_HIP(hipMalloc(&s->d_pos32, sizeof(float4) * RN));
  DPMD_HIP(hipMalloc(&s->d_force, sizeof(float) * RN * 3));
  DPMD_HIP(hipMalloc(&s->d_flags, sizeof(int) * 16));
  DPMD_HIP(hipMalloc(&s->d_thermo, sizeof(double) * 64 * nrep));
  DPMD_HIP(hipMemset(s->d_pos, 0, sizeof(double) * RN * 3));
  DPMD_HIP(hipMemset(s->d_mom, 0, sizeof(double) * RN * 3));
  DPMD_HIP(hipMemset(s->d_force, 0, sizeof(float) * RN * 3));
  DPMD_HIP(hipMemset(s->d_flags, 0, sizeof(int) * 16));
  DPMD_HIP(hipMemset(s->d_thermo, 0, sizeof(double) * 64 * nrep));
  std::vector<double> ones(natoms, 1.0);
  DPMD_HIP(hipMemcpy(s->d_mass, ones.data(), sizeof(double) * natoms, hipMemcpyHostToDevice));
  int32_t rc = sys_alloc_work(s);
  if (rc != DPMD_OK) { delete s; return rc; }
  *out = s;
  return DPMD_OK;
}

static int32_t model_forces_impl(System* s, int precision) {
  hipEvent_t ev;
  s->force_valid = true;
  DPMD_TRY(md_zero_force(s));
  timer_begin(s, 1, &ev);
  DPMD_TRY(embed_launch(s, false, precision));
  timer_end(s, 1, ev);
  timer_begin(s, 2, &ev);
  DPMD_TRY(fit_launch(s, precision));
  timer_end(s, 2, ev);
  timer_begin(s, 3, &ev);
  DPMD_TRY(embed_launch(s, true, precision));
  timer_end(s, 3, ev);
  return DPMD_OK;
}

static int32_t nl_threshold(System* s, float* thr) {
  const double rc = s->model->rcut;
  float sc = 0.f;
  for (int d = 0; d < 3; ++d) sc = std::max(sc, (float)s->ref_box[d] / (float)s->box[d]);
  const float allowed = ((float)s->rc_skin - (float)rc * sc) / 2.f;
  const double lmin = std::min(s->ref_box[0], std::min(s->ref_box[1], s->ref_box[2]));
  *thr = allowed - (float)(lmin * 1e-6);
  return DPMD_OK;
}

static int32_t nl_read_flags(System* s, int* h) {
  DPMD_HIP(hipMemcpyAsync(h, s->d_flags, sizeof(int) * 16, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

static int32_t nl_allocate(System* s, double rc_skin, double ratio, const int32_t* kover, int32_t* kout) {
  if (!(rc_skin > 0)) { set_error("rc_skin must be positive"); return DPMD_E_ARG; }
  if (s->box[0] <= 0) { set_error("box not set"); return DPMD_E_STATE; }
  for (int d = 0; d < 3; ++d)
    if (s->box[d] < 2.0 * rc_skin) {
      set_error("box length %g below 2*(rcut+skin)=%g: neighbour list needs the minimum-image regime", s->box[d], 2 * rc_skin);
      return DPMD_E_ARG;
    }
  s->rc_skin = rc_skin;
  s->buffer_ratio = ratio;
  for (int d = 0; d < 3; ++d) s->ref_box[d] = s->box[d];
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
    if (s->d_nbr[j]) hipFree(s->d_nbr[j]);
    if (s->d_nbr_cnt[j]) hipFree(s->d_nbr_cnt[j]);
    s->d_nbr[j] = nullptr;
    s->d_nbr_cnt[j] = nullptr;
    const size_t RN = (size_t)s->nrep * s->natoms;
    DPMD_HIP(hipMalloc(&s->d_nbr[j], sizeof(int) * RN * s->cap[j]));
    DPMD_HIP(hipMalloc(&s->d_nbr_cnt[j], sizeof(int) * RN));
  }
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
  const int N = s->natoms;
  out.resize((size_t)s->nrep * N * 3);
  for (int r = 0; r < s->nrep; ++r)
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) out[((size_t)r * N + i) * 3 + d] = in[((size_t)r * N + s->perm[i]) * 3 + d];
}

static void from_sorted(System* s, const double* in, double* out) {
  const int N = s->natoms;
  for (int r = 0; r < s->nrep; ++r)
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) out[((size_t)r * N + s->perm[i]) * 3 + d] = in[((size_t)r * N + i) * 3 + d];
}

static int32_t compute_impl(System* s, int precision, double* energy, double* force, double* virial, double* eatom) {
  if (!s->nl_ready) { set_error("no neighbour list: call dpmd_nlist_allocate first"); return DPMD_E_STATE; }
  Model* M = s->model;
  DPMD_TRY(md_write_pos32(s));
  DPMD_TRY(model_forces_impl(s, precision));
  int meta[2 * kMaxT];
  for (int t = 0; t < s->ntypes; ++t) { meta[t] = s->rowbase[t]; meta[kMaxT + t] = s->count[t]; }
  int* d_meta = reinterpret_cast<int*>(s->d_red + 16 * s->nrep);
  DPMD_HIP(hipMemcpyAsync(d_meta, meta, sizeof(meta), hipMemcpyHostToDevice, s->stream));
  DPMD_TRY(md_reduce_launch(s, d_meta, d_meta + kMaxT));
  std::vector<double> red((size_t)10 * s->nrep);
  DPMD_HIP(hipMemcpyAsync(red.data(), s->d_red, sizeof(double) * 10 * s->nrep, hipMemcpyDeviceToHost, s->stream));
  const size_t RN = (size_t)s->nrep * s->natoms;
  std::vector<float> f;
  if (force) {
    f.resize(RN * 3);
    DPMD_HIP(hipMemcpyAsync(f.data(), s->d_force, sizeof(float) * RN * 3, hipMemcpyDeviceToHost, s->stream));
  }
  std::vector<float> ea;
  if (eatom) {
    ea.resize(s->mrows);
    DPMD_HIP(hipMemcpyAsync(ea.data(), s->d_eatom, sizeof(float) * s->mrows, hipMemcpyDeviceToHost, s->stream));
  }
  std::vector<int> pc(RN);
  DPMD_HIP(hipMemcpyAsync(pc.data(), s->d_pairs, sizeof(int) * RN, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  double ps = 0;
  for (size_t i = 0; i < RN; ++i) ps += pc[i];
  s->last_pairs = ps / RN;
  double eb = 0;
  for (int t = 0; t < s->ntypes; ++t) eb += M->ebias[t] * s->count[t] * M->out_norm;
  for (int r = 0; r < s->nrep; ++r) {
    if (energy) energy[r] = red[r * 10] + eb;
    if (virial)
      for (int k = 0; k < 9; ++k) virial[r * 9 + k] = red[r * 10 + 1 + k];
  }
  if (force) {
    std::vector<double> fd(RN * 3);
    for (size_t i = 0; i < RN * 3; ++i) fd[i] = f[i];
    from_sorted(s, fd.data(), force);
  }
  if (eatom) {
    const int N = s->natoms;
    for (int r = 0; r < s->nrep; ++r)
      for (int i = 0; i < N; ++i) {
        const int t = s->type_sorted[i];
        const int row = s->rowbase[t] + r * s->count[t] + (i - s->off[t]);
        eatom[(size_t)r * N + s->perm[i]] = ea[row] + M->ebias[t] * M->out_norm;
      }
  }
  return DPMD_OK;
}

}  // namespace dpmd

using namespace dpmd;

extern "C" {

const char* dpmd_last_error(void) { return g_err; }
int32_t dpmd_version(void) { return 1; }

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
  for (int k = 0; k < 2; ++k) {
    if (M->d_emb_h[k]) hipFree(M->d_emb_h[k]);
    if (M->d_fit_h[k]) hipFree(M->d_fit_h[k]);
  }
  delete M;
  return DPMD_OK;
}

int32_t dpmd_system_create(dpmd_model_t model, int32_t natoms, int32_t nreplica, const int32_t* type_idx,
                           dpmd_system_t* out) {
  System* s = nullptr;
  int32_t rc = sys_create(reinterpret_cast<Model*>(model), natoms, nreplica, type_idx, &s);
  if (rc == DPMD_OK) *out = reinterpret_cast<dpmd_system_t>(s);
  return rc;
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
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !box3) { set_error("null argument"); return DPMD_E_ARG; }
  for (int d = 0; d < 3; ++d)
    if (!(box3[d] > 0)) { set_error("box lengths must be positive"); return DPMD_E_ARG; }
  for (int d = 0; d < 3; ++d) s->box[d] = box3[d];
  if (!s->nl_ready)
    for (int d = 0; d < 3; ++d) s->ref_box[d] = box3[d];
  return DPMD_OK;
}

int32_t dpmd_system_set_positions(dpmd_system_t sys, const double* pos) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !pos) { set_error("null argument"); return DPMD_E_ARG; }
  if (s->box[0] <= 0) { set_error("box not set"); return DPMD_E_STATE; }
  hipSetDevice(s->model->device);
  std::vector<double> v;
  to_sorted(s, pos, v);
  DPMD_HIP(hipMemcpyAsync(s->d_pos, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, s->stream));
  DPMD_TRY(md_write_pos32(s));
  s->force_valid = false;
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_system_get_positions(dpmd_system_t sys, double* pos) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !pos) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> v((size_t)s->nrep * s->natoms * 3);
  DPMD_HIP(hipMemcpyAsync(v.data(), s->d_pos, sizeof(double) * v.size(), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  from_sorted(s, v.data(), pos);
  return DPMD_OK;
}

static int32_t get_masses_sorted(System* s, std::vector<double>& m) {
  m.resize(s->natoms);
  DPMD_HIP(hipMemcpy(m.data(), s->d_mass, sizeof(double) * s->natoms, hipMemcpyDeviceToHost));
  return DPMD_OK;
}

int32_t dpmd_system_set_velocities(dpmd_system_t sys, const double* vel) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !vel) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> v, m;
  to_sorted(s, vel, v);
  DPMD_TRY(get_masses_sorted(s, m));
  const int N = s->natoms;
  for (int r = 0; r < s->nrep; ++r)
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) v[((size_t)r * N + i) * 3 + d] *= m[i];
  DPMD_HIP(hipMemcpyAsync(s->d_mom, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_system_get_velocities(dpmd_system_t sys, double* vel) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !vel) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> v((size_t)s->nrep * s->natoms * 3), m;
  DPMD_HIP(hipMemcpyAsync(v.data(), s->d_mom, sizeof(double) * v.size(), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  DPMD_TRY(get_masses_sorted(s, m));
  const int N = s->natoms;
  for (int r = 0; r < s->nrep; ++r)
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) v[((size_t)r * N + i) * 3 + d] /= m[i];
  from_sorted(s, v.data(), vel);
  return DPMD_OK;
}

int32_t dpmd_system_set_masses(dpmd_system_t sys, const double* mass) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !mass) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> m(s->natoms);
  for (int i = 0; i < s->natoms; ++i) {
    m[i] = mass[s->perm[i]];
    if (!(m[i] > 0)) { set_error("masses must be positive"); return DPMD_E_ARG; }
  }
  DPMD_HIP(hipMemcpy(s->d_mass, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice));
  return DPMD_OK;
}

int32_t dpmd_nlist_allocate(dpmd_system_t sys, double rc_skin, double buffer_ratio, const int32_t* knbr_override,
                            int32_t* knbr_out) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s) { set_error("null argument"); return DPMD_E_ARG; }
  hipSetDevice(s->model->device);
  return nl_allocate(s, rc_skin, buffer_ratio, knbr_override, knbr_out);
}

int32_t dpmd_nlist_update(dpmd_system_t sys, int32_t* overflow) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s) { set_error("null argument"); return DPMD_E_ARG; }
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
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !out) { set_error("null argument"); return DPMD_E_ARG; }
  if (!s->nl_ready) { set_error("neighbour list not allocated"); return DPMD_E_STATE; }
  if (j < 0 || j >= s->ntypes) { set_error("type %d out of range", j); return DPMD_E_ARG; }
  const size_t RN = (size_t)s->nrep * s->natoms;
  DPMD_HIP(hipMemcpyAsync(out, s->d_nbr[j], sizeof(int) * RN * s->cap[j], hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_nlist_capacity(dpmd_system_t sys, int32_t* knbr) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !knbr) { set_error("null argument"); return DPMD_E_ARG; }
  for (int j = 0; j < s->ntypes; ++j) knbr[j] = s->knbr[j];
  return DPMD_OK;
}

int32_t dpmd_compute(dpmd_system_t sys, int32_t precision, double* energy, double* force, double* virial,
                     double* atom_energy) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s) { set_error("null argument"); return DPMD_E_ARG; }
  if (precision != DPMD_PREC_FP32 && precision != DPMD_PREC_BF16 && precision != DPMD_PREC_FP16) {
    set_error("unknown precision %d", precision);
    return DPMD_E_ARG;
  }
  hipSetDevice(s->model->device);
  return compute_impl(s, precision, energy, force, virial, atom_energy);
}

int32_t dpmd_md_run(dpmd_system_t sys, const dpmd_md_params* p, int32_t nsteps, double* traj_pos, double* traj_vel,
                    uint32_t* error_code) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !p) { set_error("null argument"); return DPMD_E_ARG; }
  if (!s->nl_ready) { set_error("no neighbour list: call dpmd_nlist_allocate first"); return DPMD_E_STATE; }
  if (nsteps < 0) { set_error("nsteps must be >= 0"); return DPMD_E_ARG; }
  if (p->routine != DPMD_NVE && p->routine != DPMD_NVT_LANGEVIN && p->routine != DPMD_NVT_NH &&
      p->routine != DPMD_PIMD_PILE_L) {
    set_error("unknown routine %d", p->routine);
    return DPMD_E_ARG;
  }
  if (p->routine == DPMD_PIMD_PILE_L && !s->d_nm) { set_error("PIMD normal modes not set"); return DPMD_E_STATE; }
  hipSetDevice(s->model->device);
  const size_t RN = (size_t)s->nrep * s->natoms;
  const bool rec = traj_pos || traj_vel;
  if (rec) {
    const size_t need = (size_t)nsteps * RN * 3 * 2;
    if (need > s->traj_cap) {
      if (s->d_traj) hipFree(s->d_traj);
      s->d_traj = nullptr;
      DPMD_HIP(hipMalloc(&s->d_traj, sizeof(double) * need));
      s->traj_cap = need;
    }
  }
  DPMD_HIP(hipMemcpyAsync(s->d_pos_save, s->d_pos, sizeof(double) * RN * 3, hipMemcpyDeviceToDevice, s->stream));
  DPMD_HIP(hipMemcpyAsync(s->d_mom_save, s->d_mom, sizeof(double) * RN * 3, hipMemcpyDeviceToDevice, s->stream));
  if (s->nl_overflow) {
    // the current list already lacks neighbours (deepmd_jax/md.py:637-641): report
    // bit 2 before stepping so the caller reallocates
    if (error_code) *error_code |= 2u;
    return DPMD_OK;
  }
  float thr;
  nl_threshold(s, &thr);
  DPMD_TRY(nlist_clear_flags(s));
  const double dt = p->dt;
  const int prec = p->precision;
  if (!s->force_valid) {
    DPMD_TRY(md_write_pos32(s));
    DPMD_TRY(model_forces_impl(s, prec));
  }
  for (int step = 0; step < nsteps; ++step) {
    DPMD_TRY(nlist_drift_check(s, thr));
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
      timer_end(s, 4, ev);
    } else if (p->routine == DPMD_NVT_LANGEVIN) {
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_drift(s, 0.5 * dt));
      DPMD_TRY(md_langevin_o(s, dt, p->gamma, p->kT, p->seed, s->rng_counter++));
      DPMD_TRY(md_drift(s, 0.5 * dt));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_kick(s, 0.5 * dt));
    } else if (p->routine == DPMD_NVT_NH) {
      DPMD_TRY(md_nh_half(s, p->kT, dt, p->tau, p->chain_length));
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_drift(s, dt));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_nh_half(s, p->kT, dt, p->tau, p->chain_length));
    } else {
      const int P = s->nrep;
      const double* C = s->d_nm;
      const double* fr = s->d_nm + P * P;
      const double* c1 = fr + P;
      const double* c2 = c1 + P;
      DPMD_TRY(md_kick(s, 0.5 * dt));
      DPMD_TRY(md_pimd_free_ring(s, C, fr, 0.5 * dt));
      DPMD_TRY(md_pimd_o(s, C, c1, c2, p->seed, s->rng_counter++));
      DPMD_TRY(md_pimd_free_ring(s, C, fr, 0.5 * dt));
      DPMD_TRY(md_write_pos32(s));
      DPMD_TRY(model_forces_impl(s, prec));
      DPMD_TRY(md_kick(s, 0.5 * dt));
    }
    if (rec) {
      DPMD_HIP(hipMemcpyAsync(s->d_traj + (size_t)step * RN * 6, s->d_pos, sizeof(double) * RN * 3,
                              hipMemcpyDeviceToDevice, s->stream));
      DPMD_HIP(hipMemcpyAsync(s->d_traj + (size_t)step * RN * 6 + RN * 3, s->d_mom, sizeof(double) * RN * 3,
                              hipMemcpyDeviceToDevice, s->stream));
    }
  }
  int h[16];
  DPMD_TRY(nl_read_flags(s, h));
  uint32_t code = (uint32_t)(h[2] & (2 | 16));
  if (h[8]) code |= 2;
  if (error_code) *error_code |= code;
  if (code & 2) {
    // neighbour capacity exceeded during the chunk: restore the entry state so the caller
    // can reallocate and re-run (deepmd_jax/md.py:914-916)
    DPMD_HIP(hipMemcpyAsync(s->d_pos, s->d_pos_save, sizeof(double) * RN * 3, hipMemcpyDeviceToDevice, s->stream));
    DPMD_HIP(hipMemcpyAsync(s->d_mom, s->d_mom_save, sizeof(double) * RN * 3, hipMemcpyDeviceToDevice, s->stream));
    DPMD_TRY(md_write_pos32(s));
    DPMD_HIP(hipStreamSynchronize(s->stream));
    s->force_valid = false;
    s->nl_overflow = true;
    return DPMD_OK;
  }
  if (rec) {
    std::vector<double> buf((size_t)nsteps * RN * 6);
    DPMD_HIP(hipMemcpyAsync(buf.data(), s->d_traj, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, s->stream));
    DPMD_HIP(hipStreamSynchronize(s->stream));
    std::vector<double> m;
    DPMD_TRY(get_masses_sorted(s, m));
    const int N = s->natoms;
    for (int k = 0; k < nsteps; ++k) {
      double* bp = buf.data() + (size_t)k * RN * 6;
      double* bm = bp + RN * 3;
      if (traj_pos) from_sorted(s, bp, traj_pos + (size_t)k * RN * 3);
      if (traj_vel) {
        for (int r = 0; r < s->nrep; ++r)
          for (int i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d) bm[((size_t)r * N + i) * 3 + d] /= m[i];
        from_sorted(s, bm, traj_vel + (size_t)k * RN * 3);
      }
    }
  }
  return DPMD_OK;
}

int32_t dpmd_pimd_setup(dpmd_system_t sys, const double* C, const double* freqs, const double* c1, const double* c2) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !C || !freqs || !c1 || !c2) { set_error("null argument"); return DPMD_E_ARG; }
  const int P = s->nrep;
  if (P < 2) { set_error("PIMD needs nreplica >= 2"); return DPMD_E_ARG; }
  for (int k = 1; k < P; ++k)
    if (!(freqs[k] > 0)) { set_error("non-centroid normal-mode frequencies must be positive"); return DPMD_E_ARG; }
  std::vector<double> buf((size_t)P * P + 3 * P);
  std::copy(C, C + P * P, buf.begin());
  std::copy(freqs, freqs + P, buf.begin() + P * P);
  std::copy(c1, c1 + P, buf.begin() + P * P + P);
  std::copy(c2, c2 + P, buf.begin() + P * P + 2 * P);
  if (!s->d_nm) DPMD_HIP(hipMalloc(&s->d_nm, sizeof(double) * buf.size()));
  DPMD_HIP(hipMemcpy(s->d_nm, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
  return DPMD_OK;
}

int32_t dpmd_kinetic_energy(dpmd_system_t sys, double* ke) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !ke) { set_error("null argument"); return DPMD_E_ARG; }
  double* d_out = s->d_red + 10 * s->nrep;
  DPMD_TRY(md_ke_launch(s, d_out));
  DPMD_HIP(hipMemcpyAsync(ke, d_out, sizeof(double) * s->nrep, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

int32_t dpmd_thermostat_energy(dpmd_system_t sys, double* e) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !e) { set_error("null argument"); return DPMD_E_ARG; }
  std::vector<double> t((size_t)64 * s->nrep);
  DPMD_HIP(hipMemcpyAsync(t.data(), s->d_thermo, sizeof(double) * t.size(), hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  for (int r = 0; r < s->nrep; ++r) e[r] = t[(size_t)r * 64];
  return DPMD_OK;
}

int32_t dpmd_timing_enable(dpmd_system_t sys, int32_t on) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s) { set_error("null argument"); return DPMD_E_ARG; }
  timer_harvest(s);
  s->timer.on = on != 0;
  for (int w = 0; w < 8; ++w) { s->timer.ms[w] = 0; s->timer.count[w] = 0; }
  return DPMD_OK;
}

int32_t dpmd_timing_get(dpmd_system_t sys, int32_t which, double* ms, int64_t* count) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || which < 0 || which >= 8) { set_error("bad argument"); return DPMD_E_ARG; }
  timer_harvest(s);
  if (ms) *ms = s->timer.ms[which];
  if (count) *count = s->timer.count[which];
  return DPMD_OK;
}

int32_t dpmd_pair_count(dpmd_system_t sys, double* pairs) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s || !pairs) { set_error("null argument"); return DPMD_E_ARG; }
  const size_t RN = (size_t)s->nrep * s->natoms;
  std::vector<int> pc(RN);
  DPMD_HIP(hipMemcpyAsync(pc.data(), s->d_pairs, sizeof(int) * RN, hipMemcpyDeviceToHost, s->stream));
  DPMD_HIP(hipStreamSynchronize(s->stream));
  double t = 0;
  for (size_t i = 0; i < RN; ++i) t += pc[i];
  *pairs = t / RN;
  return DPMD_OK;
}

int32_t dpmd_synchronize(dpmd_system_t sys) {
  System* s = reinterpret_cast<System*>(sys);
  if (!s) { set_error("null argument"); return DPMD_E_ARG; }
  DPMD_HIP(hipStreamSynchronize(s->stream));
  return DPMD_OK;
}

}  // extern "C"

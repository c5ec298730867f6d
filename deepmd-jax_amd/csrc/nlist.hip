// Typed neighbour list on the device: cell binning, 27-cell scan, per-row ascending
// type-split output, overflow detection, and the per-step drift test that decides a
// rebuild.
//
// Semantics restated from deepmd-jax (and the jax_md neighbour list it wraps):
//   canonicalize   deepmd_jax/md_utils.py:364-373   c = ((f32(x) mod box) * scale) * f32(1-5e-7),
//                  scale = L / box (L: reference box of the list, box: current box)
//   pair test      jax_md space.periodic + square_distance (call site md_utils.py:384-415):
//                  d = (ca - cb); d = ((d + L/2) mod L) - L/2; d2 = (dx^2 + dy^2) + dz^2 < cut^2,
//                  self excluded
//   rows           md_utils.py:431-459: per neighbour type, ascending type-local indices,
//                  padded with N, knbr_j - 1 columns; overflow if a row needs knbr_j
//   drift test     deepmd_jax/md.py:693-741: max |minimg_L((f32(x) mod box)*scale - c_ref)|
//                  > ((rc+skin) - rc*max(scale)*(1.02 under NPT))/2 - 1e-6 L_min
// All float32 expressions are evaluated without contraction so the neighbour sets are
// bit-exact with a float32 evaluation of the same formulas.
// General (triclinic) cells, which the reference evaluates on its dense lattice-image path
// (deepmd_jax/utils.py:20-27,66-78): canonical coordinates are fractional in [0, 1), cells
// bin them per lattice direction (edge: perpendicular width / count >= the cutoff), and the
// pair and drift tests use the fractional minimum image mapped back through the cell.
#include "internal.h"
#include <algorithm>
#include <string.h>

namespace dpmd {

__device__ __forceinline__ void nl_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#pragma clang fp contract(off)

__device__ __forceinline__ float fmod_floor(float x, float y) {
  float r = fmodf(x, y);
  if (r != 0.f && ((r < 0.f) != (y < 0.f))) r = r + y;
  return r;
}

// One coordinate wrapped into [0, box) in FP64, rounded to FP32 and wrapped again against
// the FP32 box (a value that rounds up to the box becomes 0).  The reference keeps its FP32
// positions wrapped every step, so its canonicalisation (f32(x) mod box,
// deepmd_jax/md_utils.py:364-373) sees the correctly rounded wrapped coordinate; wrapping
// the engine's unwrapped FP64 positions first gives that same value, where rounding an
// unwrapped coordinate to FP32 first would lose the low bits of atoms that left the box.
__device__ __forceinline__ float wrap_axis_f32(double x, double Ld, float Lf) {
  double r = fmod(x, Ld);
  if (r < 0) r += Ld;
  return fmod_floor((float)r, Lf);
}

// fmod_floor(x, L) for x in (-L, 2L), the range of d + L/2 for two canonical coordinates in
// [0, L): x - L is exact there (Sterbenz), and x + L is the same rounded sum fmod_floor forms,
// so the result is bit-identical without the library fmod.
__device__ __forceinline__ float wrap_half_open(float x, float L) {
  const float r = x >= L ? x - L : x;
  return r < 0.f ? r + L : r;
}

__device__ __forceinline__ float pair_d2(float4 a, float4 b, float Lx, float Ly, float Lz) {
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const float hx = Lx * 0.5f, hy = Ly * 0.5f, hz = Lz * 0.5f;
  dx = wrap_half_open(dx + hx, Lx) - hx;
  dy = wrap_half_open(dy + hy, Ly) - hy;
  dz = wrap_half_open(dz + hz, Lz) - hz;
  const float sx = dx * dx, sy = dy * dy, sz = dz * dz;
  return (sx + sy) + sz;
}

// general cell: canonical coordinates are fractional; the fractional minimum image is the
// nearest image for every pair within the cutoff when all perpendicular widths exceed
// twice the cutoff (checked at allocation)
__device__ __forceinline__ float pair_d2_cell(float4 a, float4 b, const float* H) {
  float s[3] = {a.x - b.x, a.y - b.y, a.z - b.z}, d[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) s[e] -= rintf(s[e]);
  cell_mul(H, s, d);
  return (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
}

// fractional coordinates in [0, 1) of a position in a general cell
__device__ __forceinline__ void cell_frac01(const SegInfo& S, const double* x, float f[3]) {
  double s[3];
  cell_mul(S.Hid, x, s);
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    float v = (float)(s[e] - floor(s[e]));
    f[e] = v >= 1.f ? 0.f : v;
  }
}

struct NlArgs {
  const double* pos;   // (NA*3) type-sorted per segment
  const SegInfo* seg;
  const int* atom_seg; // (NA)
  float4* canon;       // (NA)
  int* atom_cell;      // (NA) cell inside the segment's grid
  int* cell_count;     // (ncell_total)
  int* cell_start;
  int* cell_fill;
  int* cell_atoms;     // (NA) segment-local atom indices grouped by cell
  int* nbr[kMaxT];
  int* nbr_cnt[kMaxT];
  int cap[kMaxT];
  int* rev[kMaxT];     // deterministic forces: slot of this atom in each neighbour's row
  int ntypes, NA, ncell_total;
  float cut_sq;
  int* flags;
  int conditional;
  int measure;
  int lcap;
};

__global__ void nl_prepare_kernel(NlArgs a) {
  if (a.conditional && a.flags[0] == 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.ncell_total; i += gridDim.x * blockDim.x) {
    a.cell_count[i] = 0;
    a.cell_fill[i] = 0;
  }
}

__global__ void nl_canon_kernel(NlArgs a) {
  if (a.conditional && a.flags[0] == 0) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.NA) return;
  const SegInfo& S = a.seg[a.atom_seg[g]];
  const float shrink = (float)(1.0 - 5e-7);
  float c[3];
  if (S.tri) {
    cell_frac01(S, a.pos + (size_t)g * 3, c);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // wrap in the current box, then scale to the reference box
      float x = wrap_axis_f32(a.pos[(size_t)g * 3 + d], S.boxd[d], S.box[d]);
      x = x * S.scale[d];
      c[d] = x * shrink;
    }
  }
  a.canon[g] = make_float4(c[0], c[1], c[2], 0.f);
  if (S.use_cells) {
    int ci[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float u = S.tri ? c[d] * (float)S.nc[d] : c[d] * ((float)S.nc[d] / S.L[d]);
      // clamped before the conversion, so a non-finite position (error bit 16 ends the
      // run) lands in cell 0 instead of an undefined integer
      ci[d] = u >= 0.f && u < (float)S.nc[d] ? (int)u : (u >= (float)S.nc[d] ? S.nc[d] - 1 : 0);
    }
    const int cell = (ci[2] * S.nc[1] + ci[1]) * S.nc[0] + ci[0];
    a.atom_cell[g] = cell;
    atomicAdd(&a.cell_count[S.cell0 + cell], 1);
  }
}

// one block per segment: exclusive scan over the segment's cells
__global__ void nl_scan_kernel(NlArgs a) {
  if (a.conditional && a.flags[0] == 0) return;
  __shared__ int part[1024];
  const SegInfo& S = a.seg[blockIdx.x];
  if (!S.use_cells) return;
  const int ncell = S.nc[0] * S.nc[1] * S.nc[2];
  const int* cnt = a.cell_count + S.cell0;
  int* st = a.cell_start + S.cell0;
  const int per = (ncell + blockDim.x - 1) / blockDim.x;
  const int b = threadIdx.x * per;
  int s = 0;
  for (int k = 0; k < per; ++k)
    if (b + k < ncell) s += cnt[b + k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 1; d < blockDim.x; d <<= 1) {
    int v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0;
  for (int k = 0; k < per; ++k)
    if (b + k < ncell) {
      st[b + k] = run;
      run += cnt[b + k];
    }
}

__global__ void nl_scatter_kernel(NlArgs a) {
  if (a.conditional && a.flags[0] == 0) return;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.NA) return;
  const SegInfo& S = a.seg[a.atom_seg[g]];
  if (!S.use_cells) return;
  const int cell = S.cell0 + a.atom_cell[g];
  const int slot = a.cell_start[cell] + atomicAdd(&a.cell_fill[cell], 1);
  a.cell_atoms[S.base + slot] = g - S.base;
}

// one wave per centre
__global__ void __launch_bounds__(256) nl_build_kernel(NlArgs a) {
  if (a.conditional && a.flags[0] == 0) return;
  extern __shared__ int nl_smem[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int* list = nl_smem + wv * 2 * a.lcap;
  int* sorted = list + a.lcap;
  const int g = blockIdx.x * 4 + wv;
  if (g >= a.NA) return;
  const int sid = __builtin_amdgcn_readfirstlane(a.atom_seg[g]);
  const SegInfo& S = a.seg[sid];
  const int N = S.natoms, i = g - S.base;
  const float4* cn = a.canon + S.base;
  const float4 ci = cn[i];
  const float Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
  const bool tri = S.tri != 0;
  int cnt = 0;
  // pa: the position of `atom` (read for every lane; masked lanes carry atom 0)
  auto consider = [&](int atom, bool in, float4 pa) {
    bool keep = false;
    if (in && atom != i)
      keep = (tri ? pair_d2_cell(ci, pa, S.H) : pair_d2(ci, pa, Lx, Ly, Lz)) < a.cut_sq;
    const unsigned long long bal = __ballot(keep);
    const int pre = __popcll(bal & ((1ull << lane) - 1ull));
    if (keep && cnt + pre < a.lcap) list[cnt + pre] = atom;
    cnt += __popcll(bal);
  };
  if (S.use_cells) {
    const int cell = a.atom_cell[g];
    const int nx = S.nc[0], ny = S.nc[1], nz = S.nc[2];
    const int cx = cell % nx, cy = (cell / nx) % ny, cz = cell / (nx * ny);
    const int* catoms = a.cell_atoms + S.base;
    // the 27 cells (dz outer, dx inner) with their start and count in lanes 0..26, read at once
    int cst = 0, ccnt = 0;
    if (lane < 27) {
      const int x = (cx + lane % 3 - 1 + nx) % nx;
      const int y = (cy + (lane / 3) % 3 - 1 + ny) % ny;
      const int z = (cz + lane / 9 - 1 + nz) % nz;
      const int cc = S.cell0 + (z * ny + y) * nx + x;
      cst = a.cell_start[cc];
      ccnt = a.cell_count[cc];
    }
    // the cells' atoms in pieces of 64 (cell k, offset b; k = 27: past the end), tested in
    // the order of the cells; a piece's atom indices are read two pieces ahead and their
    // positions one piece ahead, so neither dependent read waits in line
    auto settle = [&](int& k, int& b) {
      while (k < 27 && b >= __builtin_amdgcn_readlane(ccnt, k)) {
        ++k;
        b = 0;
      }
    };
    auto advance = [&](int& k, int& b) {
      if (k < 27) {
        b += 64;
        settle(k, b);
      }
    };
    auto load_atom = [&](int k, int b, bool& in) -> int {
      in = false;
      if (k >= 27) return 0;
      const int st = __builtin_amdgcn_readlane(cst, k), num = __builtin_amdgcn_readlane(ccnt, k);
      in = b + lane < num;
      return in ? catoms[st + b + lane] : 0;
    };
    int kA = 0, bA = 0;
    settle(kA, bA);
    int kB = kA, bB = bA;
    advance(kB, bB);
    bool inA, inB;
    int atomA = load_atom(kA, bA, inA);
    int atomB = load_atom(kB, bB, inB);
    float4 pA = cn[atomA];
    while (kA < 27) {
      int kC = kB, bC = bB;
      advance(kC, bC);
      bool inC;
      const int atomC = load_atom(kC, bC, inC);
      const float4 pB = cn[atomB];
      consider(atomA, inA, pA);
      kA = kB; bA = bB; atomA = atomB; inA = inB; pA = pB;
      kB = kC; bB = bC; atomB = atomC; inB = inC;
    }
  } else {
    for (int b = 0; b < N; b += 64) {
      const bool in = b + lane < N;
      const int atom = in ? b + lane : 0;
      consider(atom, in, cn[atom]);
    }
  }
  if (cnt > a.lcap) {
    if (lane == 0) atomicOr(&a.flags[8], 1);
    cnt = a.lcap;
  }
  nl_wave_sync();
  for (int e = lane; e < cnt; e += 64) {
    const int v = list[e];
    int rank = 0;
    for (int k = 0; k < cnt; ++k) rank += list[k] < v ? 1 : 0;
    sorted[rank] = v;
  }
  nl_wave_sync();
  int start = 0;
  for (int j = 0; j < a.ntypes; ++j) {
    const int o0 = S.off[j], o1 = S.off[j + 1];
    int c = 0;
    for (int e = lane; e < cnt; e += 64) {
      const int v = sorted[e];
      c += (v >= o0 && v < o1) ? 1 : 0;
    }
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) c += __shfl_xor(c, s);
    if (a.measure) {
      if (lane == 0) atomicMax(&a.flags[4 + j], c);
    } else {
      const int capj = a.cap[j];
      if (c > capj && lane == 0) atomicOr(&a.flags[2], 2);
      int* row = a.nbr[j] + (size_t)g * capj;
      for (int p = lane; p < capj; p += 64) row[p] = p < c ? sorted[start + p] - o0 : N;
      if (lane == 0) a.nbr_cnt[j][g] = c < capj ? c : capj;
    }
    start += c;
  }
}

// Reverse index of the rows for the deterministic force gather: for entry p of atom g's
// row of type t (neighbour i), the position e of g in i's row of g's type, found by binary
// search (rows hold ascending type-local indices), or -1 when i's row does not list g (a
// pair at the skin edge whose FP32 test came out differently in the two directions; such a
// pair lies beyond rcut and carries no gradient).  One wave per atom.
__global__ void __launch_bounds__(256) nl_reverse_kernel(NlArgs a) {
  if (a.conditional && a.flags[0] == 0) return;
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= a.NA) return;
  const SegInfo& S = a.seg[a.atom_seg[g]];
  const int i0 = g - S.base;
  int tg = 0;
  while (tg + 1 < a.ntypes && i0 >= S.off[tg + 1]) ++tg;
  const int me = i0 - S.off[tg];
  const int capg = a.cap[tg];
  for (int t = 0; t < a.ntypes; ++t) {
    const int capt = a.cap[t];
    const int cnt = a.nbr_cnt[t][g];
    const int* row = a.nbr[t] + (size_t)g * capt;
    int* rev = a.rev[t] + (size_t)g * capt;
    for (int p = lane; p < cnt; p += 64) {
      const int gi = S.base + S.off[t] + row[p];
      const int* ri = a.nbr[tg] + (size_t)gi * capg;
      int lo = 0, hi = a.nbr_cnt[tg][gi];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ri[mid] < me) lo = mid + 1;
        else hi = mid;
      }
      rev[p] = (lo < a.nbr_cnt[tg][gi] && ri[lo] == me) ? lo : -1;
    }
  }
}

// drift of every atom against the canonical coordinates of the last build; sets
// flags[1] when any atom of any segment moved beyond its segment's threshold.  With
// shrink_max > 0 (NPT) the thread of a segment's first atom also runs the box test of
// deepmd_jax/md.py:630-636 in FP32: the list's reference box no longer bounds the cutoff
// when max(1 - box / L_ref) > shrink_max (0.8 skin / (rc + skin)) or min(box) < box_min
// (2 (rc + skin) 1.02); that sets error bit 1, which the caller resolves by widening the
// skin and reallocating.
__global__ void nl_drift_kernel(const double* pos, const float4* canon, const SegInfo* seg, const int* atom_seg,
                                int total, int* flags, float shrink_max, float box_min) {
  bool over = false;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    const SegInfo& S = seg[atom_seg[g]];
    if (shrink_max > 0.f && g == S.base) {
      float shrink = 1.f - S.box[0] / S.L[0];
      shrink = fmaxf(shrink, 1.f - S.box[1] / S.L[1]);
      shrink = fmaxf(shrink, 1.f - S.box[2] / S.L[2]);
      const float bmin = fminf(S.box[0], fminf(S.box[1], S.box[2]));
      if (shrink > shrink_max || bmin < box_min) atomicOr(&flags[2], 1);
    }
    const float4 r = canon[g];
    const float rv[3] = {r.x, r.y, r.z};
    float d[3];
    if (S.tri) {
      float f[3];
      cell_frac01(S, pos + (size_t)g * 3, f);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        f[k] -= rv[k];
        f[k] -= rintf(f[k]);
      }
      cell_mul(S.H, f, d);
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float x = wrap_axis_f32(pos[(size_t)g * 3 + k], S.boxd[k], S.box[k]) * S.scale[k];
        float dd = x - rv[k];
        const float h = S.L[k] / 2.f;
        dd = fmod_floor(dd - h, S.L[k]) - h;
        d[k] = dd;
      }
    }
    const float n = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    over |= n > S.thr;
  }
  if (__ballot(over) && (threadIdx.x & 63) == 0) atomicOr(&flags[1], 1);
}

// flags[9] counts the rebuilds the drift test triggered since the flags were cleared
// flags[16] (not cleared per chunk) holds, as float bits, the moving average of the rebuild
// decisions (profile * (1 - 1/100) + update / 100, deepmd_jax/md.py:733)
__global__ void nl_decide_kernel(int* flags) {
  flags[0] = flags[1] != 0 ? 1 : 0;
  flags[9] += flags[0];
  flags[1] = 0;
  float* prof = reinterpret_cast<float*>(flags + 16);
  prof[0] = prof[0] * (1.f - 1.f / 100.f) + 1.f / 100.f * (float)flags[0];
}

__global__ void nl_set_flag_kernel(int* flags, int v) { flags[0] = v; }

__global__ void nl_clear_kernel(int* flags, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) flags[i] = 0;
}

#pragma clang fp contract(on)

static void fill_nl(System* s, NlArgs& a) {
  memset(&a, 0, sizeof(a));
  a.pos = s->d_pos;
  a.seg = s->d_seg;
  a.atom_seg = s->d_atom_seg;
  a.canon = s->d_canon;
  a.atom_cell = s->d_atom_cell;
  a.cell_count = s->d_cell_count;
  a.cell_start = s->d_cell_start;
  a.cell_fill = s->d_cell_fill;
  a.cell_atoms = s->d_cell_atoms;
  for (int j = 0; j < s->ntypes; ++j) {
    a.nbr[j] = s->d_nbr[j];
    a.nbr_cnt[j] = s->d_nbr_cnt[j];
    a.cap[j] = s->cap[j];
    a.rev[j] = s->d_rev[j];
  }
  a.ntypes = s->ntypes;
  a.NA = s->NA;
  a.ncell_total = s->ncell_total;
  const float cut = (float)s->rc_skin;
  a.cut_sq = cut * cut;
  a.flags = s->d_flags;
  a.lcap = s->nl_lds_list;
}

// Per-segment cell grids (edge >= cutoff with a margin; the 27-cell stencil needs >= 3
// cells per axis, otherwise the segment is scanned brute force) and the drift thresholds.
int32_t nlist_alloc_cells(System* s) {
  const double rc = s->model->rcut;
  int cell0 = 0;
  for (SegInfo& S : s->seg) {
    bool cells = true;
    for (int d = 0; d < 3; ++d) {
      const int nc = (int)(S.L[d] / (s->rc_skin * (1.0 + 1e-4) + 1e-3));
      S.nc[d] = nc;
      if (nc < 3) cells = false;
    }
    S.use_cells = cells ? 1 : 0;
    if (!cells) S.nc[0] = S.nc[1] = S.nc[2] = 1;
    S.cell0 = cell0;
    cell0 += cells ? S.nc[0] * S.nc[1] * S.nc[2] : 0;
    seg_set_threshold(S, s->rc_skin, rc, false);
  }
  s->ncell_total = std::max(1, cell0);
  s->seg_dirty = true;
  const size_t NA = (size_t)s->NA;
  DPMD_TRY(sys_ensure(s, (void**)&s->d_cell_count, sizeof(int) * s->ncell_total));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_cell_start, sizeof(int) * s->ncell_total));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_cell_fill, sizeof(int) * s->ncell_total));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_cell_atoms, sizeof(int) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_atom_cell, sizeof(int) * NA));
  DPMD_TRY(sys_ensure(s, (void**)&s->d_canon, sizeof(float4) * NA));
  DPMD_HIP(hipMemsetAsync(s->d_atom_cell, 0, sizeof(int) * NA, s->stream));
  return DPMD_OK;
}

// measure: record the per-type maximum counts in flags[4..]; conditional: every kernel
// returns immediately unless flags[0] was set by the drift test.
int32_t nlist_build(System* s, bool measure, bool conditional) {
  DPMD_TRY(sys_upload_segments(s));
  NlArgs a;
  fill_nl(s, a);
  a.measure = measure ? 1 : 0;
  a.conditional = conditional ? 1 : 0;
  if (!conditional) hipLaunchKernelGGL(nl_set_flag_kernel, dim3(1), dim3(1), 0, s->stream, s->d_flags, 1);
  const int NA = s->NA;
  hipEvent_t ev = nullptr;
  timer_begin(s, 0, &ev);
  hipLaunchKernelGGL(nl_prepare_kernel, dim3(std::max(1, std::min(1024, (a.ncell_total + 255) / 256))), dim3(256),
                     0, s->stream, a);
  hipLaunchKernelGGL(nl_canon_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, a);
  hipLaunchKernelGGL(nl_scan_kernel, dim3(s->nseg), dim3(1024), 0, s->stream, a);
  hipLaunchKernelGGL(nl_scatter_kernel, dim3((NA + 255) / 256), dim3(256), 0, s->stream, a);
  const size_t lds = sizeof(int) * 2 * a.lcap * 4;
  hipLaunchKernelGGL(nl_build_kernel, dim3((NA + 3) / 4), dim3(256), lds, s->stream, a);
  if (s->deterministic && !measure && s->d_rev[0])
    hipLaunchKernelGGL(nl_reverse_kernel, dim3((NA + 3) / 4), dim3(256), 0, s->stream, a);
  timer_end(s, 0, ev);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

// reverse index of the current rows (after a change of the deterministic setting)
int32_t nlist_reverse(System* s, bool conditional) {
  DPMD_TRY(sys_upload_segments(s));
  NlArgs a;
  fill_nl(s, a);
  a.conditional = conditional ? 1 : 0;
  hipLaunchKernelGGL(nl_reverse_kernel, dim3((s->NA + 3) / 4), dim3(256), 0, s->stream, a);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t nlist_drift_check(System* s, bool npt) {
  const int NA = s->NA;
  const int blocks = std::min(1024, (NA + 255) / 256);
  const double skin = s->rc_skin - s->model->rcut;
  const float shrink_max = npt ? (float)(0.8 * skin / s->rc_skin) : 0.f;
  const float box_min = npt ? (float)(2.0 * s->rc_skin * 1.02) : 0.f;
  hipLaunchKernelGGL(nl_drift_kernel, dim3(blocks), dim3(256), 0, s->stream, s->d_pos, s->d_canon, s->d_seg,
                     s->d_atom_seg, NA, s->d_flags, shrink_max, box_min);
  hipLaunchKernelGGL(nl_decide_kernel, dim3(1), dim3(1), 0, s->stream, s->d_flags);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

int32_t nlist_clear_flags(System* s) {
  hipLaunchKernelGGL(nl_clear_kernel, dim3(1), dim3(64), 0, s->stream, s->d_flags, 16);
  DPMD_HIP(hipGetLastError());
  return DPMD_OK;
}

}  // namespace dpmd

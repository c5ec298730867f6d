// Internal declarations shared by the HIP translation units of libdpmd_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "dpmd.h"

struct hipfftHandle_t;

namespace dpmd {

void set_error(const char* fmt, ...);

// PPPM long-range electrostatics on one device (pppm.hip): grid, half spectrum, influence
// function and hipFFT plans for a grid size (set-up reallocates when the size changes).
struct Pppm {
  int M[3] = {0, 0, 0};
  double beta = 0.0;
  hipfftHandle_t* fwd = nullptr;
  hipfftHandle_t* inv = nullptr;
  double* grid = nullptr;   // M0 M1 M2 charges, then the potential gradient phi
  void* spec = nullptr;     // M0 M1 (M2/2+1) complex doubles
  double* kfac = nullptr;   // influence function over the half spectrum
  double* part = nullptr;   // energy partials
  Pppm() = default;
  Pppm(const Pppm&) = delete;
  Pppm& operator=(const Pppm&) = delete;
  ~Pppm();
  void release();
  int32_t setup(const int* grid_size, double beta);
  // n charges xyz (n x 3) q (n) in the box d_box (device, 3 lengths): d_out[0] energy,
  // d_out[1..10) the 3 x 3 reciprocal virial; -dE/dx added to force (n x 3)
  int32_t compute(hipStream_t st, const double* d_box, const double* xyz, const double* q, int n, double* d_out,
                  double* force);
};

#define DPMD_HIP(x)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      ::dpmd::set_error("HIP error '%s' at %s:%d (%s)", hipGetErrorString(e_), __FILE__, \
                        __LINE__, #x);                                                 \
      return DPMD_E_HIP;                                                               \
    }                                                                                  \
  } while (0)

#define DPMD_TRY(x)              \
  do {                           \
    int32_t r_ = (x);            \
    if (r_ != DPMD_OK) return r_; \
  } while (0)

constexpr int kMaxT = DPMD_MAX_TYPES;
// DPLR per-segment record: [0] long-range energy, [1..10) its virial, then 64 x 9 block
// partials of the virial reduction
constexpr int kLrBlocks = 64;
constexpr int kLrStride = 10 + 9 * kLrBlocks;
constexpr int kKeBlocks = 64;   // most blocks per segment of the kinetic-energy reductions
constexpr int kWave = 64;

// Embedding-net feature layout.  An MFMA 16x16 accumulator tile holds, in lane
// (q = lane/16, n = lane%16), the rows 4q+i (i<4) of column n.  With features as rows
// and neighbours as columns, lane group q owns registers r = 4*tile + i of every layer.
// Register r of layer L in group q carries logical feature base[L][q] + r; groups take
// contiguous runs of first-layer features and their repeat-residual children, so the
// residual parent of register r is register r / Z in the same lane.
struct EmbedLayout {
  int w[3];         // widths of the three layers
  int z[3];         // z[1] = w1/w0, z[2] = w2/w1 (1 or 2)
  int tiles[3];     // 16-row MFMA tiles per layer
  int base[3][4];
  int size[3][4];
  // first width 4s + 1 (lane group 0 holds one feature more per layer) with the last tile of
  // layers 2 and 3 holding only group 0's extra features: the FP32 kernels fold those
  bool fold;
  bool init(const int* widths);
  // physical slot (16*tile + 4*q + i) -> logical feature or -1
  int logical(int layer, int slot) const;
};

// Per-net device tables (fp32), see embed.hip for the meaning of each block.
struct EmbedNetOffsets {
  int k0, b0, b1, b2, w1a, w2a, w2b, w1b;
  int f2a, f2b, f1b, f1a, s3a;  // fold-form tables (zero unless the layout folds)
  int total;
};
EmbedNetOffsets embed_net_offsets(const EmbedLayout& L);
void embed_pack_net(const EmbedLayout& L, const float* K0, const float* B0, const float* K1,
                    const float* B1, const float* K2, const float* B2, float* out);
// bf16 copies of the A-operand tables (w1a, w2a, w2b, w1b), packed per MFMA step
struct EmbedNetOffsetsH {
  int w1a, w2a, w2b, w1b, total;
};
EmbedNetOffsetsH embed_net_offsets_h(const EmbedLayout& L);
void embed_pack_net_h(const EmbedLayout& L, const float* packed_f32, uint16_t* out, bool fp16);

struct Model {
  int device = 0;
  int ntypes = 0;
  EmbedLayout lay;
  int fitw[3] = {0, 0, 0};
  int axis = 0;
  int wout = 0;  // embedding output width W
  double rcut = 0;
  double mean[kMaxT], istd[kMaxT], ebias[kMaxT];
  double nnbrs = 1, out_norm = 1;
  float* d_tbias = nullptr;        // (W)
  float* d_emb = nullptr;          // packed nets, ntypes^2 * off.total
  uint16_t* d_emb_h[2] = {nullptr, nullptr};  // bf16 / fp16 operand tables
  EmbedNetOffsets off;
  EmbedNetOffsetsH offh;
  // fitting weights per type: K (in,out) padded to out multiple of 16, transposes
  float* d_fit = nullptr;
  size_t fit_off[kMaxT][16];  // see api.hip
  size_t fit_total = 0;
  float fit_b3[kMaxT] = {0, 0, 0, 0};
  int fit_n = 0, fit_k = 0;  // output width (x16) and K stride (x32) of the hidden GEMMs
  uint16_t* d_fit_h[3] = {nullptr, nullptr, nullptr};  // BF16, FP16 copies of the kernels; BF16 low parts
  // tabulated embedding nets (compressed model): per net [interval][6 powers][16*tiles3]
  // in the layer-3 physical slot order, so one float4 load fills four accumulator rows
  int tab_n = 0;                   // intervals; 0 = evaluate the MLPs
  int tab_stride = 0;              // floats per net
  float tab_srmin[kMaxT * kMaxT], tab_delta[kMaxT * kMaxT];
  float* d_tab = nullptr;
  // atomic model: hidden fitting layers only, output for the centre types in nsel_mask;
  // 1 vector (Wannier centroids), 2 symmetric tensor (atomic_t2), 0 energy model
  int atomic = 0;
  int nsel_mask = 0;
};
void embed_pack_table(const EmbedLayout& L, int ngrids, const float* coeff, float* out);

struct Timer {
  bool on = false;
  double ms[8] = {0};
  int64_t count[8] = {0};
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[8];
};

// One independent periodic system inside a System batch (a request, a bead, a frame).
// Atoms of all segments are concatenated; inside a segment they are in stable
// type-sorted order.  Device-visible (copied to System::d_seg).
struct SegInfo {
  int base;            // first atom (global index)
  int natoms;
  int off[kMaxT + 1];  // type offsets inside the segment
  int row[kMaxT];      // fitting row of the segment's first atom of each type (relative to rowbase)
  int blk0, nblk;      // embedding blocks (4 centres per block)
  int cell0;           // first cell of the segment's cell grid
  int nc[3];
  int use_cells;
  int nimg;            // periodic images held (supercell of a small box), 1 otherwise
  float thr;           // neighbour-list drift threshold
  float L[3];          // neighbour-list reference box
  float scale[3];      // reference box / box
  float box[3];        // current box, model minimum image (general cell: perpendicular widths)
  double boxd[3];      // current box, FP64 wrap (general cell: perpendicular widths)
  unsigned long long seed;  // stochastic stream of the segment
  unsigned long long step;  // integrator steps taken (counter of the stream)
  // General (triclinic) cell, tri != 0: rows of H are the lattice vectors, x = s H and
  // s = x Hi for fractional s.  Positions wrap to s in [0, 1), pair vectors take the
  // fractional minimum image, and the neighbour list bins fractional coordinates.
  int tri;
  float H[9], Hi[9];
  double Hd[9], Hid[9];
  // images per axis of a supercell segment (0 when only their count is known) and the
  // primary cell's lengths (box / images), which the DPLR long-range term reads
  int kimg[3];
  double cellbox[3];
};

// s = x Hi and x = s H for row-vector coordinates and a row-major 3 x 3 matrix.
template <class T>
__host__ __device__ __forceinline__ void cell_mul(const T* M, const T x[3], T y[3]) {
#pragma unroll
  for (int e = 0; e < 3; ++e) y[e] = x[0] * M[e] + x[1] * M[3 + e] + x[2] * M[6 + e];
}

// Rebuild threshold of the drift test in FP32, as deepmd_jax/md.py:724-732 evaluates it:
// scale_max = max(L_ref / box) (x 1.02 under NPT), allowed = (f32(rc + skin) - f32(rc)
// scale_max) / 2, threshold = allowed - L_ref,min * f32(1e-6); a list is rebuilt when an
// atom's drift exceeds it.
__host__ __device__ __forceinline__ void seg_set_threshold(SegInfo& S, double rc_skin, double rc, bool npt) {
#pragma clang fp contract(off)  // one rounding per FP32 operation, no fused multiply-add
  float sc = S.scale[0] > S.scale[1] ? S.scale[0] : S.scale[1];
  sc = sc > S.scale[2] ? sc : S.scale[2];
  const float safe = npt ? sc * 1.02f : sc;
  const float allowed = ((float)rc_skin - (float)rc * safe) / 2.f;
  float lmin = S.L[0] < S.L[1] ? S.L[0] : S.L[1];
  lmin = lmin < S.L[2] ? lmin : S.L[2];
  S.thr = allowed - lmin * 1e-6f;
}

struct RowBase {
  int r[kMaxT];
};

// neighbour rows, their counts and reverse index, by type
struct NbrRef {
  const int* nbr[kMaxT];
  const int* cnt[kMaxT];
  const int* rev[kMaxT];
  int cap[kMaxT];
};

// NPT step parameters (one segment): target pressure, thermostat temperature, step,
// barostat time, cutoffs for the drift threshold, barostat chain length, group count,
// atoms, per-axis group (-1 fixed) and axes per group.
struct NptArgs {
  double pressure, kT, dt, tau_p, rc, rc_skin;
  int C, G, natoms;
  int cs, sy;  // barostat chain substeps (chain_steps_p, sy_steps_p)
  int gid[3], nper[3];
};

struct System {
  Model* model = nullptr;
  hipStream_t stream = nullptr;
  // trajectory frames leave the device on their own stream while later steps compute
  hipStream_t copy_stream = nullptr;
  double* h_traj = nullptr;     // pinned staging of recorded frames
  size_t h_traj_bytes = 0;
  std::vector<hipEvent_t> ev_frame, ev_copied;
  int ntypes = 0;
  int nseg = 0;
  int NA = 0;                   // atoms over all segments
  bool uniform = true;          // every segment has the same composition (replicas / beads)
  std::vector<SegInfo> seg;
  bool seg_dirty = true;        // host copy changed since the last upload
  int count[kMaxT];             // type counts summed over segments
  std::vector<int> perm;        // global sorted index -> global caller index
  std::vector<int> type_sorted; // per global sorted index
  std::vector<int> blk_seg;     // embedding block -> segment
  std::vector<double> mass;     // per global sorted atom
  // device state
  SegInfo* d_seg = nullptr;
  int* d_atom_seg = nullptr;    // (NA)
  int* d_blk_seg = nullptr;     // (nblocks)
  double* d_pos = nullptr;      // (NA*3), unwrapped
  double* d_mom = nullptr;
  double* d_mass = nullptr;     // (NA)
  double* d_pos_save = nullptr;
  double* d_mom_save = nullptr;
  float4* d_pos32 = nullptr;    // wrapped fp32 mirror for the model
  float* d_force = nullptr;     // (NA*3)
  double* d_scratch = nullptr;  // (NA) per-atom work
  // neighbour list
  bool nl_ready = false;
  bool nl_overflow = false;
  double rc_skin = 0, buffer_ratio = 1.2;
  int knbr[kMaxT];
  int cap[kMaxT];               // knbr - 1 columns
  int* d_nbr[kMaxT] = {nullptr};
  int* d_nbr_cnt[kMaxT] = {nullptr};
  // deterministic forces: per skin slot the pair gradient dE/dx_ij of the last backward
  // (NA, sum cap, 3), per atom its centre part (NA, 3), and per row entry the entry's slot in
  // the neighbour's own row (d_rev[j]: (NA, cap_j), -1 when absent)
  bool deterministic = false;
  bool fit_fused = true;        // fitting hidden layers in one kernel (fit_hidden_kernel)
  float* d_pgrad = nullptr;
  float* d_fcen = nullptr;
  int* d_rev[kMaxT] = {nullptr};
  float4* d_canon = nullptr;    // canonical coordinates of the last build (reference_position)
  int ncell_total = 0;
  int* d_cell_count = nullptr;
  int* d_cell_start = nullptr;
  int* d_cell_fill = nullptr;
  int* d_cell_atoms = nullptr;
  int* d_atom_cell = nullptr;
  int nl_lds_list = 512;
  int* d_flags = nullptr;       // [0] rebuild flag, [1] drift flag, [2] error bits,
                                // [4..4+kMaxT) max count per type, [8] measure-list overflow,
                                // [9] drift-triggered rebuilds since the last clear
  int64_t nl_rebuilds = 0;      // drift-triggered rebuilds since the list was allocated
  // model work buffers
  int rowbase[kMaxT + 1];       // fitting rows per type (all segments), padded to 64
  int mrows = 0;
  int nblocks = 0;              // embedding blocks over all segments
  float* d_T = nullptr;         // (NA, 10, W)
  void* d_G = nullptr;          // (mrows, A*W)
  float* d_dG = nullptr;        // (mrows, A*W)
  int dg16 = 0;                 // d_dG holds FP16 values (written by a plain 16-bit fitting)
  float* d_fitbuf = nullptr;    // 7 activations (mrows, fstride)
  float* d_eatom = nullptr;     // (mrows)
  double* d_partial = nullptr;  // per-block virial partials (nblocks, 9)
  double* d_red = nullptr;      // 16 doubles per segment
  int* d_pairs = nullptr;       // (NA)
  int fpad = 0;                 // fitting output width
  int fstride = 0;              // fitting activation row stride
  bool force_valid = false;
  int force_prec = 0;           // precision the valid forces were computed in
  // integrator
  double* d_thermo = nullptr;   // 64 doubles per segment
  double* d_traj = nullptr;     // device trajectory staging: positions then velocities,
                                // (steps, NA, 3) each, caller order
  int* d_perm = nullptr;        // (NA) sorted index -> caller index
  double* d_nm = nullptr;       // PIMD normal modes
  double* d_kepart = nullptr;   // kinetic-energy partial sums (nseg, kKeBlocks, 3)
  float* d_mobile = nullptr;    // (NA) 1 mobile, 0 fixed (fixed_indices)
  bool has_fixed = false;
  int n_mobile = 0;
  double* d_npt = nullptr;      // NPT barostat state (64 doubles) and its chunk-entry copy
  bool npt_ready = false;       // NPT state initialised for the current layout
  int npt_kind = 0;             // routine that initialised it (DPMD_NPT_NH or DPMD_PIMD_NPT)
  double* d_box_traj = nullptr; // per-step boxes of the last recorded chunk (steps, nseg, 3)
  int box_traj_steps = 0;
  uint64_t rng_counter = 0;
  // DPLR (dpmd_system_set_dplr): centroid model, charges and one PPPM per segment.  Charge
  // layout: segment by segment, the segment's atoms in sorted order, then its centroids in
  // sorted atom order (block r holds q_n[r] charges from q_base[r]).
  Model* wc_model = nullptr;
  Pppm* pppm = nullptr;
  int n_wc = 0;
  int* d_wc_atom = nullptr;     // (n_wc) sorted atom of each centroid
  int* d_wc_row = nullptr;      // (n_wc) its fitting row
  int* d_atom_wc = nullptr;     // (NA) centroid of each atom or -1
  int dplr_grid[3] = {0, 0, 0};
  double dplr_beta = 0.0;
  int* d_atom_q = nullptr;      // (NA) charge index of each atom
  int* d_wc_q = nullptr;        // (n_wc) charge index of each centroid
  int* d_wc_seg = nullptr;      // (nseg + 1) first centroid of each segment
  std::vector<int> q_base, q_n; // charge block of each segment
  double* d_qxyz = nullptr;     // (NA + n_wc, 3) charge positions
  double* d_q = nullptr;        // (NA + n_wc) charges
  double* d_qf = nullptr;       // (NA + n_wc, 3) PPPM forces
  double* d_lr = nullptr;       // per segment kLrStride doubles: long-range energy, virial, partials
  float* d_wc_g = nullptr;      // (NA, 3) dE_lr / d(centroid) of each atom's centroid (0 if none)
  double* d_partial_wc = nullptr;  // virial partials of the centroid backward (not used)
  float* d_wc_fit = nullptr;    // 10 activation buffers (mrows x fit_k) of the centroid model,
                                // zeroed once: columns past the fitting width stay zero
  std::vector<int> wc_perm;     // centroid k -> rank among the selected atoms in caller order
  const Model* atomic_maps_of = nullptr;  // model whose selection the maps above follow
  // embedding-backward hooks, set only while the centroid model's backward runs: the extra
  // dE/dT3[c][w] = wc_norm wc_g[atom][c] wc_y3[row][w], and the virial partial output
  const float* wc_g = nullptr;
  const float* wc_y3 = nullptr;
  int wc_ldy = 0;
  float wc_norm = 0.f;
  double* partial_out = nullptr;
  Timer timer;
  std::vector<std::pair<void**, size_t>> caps;  // allocation capacities (grow-only buffers)
  ~System();
};

// grow-only device allocation: keeps the buffer when it is large enough
int32_t sys_ensure(System* s, void** p, size_t bytes);
int32_t sys_upload_segments(System* s);

// kernels / launchers (return DPMD_OK or error)
int32_t nlist_build(System* s, bool measure, bool conditional);
int32_t nlist_drift_check(System* s, bool npt = false);
int32_t nlist_reverse(System* s, bool conditional);
int32_t force_gather(System* s);
int32_t nlist_alloc_cells(System* s);
int32_t nlist_clear_flags(System* s);
// precision: DPMD_PREC_FP32 / BF16 / FP16 of the embedding GEMMs; g16 (forward): the format
// the descriptor is stored in for the fitting (0 FP32, 1 BF16, 2 FP16; -1: as precision)
int32_t embed_launch(System* s, bool backward, int precision, int g16 = -1);
// precision: FP32 / BF16 / FP16, or DPMD_PREC_BF16X3 (the three-product BF16 split)
int32_t fit_launch(System* s, int precision);
// the embedding and fitting precisions of a DPMD_PREC_* code, and whether it is one
inline int prec_embed(int p) { return p == DPMD_PREC_BF16X3 || p == DPMD_PREC_BF16_FIT32 ? DPMD_PREC_BF16 : p; }
inline int prec_fit(int p) { return p == DPMD_PREC_BF16_FIT32 ? DPMD_PREC_FP32 : p; }
inline bool prec_valid(int p) { return p >= DPMD_PREC_FP32 && p <= DPMD_PREC_BF16_FIT32; }
// atomic (centroid) model hidden layers: forward into the activation buffers, and the backward
// from dE/d(last hidden layer) (buffer DY3) to dG; buffers laid out by fit_atomic_buffers
struct AtomicFitBufs {
  float *Y1, *T2, *Y2, *T3, *Y3, *DY3, *DZ3, *DY2, *DZ2, *DZ1;
  int FK;
};
int32_t fit_atomic_buffers(System* s, const Model* M, AtomicFitBufs& b);
int32_t fit_atomic_fwd(System* s, const AtomicFitBufs& b);
int32_t fit_atomic_bwd(System* s, const AtomicFitBufs& b);
int32_t dplr_forces(System* s);
// maps and buffers of an atomic model's outputs on the system's layout (selected atoms, their
// fitting rows, caller-order ranks)
int32_t atomic_setup(System* s, Model* W);
int32_t dplr_add_energy(System* s);
int32_t md_zero_force(System* s);
int32_t md_kick(System* s, double dt_half);
int32_t md_drift(System* s, double dt);
int32_t md_write_pos32(System* s);
int32_t md_traj_record(System* s, double* tpos, double* tvel);
int32_t md_npt_init(System* s, const NptArgs& a, bool refresh_only);
int32_t md_npt_baro_half(System* s, const NptArgs& a);
int32_t md_npt_box(System* s, const NptArgs& a, int phase);
int32_t md_npt_particles(System* s, double dt, bool move);
int32_t md_box_record(System* s, double* out);
int32_t md_pnpt_init(System* s, const NptArgs& a);
int32_t md_pnpt_dudv(System* s, const NptArgs& a);
int32_t md_pnpt_kick(System* s, const NptArgs& a);
int32_t md_pnpt_move(System* s, const NptArgs& a, double h);
int32_t md_pnpt_box_o(System* s, const NptArgs& a, double gamma, uint64_t seed, uint64_t counter);
int32_t md_pnpt_il2(System* s, const NptArgs& a);
int32_t md_pnpt_drift_ring(System* s, const NptArgs& a, const double* C, const double* freqs, double h);
int32_t md_mask_forces(System* s);
int32_t md_mask_momenta(System* s);
int32_t md_remove_com(System* s);
int32_t md_reduce_launch(System* s);
int32_t md_ke_launch(System* s, double* d_out);
int32_t md_langevin_o(System* s, double dt, double gamma, double kT, uint64_t seed, uint64_t step);
int32_t md_nh_half(System* s, double kT, double dt, double tau, int chain, int cs, int sy);
int32_t md_pimd_free_ring(System* s, const double* cmat, const double* freqs, double dt_half);
int32_t md_pimd_o(System* s, const double* cmat, const double* c1, const double* c2, uint64_t seed,
                  uint64_t counter);

void timer_begin(System* s, int which, hipEvent_t* ev);
void timer_end(System* s, int which, hipEvent_t ev);

}  // namespace dpmd

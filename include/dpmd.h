/*
 * dpmd.h -- C ABI of libdpmd_hip.so, the MI355X (gfx950) engine for deepmd-jax's
 * Deep Potential se_a energy/force/virial evaluation and the MD step around it.
 *
 * Plain C types only: host pointers, sizes, int32 status codes.  Every entry point
 * returns 0 on success and a negative DPMD_E* code on failure; dpmd_last_error()
 * then holds a thread-local message.  Device-side conditions (neighbour-list
 * overflow, NaN/Inf) are reported as the reference's bit-coded error word
 * (deepmd_jax/md.py:616-642): 2 = neighbour buffer overflow, 16 = NaN or Inf.
 *
 * Arrays given in "caller order" follow the caller's atom order; the library keeps
 * its device state in the reference's stable type-sorted order
 * (deepmd_jax/utils.py:329-342) and permutes at the boundary.
 *
 * The reference is pure Python on JAX; it has no FFI.  Each entry point below names
 * the reference function it replaces.  The Python host package
 * (deepmd-jax_amd/deepmd_jax) binds them with ctypes.
 */
#ifndef DPMD_H
#define DPMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPMD_MAX_TYPES 4

#define DPMD_OK 0
#define DPMD_E_ARG -1      /* invalid argument (ValueError in the reference) */
#define DPMD_E_HIP -2      /* HIP runtime error */
#define DPMD_E_STATE -3    /* call out of order (e.g. no neighbour list yet) */
#define DPMD_E_UNSUPPORTED -4 /* model topology the kernels do not implement */

/* Arithmetic of the embedding / fitting GEMMs.  Switch, environment matrix and
 * descriptor contraction are always FP32; energy / virial reductions FP64. */
#define DPMD_PREC_FP32 0
#define DPMD_PREC_BF16 1
#define DPMD_PREC_FP16 2
/* Mixed forms: BF16 embedding GEMMs with fitting GEMMs as three BF16 products (operands
 * split into BF16 high and low parts: hi*hi + hi*lo + lo*hi, FP32 accumulation), or with
 * FP32 fitting GEMMs.  The fitting net's operand rounding sets the energy bias whose change
 * along a trajectory shows as NVE drift; the split keeps its GEMMs on BF16 MFMAs. */
#define DPMD_PREC_BF16X3 3
#define DPMD_PREC_BF16_FIT32 4

/* Integrators for dpmd_md_run (deepmd_jax/md.py:313-383). */
#define DPMD_NVE 0          /* velocity Verlet, deepmd_jax/routine.py:139-144 */
#define DPMD_NVT_NH 1       /* Nose-Hoover chain, deepmd_jax/routine.py:147-158 */
#define DPMD_NVT_LANGEVIN 2 /* BAOAB, deepmd_jax/routine.py:46-107 */
#define DPMD_PIMD_PILE_L 3  /* ring-polymer BAOAB + PILE-L, deepmd_jax/pimd.py:114-207 */
#define DPMD_NPT_NH 4       /* MTK Nose-Hoover barostat + thermostat, deepmd_jax/routine.py:161-423 */
#define DPMD_PIMD_NPT 5     /* path-integral NPT: MTK on the centroid, exact internal ring modes,
                               PILE-L and a Langevin barostat, deepmd_jax/pimd.py:297-518 */

typedef struct dpmd_model* dpmd_model_t;
typedef struct dpmd_system* dpmd_system_t;

/* Host description of an se_a energy model, i.e. model.params + variables['params']
 * of a deepmd-jax model.pkl (deepmd_jax/utils.py:225-237, deepmd_jax/train.py:288-304).
 * Dense kernels are flax layout (in, out), row-major fp32.  Embedding net k = i*ntypes + j
 * maps centre type i and neighbour type j (deepmd_jax/dpmodel.py:50-51). */
typedef struct {
  int32_t ntypes;
  int32_t embed_widths[3];
  int32_t fit_widths[3];
  int32_t axis;
  double rcut;
  double sr_mean[DPMD_MAX_TYPES];
  double sr_std[DPMD_MAX_TYPES];
  double nnbrs;
  double ebias[DPMD_MAX_TYPES];
  double out_norm;
  const float* tbias;                                              /* (W) */
  const float* emb_kernel[DPMD_MAX_TYPES * DPMD_MAX_TYPES][3];
  const float* emb_bias[DPMD_MAX_TYPES * DPMD_MAX_TYPES][3];
  const float* fit_kernel[DPMD_MAX_TYPES][4];
  const float* fit_bias[DPMD_MAX_TYPES][4];
  const float* fit_dt[DPMD_MAX_TYPES][3];                          /* NULL where no residual */
  /* Compressed (tabulated) embedding nets, variables['compress_var'] of a model that went
   * through compress_model (deepmd_jax/utils.py:250-297).  tab_ngrids > 0 selects them:
   * net k then evaluates the piecewise quintic tab_coeff[k] (tab_ngrids, 6, W), row c the
   * coefficient of x0^(5-c), over [tab_srmin[k], tab_srmax[k]] (deepmd_jax/utils.py:126-151),
   * and emb_kernel / emb_bias may be NULL.  tab_ngrids = 0: the MLPs above are evaluated. */
  int32_t tab_ngrids;
  float tab_srmin[DPMD_MAX_TYPES * DPMD_MAX_TYPES];
  float tab_srmax[DPMD_MAX_TYPES * DPMD_MAX_TYPES];
  const float* tab_coeff[DPMD_MAX_TYPES * DPMD_MAX_TYPES];
  /* Atomic models (params['atomic'], deepmd_jax/dpmodel.py:83-99): 1 = type 'atomic' (a vector
   * per atom, the Wannier-centroid model of DPLR), 2 = type 'atomic_t2' (a symmetric 3x3 tensor
   * per atom), 0 = energy model.  The fitting nets are hidden layers only (fit_kernel[t][3] unused,
   * fit_widths[2] == embed_widths[2]); an atom of a type in nsel_mask (bit t) gets
   * out_norm * sum_w f_w Y[:, w] with Y = T3 (vector) or T6 + 3to6(T3) + TW (1,1,1,0,0,0)
   * mapped 6 -> 9 (tensor).  Nets of centre types outside the mask are not evaluated (any weights
   * of the right shapes may be passed for them). */
  int32_t atomic;
  int32_t nsel_mask;
} dpmd_model_desc;

/* Integrator parameters for dpmd_md_run. Units: fs, eV, Angstrom. */
typedef struct {
  int32_t routine;      /* DPMD_NVE, ... */
  int32_t precision;    /* DPMD_PREC_* for the force evaluations */
  double dt;
  double kT;            /* thermostat temperature (eV) */
  double gamma;         /* Langevin friction (1/fs) of NVT_langevin; DPMD_PIMD_NPT: of the box
                           momenta (1 / tau_p in deepmd_jax/md.py:374) */
  double tau;           /* Nose-Hoover relaxation time (fs) */
  int32_t chain_length; /* Nose-Hoover chain length */
  int32_t record_every; /* >0: store positions/velocities every step into traj buffers */
  uint64_t seed;        /* stochastic integrators: counter-based normal stream */
  double pimd_lambda;   /* PILE-L internal-mode damping scale */
  /* DPMD_NPT_NH (deepmd_jax/md.py:357-366, deepmd_jax/routine.py:161-423) and DPMD_PIMD_NPT
   * (deepmd_jax/md.py:368-381; the chain length is not used there): */
  double pressure;      /* target pressure (eV / Angstrom^3) */
  double tau_p;         /* barostat relaxation time (fs) */
  int32_t chain_length_p; /* barostat chain length */
  int32_t couple[3];    /* coupled group of each axis (0, 1, 2; groups numbered from 0 in
                           first-appearance order) or -1 for a fixed axis (couple_axes) */
  int32_t remove_com;   /* nonzero: remove the centre-of-mass velocity of the mobile atoms after
                           every step (Simulation(remove_com_motion=True), deepmd_jax/md.py:786-790) */
  int32_t chain_steps;  /* Nose-Hoover chain substeps per half step of the thermostat and their */
  int32_t sy_steps;     /* Suzuki-Yoshida order (1, 3, 5, 7), jax_md's chain_steps / sy_steps;  */
  int32_t chain_steps_p;/* _p: the NPT barostat chain.  0 reads as 1 (deepmd_jax/md.py:80-87)     */
  int32_t sy_steps_p;
} dpmd_md_params;

const char* dpmd_last_error(void);
int32_t dpmd_version(void);

/* Replaces load_model's device_put of variables (deepmd_jax/utils.py:230-237). */
int32_t dpmd_model_create(const dpmd_model_desc* desc, int32_t device, dpmd_model_t* out);
int32_t dpmd_model_destroy(dpmd_model_t model);

/* One simulated system of natoms atoms x nreplica replicas (ring-polymer beads share
 * the composition).  Replaces the state set-up of Simulation.__init__
 * (deepmd_jax/md.py:61-283).  type_idx: (natoms) in [0, ntypes).  Replica r is segment r
 * of the ragged layout below. */
int32_t dpmd_system_create(dpmd_model_t model, int32_t natoms, int32_t nreplica,
                           const int32_t* type_idx, dpmd_system_t* out);
/* nseg independent periodic systems ("segments") of uneven size evaluated and integrated
 * together: every kernel launch covers all segments.  seg_natoms (nseg); type_idx
 * (sum seg_natoms) caller order, segments concatenated.  Generalises the per-frame batch
 * of evaluate() (deepmd_jax/train.py:703-788) to frames of different sizes. */
int32_t dpmd_system_create_ragged(dpmd_model_t model, int32_t nseg, const int32_t* seg_natoms,
                                  const int32_t* type_idx, dpmd_system_t* out);
/* Re-lay the system out for a new set of segments, reusing its device buffers when they
 * are large enough (grow-only).  All state (positions, velocities, masses, boxes, thermostat,
 * neighbour list) is reset and must be set again.  No reference counterpart: the reference
 * re-traces its jit functions per shape. */
int32_t dpmd_system_reshape(dpmd_system_t sys, int32_t nseg, const int32_t* seg_natoms,
                            const int32_t* type_idx);
int32_t dpmd_system_destroy(dpmd_system_t sys);

/* Orthorhombic box lengths (3), applied to every segment. */
int32_t dpmd_system_set_box(dpmd_system_t sys, const double* box3);
/* Per-segment box lengths (nseg, 3). */
int32_t dpmd_system_set_boxes(dpmd_system_t sys, const double* boxes);
/* Per-segment general (triclinic) cells (nseg, 3, 3), rows the lattice vectors, as the
 * reference's (3,3) boxes (deepmd_jax/utils.py:20-27 shift, deepmd_jax/md.py:208-215): a
 * diagonal cell is the orthorhombic box; otherwise pair vectors take the fractional minimum
 * image, every perpendicular width must exceed 2 (rcut + skin) (dpmd_system_set_images holds
 * the supercell of a smaller cell), and dpmd_system_get_boxes reports the widths.  Setting a
 * different general cell drops the neighbour list (call dpmd_nlist_allocate again).  NPT,
 * path-integral NPT and DPLR reject general cells, as the reference does
 * (deepmd_jax/md.py:391-392, deepmd_jax/routine.py:161-170). */
int32_t dpmd_system_set_cells(dpmd_system_t sys, const double* cells);
/* Per-segment count of periodic images held (nseg): a segment that holds the k^3 images
 * of a box smaller than 2 rc (the supercell equivalent of the reference's lattice-image
 * branch, deepmd_jax/utils.py:66-78) lists its atoms image-major; energies, virials,
 * kinetic energies and thermostat work are then reported per image and all images draw
 * the same thermostat noise. */
int32_t dpmd_system_set_images(dpmd_system_t sys, const int32_t* nimg);
/* Per-segment images per axis (nseg, 3): dpmd_system_set_images with their product, and the
 * primary cell (box / images per axis) on which a DPLR term of the segment evaluates its
 * long-range energy (the reference's lattice path keeps PPPM on the cell, deepmd_jax/md.py:404-415). */
int32_t dpmd_system_set_image_grid(dpmd_system_t sys, const int32_t* k3);
/* Per-segment stochastic stream (nseg each, either may be NULL): the Langevin normals of
 * a segment depend only on (params.seed, seed, step, atom), so a segment draws the same
 * stream alone or in any batch.  steps counts integrator steps already taken. */
int32_t dpmd_system_set_rng(dpmd_system_t sys, const uint64_t* seeds, const uint64_t* steps);
int32_t dpmd_system_get_rng(dpmd_system_t sys, uint64_t* seeds, uint64_t* steps);
/* (sum seg_natoms, 3) caller order.  setPosition / getPosition (deepmd_jax/md.py:969-1014). */
int32_t dpmd_system_set_positions(dpmd_system_t sys, const double* pos);
int32_t dpmd_system_get_positions(dpmd_system_t sys, double* pos);
/* Velocities (Angstrom/fs); setVelocity / getVelocity (deepmd_jax/md.py:1016-1031). */
int32_t dpmd_system_set_velocities(dpmd_system_t sys, const double* vel);
int32_t dpmd_system_get_velocities(dpmd_system_t sys, double* vel);
/* Per-atom masses (sum seg_natoms) in eV fs^2 / Angstrom^2, caller order.  Set masses
 * before velocities (momenta are stored). */
int32_t dpmd_system_set_masses(dpmd_system_t sys, const double* mass);
/* Fixed atoms (fixed_indices of deepmd_jax/md.py:221-234): fixed (sum seg_natoms) caller order,
 * nonzero = fixed, NULL = none.  Fixed atoms get zero force (their coordinates carry no
 * gradient, :428-430), keep their positions, have their momenta zeroed after each NVE / NVT
 * step and receive no Langevin noise; the Nose-Hoover chain counts 3 x mobile degrees of
 * freedom (deepmd_jax/routine.py:110-136).  NVE, NVT and NVT_langevin only. */
int32_t dpmd_system_set_fixed(dpmd_system_t sys, const int32_t* fixed);

/* Typed neighbour list at cutoff rc_skin.  Replaces typed_neighbor_list's allocate_fn
 * (deepmd_jax/md_utils.py:376-421): measures the per-type maximum neighbour count,
 * sizes each row to knbr_j = 1 + floor(k_j*ratio + max(20-k_j,0)*max(ratio-1.2,0)),
 * builds.  knbr_override (ntypes) skips the measurement when non-NULL.  knbr_out (ntypes). */
int32_t dpmd_nlist_allocate(dpmd_system_t sys, double rc_skin, double buffer_ratio,
                            const int32_t* knbr_override, int32_t* knbr_out);
/* Rebuild at the allocated capacity; update_fn (deepmd_jax/md_utils.py:424-429).
 * *overflow = 1 when some row needed more than knbr_j - 1 entries. */
int32_t dpmd_nlist_update(dpmd_system_t sys, int32_t* overflow);
/* Export rows for neighbour type j as nbrs_nm (deepmd_jax/md_utils.py:431-459):
 * out (sum seg_natoms, knbr_j - 1) int32, centres in type-sorted order per segment,
 * entries are ascending type-local indices (inside the segment) padded with the
 * segment's atom count. */
int32_t dpmd_nlist_export(dpmd_system_t sys, int32_t j, int32_t* out);
/* Current capacities knbr (ntypes). */
int32_t dpmd_nlist_capacity(dpmd_system_t sys, int32_t* knbr);
/* 1 when the last build truncated a row at its capacity (the reference's did_buffer_overflow,
 * deepmd_jax/md_utils.py:455-458), else 0.  dpmd_compute refuses such a list (DPMD_E_STATE)
 * until dpmd_nlist_allocate measures new capacities. */
int32_t dpmd_nlist_overflow(dpmd_system_t sys, int32_t* overflow);

/* Energy, forces and virial of the current positions.  Replaces
 * DPModel.energy_and_force (deepmd_jax/dpmodel.py:106-108) and the strain derivative
 * of iso_pressure (deepmd_jax/md_utils.py:139-159).  Needs an up-to-date neighbour
 * list (dpmd_nlist_allocate / _update).
 * energy (nseg); force (sum seg_natoms, 3) caller order; virial (nseg, 9)
 * row-major, virial = -sum_pairs x_ij (outer) dE/dx_ij; atom_energy (sum seg_natoms).
 * Any output pointer may be NULL. */
int32_t dpmd_compute(dpmd_system_t sys, int32_t precision, double* energy, double* force,
                     double* virial, double* atom_energy);

/* Deterministic forces on (1) or off (0, the default).  On, the backward writes each pair's
 * gradient dE/dx_ij once (by its slot in the centre's neighbour row) and a gather sums, per
 * atom and in row order, its centre part and the gradients its neighbours stored, through a
 * reverse index built with every neighbour-list rebuild; two evaluations of the same
 * positions then give bit-identical forces.  Off, neighbour contributions are FP32 device
 * atomics, whose order varies from run to run.  Replaces the scatter-add that jax.grad makes
 * of the neighbour gather (deepmd_jax/utils.py:80); the reference runs it in whatever order
 * XLA chooses.  DPLR systems keep atomics for their long-range and centroid terms. */
int32_t dpmd_system_set_deterministic(dpmd_system_t sys, int32_t on);

/* Allocation-free variant of dpmd_compute (same energy, force and virial outputs, no
 * per-atom energies).  It has no counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call makes no device allocation: its work
 * buffers (T, G, dE/dG, fitting activations, partial sums, staged outputs) live in one
 * 256 MiB library-owned device array per GPU, held from the library's load until the
 * process ends, instead of the system's own buffers.  A system whose work needs more
 * (about 10,000 atoms of the 25/50/100, 240x3, axis-16 model) gets DPMD_E_ARG; DPLR
 * systems get DPMD_E_UNSUPPORTED.  Every call on a GPU shares that GPU's array and
 * leaves its contents behind; calls that overlap on one GPU would race on it, so the
 * calls of one process are serialised. */
int32_t dpmd_compute_noalloc(dpmd_system_t sys, int32_t precision, double* energy, double* force,
                             double* virial);

/* nsteps MD steps fully on the device: per step the drift test and conditional
 * neighbour rebuild (deepmd_jax/md.py:693-741), the integrator step with one force
 * evaluation, minimum-image unwrap (deepmd_jax/md_utils.py:162-169).  Replaces the
 * jit(lax.scan(inner_step)) chunk of Simulation.run (deepmd_jax/md.py:759-804).
 * traj_pos / traj_vel: (nsteps, sum seg_natoms, 3) caller order or NULL.
 * *error_code accumulates the reference's bit codes. */
int32_t dpmd_md_run(dpmd_system_t sys, const dpmd_md_params* p, int32_t nsteps,
                    double* traj_pos, double* traj_vel, uint32_t* error_code);
/* Ring-polymer constants for DPMD_PIMD_PILE_L and DPMD_PIMD_NPT (deepmd_jax/pimd.py:22-44,92-111,160-179):
 * C (P*P, row-major, primitive <- normal mode), freqs (P), and the per-mode O-step
 * factors c1 = exp(-gamma_k dt), c2 = sqrt(P kT (1 - c1^2)) (P).  P = nreplica. */
int32_t dpmd_pimd_setup(dpmd_system_t sys, const double* C, const double* freqs, const double* c1,
                        const double* c2);
/* Kinetic energy (nseg) in FP64 from the device momenta. */
int32_t dpmd_kinetic_energy(dpmd_system_t sys, double* ke);
/* Thermostat work / chain state scalar used by the reference invariants (nseg):
 * thermostat_work of NVTLangevinStateWithWork (deepmd_jax/routine.py:27-107), the chain
 * energy of nvt_nose_hoover_invariant (deepmd_jax/routine.py:9-25); after NPT steps also
 * the barostat chain, P V and box kinetic energy of npt_nose_hoover_invariant
 * (deepmd_jax/routine.py:426-456), so that PE + KE + this is conserved. */
int32_t dpmd_thermostat_energy(dpmd_system_t sys, double* e);
/* Whole thermostat state, 64 doubles per segment, to carry a segment into another system
 * (no reference counterpart: the reference keeps it in its scan carry). */
int32_t dpmd_thermostat_state_get(dpmd_system_t sys, double* state);
int32_t dpmd_thermostat_state_set(dpmd_system_t sys, const double* state);

/* Profiling: enable HIP-event timing of each kernel class on the system's stream;
 * which = 0 nlist, 1 embed_fwd, 2 fit, 3 embed_bwd, 4 integrator.  ms = summed event
 * time, count = launches timed. */
int32_t dpmd_timing_enable(dpmd_system_t sys, int32_t on);
int32_t dpmd_timing_get(dpmd_system_t sys, int32_t which, double* ms, int64_t* count);
/* Rebuilds triggered by the drift test inside dpmd_md_run since the last dpmd_nlist_allocate
 * (the update count behind the reference's neighbour-list profile, deepmd_jax/md.py:731). */
/* Moving average (weight 1/100 per step) of the per-step rebuild decisions, the NbrUpdateRate
 * column of Simulation(debug=True) (deepmd_jax/md.py:586-592,733). */
int32_t dpmd_nlist_update_rate(dpmd_system_t sys, double* rate);
int32_t dpmd_nlist_rebuilds(dpmd_system_t sys, int64_t* count);
/* Mean neighbours within rcut per centre of the last evaluation (algorithmic FLOP count). */
int32_t dpmd_pair_count(dpmd_system_t sys, double* pairs_within_rcut);
int32_t dpmd_synchronize(dpmd_system_t sys);

/* Current box lengths (nseg, 3): state.box of the reference NPT state
 * (deepmd_jax/routine.py:189-194), equal to the set box outside NPT. */
int32_t dpmd_system_get_boxes(dpmd_system_t sys, double* boxes);
/* Per-step boxes (steps, nseg, 3) of the last dpmd_md_run that recorded a trajectory: the
 * box trajectory of Simulation.run (deepmd_jax/md.py:846,924). */
int32_t dpmd_md_box_trajectory(dpmd_system_t sys, double* boxes);
/* DPLR long-range term (deepmd_jax/md.py:404-415, deepmd_jax/utils.py:174-222,355-366): wc_model (an
 * atomic model on the same device, with the system model's embedding widths and axis) predicts a
 * Wannier centroid x_i + dx_i for every atom of its selected types; the PPPM energy of charge
 * q_atoms[type] at every atom and q_wc[type] at every centroid (beta, grid[3]) joins the energy, and
 * its gradient through the centroid model joins the forces of every later evaluation and MD step.
 * Every segment (a frame, a bead) carries its own PPPM in its own box, as the reference vmaps
 * the per-configuration energy over beads (deepmd_jax/md.py:436-440); segments hold no periodic
 * images, boxes are orthorhombic and all share the grid.  The virial includes the long-range
 * term (PPPM reciprocal virial, the centroid model's strain response and the centroid offsets),
 * so NPT and path-integral NPT run with it.  wc_model NULL removes the term. */
int32_t dpmd_system_set_dplr(dpmd_system_t sys, dpmd_model_t wc_model, const double* q_atoms, const double* q_wc,
                             double beta, const int32_t* grid);
/* Outputs of an atomic model (the system's model has desc.atomic != 0) at the current positions:
 * for every atom of the selected types, segment by segment in caller order, 3 components
 * (vector) or 9 (symmetric tensor, row-major) in FP64 (FP32 arithmetic).  Replaces the
 * atomic branch of DPModel.apply / evaluate (deepmd_jax/dpmodel.py:83-99, train.py:765-788). */
int32_t dpmd_atomic_predict(dpmd_system_t sys, double* out);
/* Of the last evaluation with DPLR set: the long-range energy of every segment (nseg) and the
 * centroid positions (n_wc x 3, segment by segment, caller order among the atoms of the selected
 * types, as wc_predict returns them), either may be NULL. */
int32_t dpmd_dplr_get(dpmd_system_t sys, double* e_long_range, double* centroids);
/* PPPM long-range energy and forces of n point charges in an orthorhombic periodic box, the
 * get_p3mlr_fn of deepmd_jax/utils.py:177-222 (TSC assignment onto grid[0] x grid[1] x grid[2],
 * E = sum_k K(k)|S(k)|^2 in eV for charges in e and lengths in Angstrom), evaluated on GPU
 * `device` in FP64.  xyz (n x 3), q (n); force (n x 3, -dE/dx) and virial (3 x 3, -dE/dstrain
 * with the charges scaling with the box) may be NULL. */
int32_t dpmd_pppm_compute(int32_t device, const double* box3, double beta, const int32_t* grid, int32_t n,
                          const double* xyz, const double* q, double* energy, double* force, double* virial);
/* NPT barostat state (64 doubles): [0..3) log box scale per group, [3..6) box momenta,
 * [6..9) box masses, [9..12) dU/deps per group, [12..15) reference box, [16..24) chain
 * positions, [24..32) chain momenta, [32..40) chain masses, [43] barostat invariant terms
 * (CachedNPTNoseHooverState, deepmd_jax/routine.py:161-194).  After DPMD_PIMD_NPT steps
 * (PIMDNPTLangevinState, deepmd_jax/pimd.py:255-290): [9..12) the centroid-virial dU/deps,
 * [40..43) the per-axis centroid 2KE, [43] KE_box + P V - box work, [44] box work; the
 * PILE-L work is dpmd_thermostat_energy of segment 0. */
int32_t dpmd_npt_state_get(dpmd_system_t sys, double* state);

#ifdef __cplusplus
}
#endif
#endif

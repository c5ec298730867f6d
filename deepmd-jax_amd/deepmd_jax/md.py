"""`Simulation`: the MD driver surface of deepmd_jax/md.py:48-1090 on the GPU engine.

The host loop follows the reference's run protocol (deepmd_jax/md.py:874-964):
chunks of max(10, min(100, 100000 // (N P))) steps, each chunk one device call
(dpmd_md_run: drift test, conditional neighbour rebuild, integrator, force
evaluation per step, no host synchronisation inside), then the error-code check
with overflow recovery and chunk re-run (deepmd_jax/md.py:644-676,906-916), the
trajectory copy and the periodic report (deepmd_jax/md.py:473-595).

Routines: 'NVE' (velocity Verlet), 'NVT' (Nose-Hoover chain; with n_bead > 1
ring-polymer PILE-L), 'NVT_langevin' (BAOAB), 'NPT' (MTK Nose-Hoover barostat over
couple_axes groups with a Nose-Hoover chain thermostat, deepmd_jax/routine.py:161-456;
with n_bead > 1 MTK on the centroid with PILE-L and a Langevin barostat,
deepmd_jax/pimd.py:297-518), in the neighbour-list regime.  Positions and momenta are
FP64 on the device.
Fixed atoms (fixed_indices) and per-step centre-of-mass removal (remove_com_motion) run
on the device (deepmd_jax/md.py:221-234,786-790, deepmd_jax/routine.py:110-136).
XYZ dumps (run(dump_prefix=...)) stream from a writer thread (deepmd_jax/dumps.py).
Model deviation (model_deviation_paths) evaluates each listed model on its own device
system at report time.  DPLR models (type 'dplr') add the Wannier-centroid PPPM term on the
device, with its virial, per bead under PIMD (orthorhombic boxes; below 2 rcut on the
supercell of images with the long-range term of the cell).
"""
import time

import numpy as np

from . import _lib
from .constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION, PRESS_UNIT_CONVERSION
from .dumps import CHEMICAL_SYMBOLS, DumpWriter, prepare_dumps, target_label
from .evaluator import canonical_box, cell_volume, cell_widths, dplr_grid, dplr_setup
from .constants import HBAR
from .pimd import normal_mode_transform, pile_l_factors, centroid, spring_energy
from .utils import load_model


def print_run_profile(steps, elapsed_time, natoms, dt):
    """Per-run timing summary (deepmd_jax/md_utils.py:172-179)."""
    steps_per_microsecond_per_atom = 1e-6 * steps * natoms / (elapsed_time + 1e-6)
    nanosecond_per_day = 1e-6 * steps * dt * (86400 / elapsed_time)
    print('# Finished %d steps in %dh %dm %ds.' %
          (steps, elapsed_time // 3600, (elapsed_time % 3600) // 60, elapsed_time % 60))
    print('# Performance: %.3f ns/day, %.3f step/μs/atom' % (nanosecond_per_day, steps_per_microsecond_per_atom))


def parse_couple_axes(couple_axes):
    """Validate couple_axes (deepmd_jax/routine.py:196-223): a tuple of tuples partitioning
    a subset of {0,1,2} into coupled groups; unlisted axes are fixed.  Returns the groups
    and the per-axis group id (-1 fixed)."""
    try:
        groups = tuple(tuple(int(a) for a in g) for g in couple_axes)
    except TypeError:
        raise ValueError('couple_axes must be a tuple of tuples of ints; got %s' % (couple_axes,))
    axes = [a for g in groups for a in g]
    if any(a not in (0, 1, 2) for a in axes):
        raise ValueError('couple_axes axes must be in {0,1,2}; got %s' % (couple_axes,))
    if len(set(axes)) != len(axes):
        raise ValueError('couple_axes must be a partition (no repeats); got %s' % (couple_axes,))
    if any(len(g) == 0 for g in groups):
        raise ValueError('couple_axes groups must be non-empty; got %s' % (couple_axes,))
    if len(groups) == 0:
        raise ValueError('couple_axes must contain at least one group; use NVT instead')
    gid = np.full(3, -1, np.int32)
    for g, grp in enumerate(groups):
        for a in grp:
            gid[a] = g
    return groups, gid


class _ImageSystem:
    """A DeviceSystem seen through its primary cell: with k = (1, 1, 1) a plain pass-through;
    otherwise each replica holds the prod(k) images of the box (image-major), which the
    device integrates together (same masses, velocities and thermostat noise), and every
    getter returns the first image.  Energies, virials and kinetic energies come back per
    image from the library."""

    def __init__(self, model, type_idx, nrep, box, k):
        from .evaluator import replicate
        self.k = np.asarray(k, int)
        self.nimg = int(np.prod(self.k))
        self.N = len(type_idx)
        self.box = np.asarray(box, np.float64)
        self._shift, sbox, _ = replicate(np.zeros((self.N, 3)), self.box, type_idx, self.k)
        types = np.tile(np.asarray(type_idx), self.nimg)
        self.dev = _lib.DeviceSystem(model, types, nrep)
        self.dev.set_box(sbox)
        if self.nimg > 1:
            self.dev.set_image_grid(self.k)

    def _up(self, x):  # (P, N, 3) -> (P, nimg N, 3)
        x = np.asarray(x, np.float64)
        return x if self.nimg == 1 else np.tile(x, (1, self.nimg, 1))

    def set_positions(self, pos):
        p = np.asarray(pos, np.float64)
        self.dev.set_positions(p if self.nimg == 1 else self._up(p) + self._shift[None])

    def get_positions(self):
        return self.dev.get_positions()[:, :self.N]

    def set_velocities(self, vel):
        self.dev.set_velocities(self._up(vel))

    def get_velocities(self):
        return self.dev.get_velocities()[:, :self.N]

    def set_masses(self, mass):
        self.dev.set_masses(np.tile(np.asarray(mass, np.float64), self.nimg))

    def set_fixed(self, fixed):
        self.dev.set_fixed(None if fixed is None else np.tile(np.asarray(fixed, np.int32), self.nimg))

    def compute(self, precision):
        e, f, v, ea = self.dev.compute(precision)
        return e, f[:, :self.N], v, ea

    def md_run(self, params, nsteps, record=True, out=None):
        """out = (pos, vel) arrays of shape (nsteps, nrep, N, 3) receive the frames; with
        images held, the primary cell is copied out of the device frames."""
        if self.nimg == 1:
            return self.dev.md_run(params, nsteps, record, out)
        tp, tv, err = self.dev.md_run(params, nsteps, record)
        if record:
            tp, tv = tp[:, :, :self.N], tv[:, :, :self.N]
            if out is not None:
                out[0][...] = tp
                out[1][...] = tv
                tp, tv = out
        return tp, tv, err

    @property
    def knbr(self):
        return self.dev.knbr

    def nlist_allocate(self, rc_skin, ratio, knbr=None):
        return self.dev.nlist_allocate(rc_skin, ratio, knbr)

    def nlist_update(self):
        return self.dev.nlist_update()

    def pimd_setup(self, *a):
        return self.dev.pimd_setup(*a)

    def kinetic_energy(self):
        return self.dev.kinetic_energy()

    def thermostat_energy(self):
        return self.dev.thermostat_energy()


class Simulation:
    step: int = 0

    def __init__(self, model_path, box, type_idx, mass, routine, dt, initial_position, initial_velocity=None,
                 report_interval=200, temperature=None, pressure=None, debug=False, seed=0,
                 model_deviation_paths=[], use_neighbor_list_when_possible=True, neighbor_skin=None,
                 neighbor_buffer_ratio=1.2, remove_com_motion=False, tau_t=100., tau_p=1000., chain_length_t=3,
                 chain_length_p=3, chain_steps_t=1, chain_steps_p=1, sy_steps_t=1, sy_steps_p=1,
                 fixed_indices=None, couple_axes=((0, 1, 2),), n_bead=1, type_symbols=None, pimd_lambda=1.0,
                 pimd_print_invariant=False, precision='fp32', device=0, model=None, deterministic=False):
        self.report_interval = int(report_interval)
        self.log = []
        self.step = 0
        self._n_bead = int(n_bead)
        self._is_pimd = self._n_bead > 1
        self._pimd_lambda = float(pimd_lambda)
        if self._pimd_lambda < 0:
            raise ValueError('pimd_lambda must be non-negative.')
        if routine not in ('NVE', 'NVT', 'NVT_langevin', 'NPT'):
            raise NotImplementedError("routine currently limited to 'NVE', 'NVT', 'NVT_langevin', 'NPT'")
        if fixed_indices is not None and routine not in ('NVT', 'NVE', 'NVT_langevin'):
            # deepmd_jax/md.py:228-229
            raise NotImplementedError('Fixing atoms currently limited to NVE, NVT, and NVT_langevin routines.')
        if fixed_indices is not None and self._is_pimd:
            raise NotImplementedError('fixed atoms with path-integral MD are not provided by this engine.')
        if len(model_deviation_paths) > 0 and self._is_pimd:
            # deepmd_jax/md.py:148-149
            raise NotImplementedError('model_deviation_paths is not supported with n_bead > 1.')
        if remove_com_motion and self._is_pimd:
            raise NotImplementedError('per-step remove_com_motion with path-integral MD is not provided by this engine.')
        self._remove_com_motion = bool(remove_com_motion)
        if not (isinstance(couple_axes, tuple) and couple_axes == ((0, 1, 2),)) and routine != 'NPT':
            raise ValueError("couple_axes is only meaningful for routine='NPT'")
        for v in (sy_steps_t, sy_steps_p):
            if v not in (1, 3, 5, 7):
                raise ValueError('sy_steps must be 1, 3, 5 or 7')
        if min(chain_steps_t, chain_steps_p) < 1:
            raise ValueError('chain_steps must be >= 1')
        self._couple_axes = ((0, 1, 2),)
        if routine == 'NPT':
            if pressure is None:
                raise ValueError("Missing argument 'pressure' (in bar) for routine 'NPT'")
            if temperature is None:
                raise ValueError("Missing argument 'temperature' for routine 'NPT'")
            self._couple_axes, self._group_ids = parse_couple_axes(couple_axes)
            self._pressure = float(pressure)
        pos = np.asarray(initial_position, np.float64)
        if self._is_pimd:
            if routine not in ('NVT', 'NPT'):
                raise NotImplementedError("Path-integral MD (n_bead > 1) currently supports routine='NVT' or 'NPT'.")
            if temperature is None:
                raise ValueError('PIMD needs a temperature.')
            if pos.ndim == 2:
                pos = np.broadcast_to(pos, (self._n_bead,) + pos.shape).copy()
            elif pos.ndim == 3:
                if pos.shape[0] != self._n_bead:
                    raise ValueError('initial_position leading axis %d does not match n_bead=%d.'
                                     % (pos.shape[0], self._n_bead))
            else:
                raise ValueError('initial_position must have shape (N,3) or (P,N,3).')
            self._natoms = pos.shape[1]
        else:
            if pos.ndim != 2 or pos.shape[1] != 3:
                raise ValueError('initial_position must have shape (N,3).')
            self._natoms = pos.shape[0]
            pos = pos[None]
        self._temperature = temperature
        self._pimd_print_invariant = bool(pimd_print_invariant)
        self._dt = float(dt)
        self._routine = routine
        self._precision = precision
        if model is None:
            self._model, self._variables = load_model(model_path)
            self._device_model = _lib.DeviceModel(self._model.params, self._variables, device)
        else:
            self._model, self._variables, self._device_model = model
        params = self._model.params
        type_idx_np = np.asarray(type_idx).astype(int)
        ct = params.get('chemical_types')
        if ct is not None:
            unknown = set(type_idx_np.tolist()) - set(ct)
            if unknown:
                raise ValueError('Atomic numbers %s in type_idx are not in model.params["chemical_types"]=%s'
                                 % (sorted(unknown), ct))
            z2i = {z: i for i, z in enumerate(ct)}
            type_idx_np = np.array([z2i[z] for z in type_idx_np], dtype=int)
        if len(type_idx_np) != self._natoms:
            raise ValueError('type_idx length does not match the number of atoms.')
        if type_idx_np.min() < 0 or type_idx_np.max() >= params['ntypes']:
            raise ValueError('type index out of trained scope of model (ntypes=%d)' % params['ntypes'])
        self._type_idx = type_idx_np
        self._mass_amu = np.asarray(mass, np.float64)[self._type_idx]
        self._mass = self._mass_amu * MASS_UNIT_CONVERSION
        # symbols for XYZ dumps: given, or from the model's atomic numbers (deepmd_jax/md.py:186-194)
        if type_symbols is None and ct is not None:
            type_symbols = [CHEMICAL_SYMBOLS[int(z)] for z in ct]
        if type_symbols is not None and len(type_symbols) < int(self._type_idx.max()) + 1:
            raise ValueError('type_symbols has length %d but type_idx references types up to %d.'
                             % (len(type_symbols), int(self._type_idx.max())))
        self._type_symbols = list(type_symbols) if type_symbols is not None else None
        self._debug = debug
        self._step_chunk_size = max(10, min(100, 100000 // (self._natoms * self._n_bead)))
        self._neighbor_skin = (0.5 if routine == 'NPT' else 0.3) if neighbor_skin is None else float(neighbor_skin)
        self._neighbor_buffer_ratio = float(neighbor_buffer_ratio)
        # (3,) lengths of an orthorhombic box, else the (3,3) cell (deepmd_jax/md.py:208-215)
        self._box = canonical_box(box)
        self._general_cell = self._box.shape == (3, 3)
        if self._general_cell and routine == 'NPT':
            raise NotImplementedError('NPT supports orthorhombic boxes (deepmd_jax/routine.py:161-170).')
        # distances between opposite faces: the box lengths of an orthorhombic box
        widths = cell_widths(self._box)
        self._rcut = float(params['rcut'])
        # The reference switches to its dense lattice-image path below 2*(rcut+skin)
        # (deepmd_jax/utils.py:66-78).  Here the neighbour list is exact for any skin with
        # 2*(rcut+skin) <= L, so the skin is shrunk to fit; a box below 2*rcut is held as a
        # k1 x k2 x k3 periodic supercell whose images move together (the lattice-image sum).
        k = np.ones(3, int)
        # the reference's own neighbour-list condition (deepmd_jax/md.py:450-456): its debug
        # report shows the neighbour update rate only then
        self._reference_nlist = bool(use_neighbor_list_when_possible) and not self._general_cell and widths.min() > 2 * (
            self._rcut + self._neighbor_skin) * (1.02 if routine == 'NPT' else 1.0)
        if routine == 'NPT' and (widths.min() <= 2 * (self._rcut + self._neighbor_skin) * 1.02
                                 or not use_neighbor_list_when_possible):
            # the reference's NPT condition for its neighbour list (deepmd_jax/md.py:450-456)
            raise NotImplementedError('NPT needs box lengths above 2*(rcut+skin)*1.02 = %.3f A (neighbour-list '
                                      'regime); smaller NPT boxes are not provided by this engine.'
                                      % (2 * (self._rcut + self._neighbor_skin) * 1.02))
        if widths.min() < 2 * self._rcut:
            if self._is_pimd:
                raise NotImplementedError('path-integral MD needs box widths >= 2*rcut (%.3f)' % (2 * self._rcut))
            k = np.maximum(1, np.ceil(2 * (self._rcut + self._neighbor_skin) / widths * (1 + 1e-9))).astype(int)
        elif widths.min() <= 2 * (self._rcut + self._neighbor_skin) or not use_neighbor_list_when_possible:
            self._neighbor_skin = max(0.0, 0.5 * widths.min() * (1 - 1e-9) - self._rcut)
        self._sys = _ImageSystem(self._device_model, self._type_idx, self._n_bead, self._box, k)
        # engine option: bit-reproducible forces (pair gradients gathered in a fixed order)
        self._sys.dev.set_deterministic(bool(deterministic))
        self._sys.set_masses(self._mass)
        # DPLR: the centroid model and PPPM term join every force evaluation (deepmd_jax/md.py:404-415)
        self._dplr = dplr_setup(params, device)
        if self._dplr is not None:
            if self._general_cell:
                raise NotImplementedError('dplr model currently only supports orthorhombic box')
        # mobile-atom mask (deepmd_jax/md.py:220-226); the device zeroes the forces on fixed
        # atoms, keeps their positions and masks their momenta after each NVE / NVT step
        self._mobile = None
        if fixed_indices is not None:
            mobile = np.ones(self._natoms, bool)
            mobile[np.asarray(fixed_indices)] = False
            if not mobile.any():
                raise ValueError('fixed_indices fixes every atom.')
            self._mobile = mobile
            self._sys.set_fixed(~mobile)
        self._sys.set_positions(pos)
        self._allocate_nlist()
        if self._dplr is not None:
            wdm, qa, qw, beta, res = self._dplr
            self._sys.dev.set_dplr(wdm, qa, qw, beta, dplr_grid(self._box, beta, res))
        # models of the deviation column (deepmd_jax/md.py:299-302,559-567): each on its own
        # device system over the same atoms, evaluated at report time
        self._deviation = []
        for path in model_deviation_paths:
            dmodel, dvar = load_model(path)
            ds = _ImageSystem(_lib.DeviceModel(dmodel.params, dvar, device), self._type_idx, 1, self._box, k)
            ds.set_masses(self._mass)
            if self._mobile is not None:
                ds.set_fixed(~self._mobile)
            ds.set_positions(pos)
            ds.nlist_allocate(float(dmodel.params['rcut']) + self._neighbor_skin, self._neighbor_buffer_ratio)
            # a DPLR deviation model carries its long-range term, as its energy function does in the
            # reference (deepmd_jax/md.py:299-302,385-415)
            dd = dplr_setup(dmodel.params, device)
            if dd is not None:
                if self._general_cell:
                    raise NotImplementedError('dplr model currently only supports orthorhombic box')
                wdm, qa, qw, beta, res = dd
                ds.dev.set_dplr(wdm, qa, qw, beta, dplr_grid(self._box, beta, res))
            self._deviation.append((ds, float(dmodel.params['rcut'])))
        self._kT = (temperature if temperature is not None else 0.0) * TEMP_UNIT_CONVERSION
        self._params = _lib.MDParams()
        self._params.precision = _lib.PREC[precision]
        self._params.dt = self._dt
        self._params.kT = self._kT
        self._params.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._params.tau = float(tau_t)
        self._params.chain_length = int(chain_length_t)
        self._params.gamma = 1.0 / float(tau_t)
        self._params.pimd_lambda = self._pimd_lambda
        self._params.remove_com = int(self._remove_com_motion)
        # Nose-Hoover chain substeps (jax_md chain_steps / sy_steps, deepmd_jax/md.py:80-87)
        self._params.chain_steps, self._params.sy_steps = int(chain_steps_t), int(sy_steps_t)
        self._params.chain_steps_p, self._params.sy_steps_p = int(chain_steps_p), int(sy_steps_p)
        if routine == 'NPT':
            # deepmd_jax/md.py:357-366: pressure in eV/A^3, barostat chain of tau_p
            self._params.pressure = self._pressure * PRESS_UNIT_CONVERSION
            self._params.tau_p = float(tau_p)
            self._params.chain_length_p = int(chain_length_p)
            for a in range(3):
                self._params.couple[a] = int(self._group_ids[a])
        if self._is_pimd:
            # PILE-L of the bead momenta (deepmd_jax/md.py:341-351,368-381); path-integral NPT
            # adds a Langevin barostat of friction 1 / tau_p (deepmd_jax/md.py:374)
            self._params.routine = _lib.ROUTINE['PIMD_NPT' if routine == 'NPT' else 'PIMD']
            if routine == 'NPT':
                self._params.gamma = 1.0 / float(tau_p)
            freqs, C = normal_mode_transform(self._n_bead, self._kT)
            _, c1, c2 = pile_l_factors(freqs, self._dt, self._kT, self._n_bead, float(tau_t), self._pimd_lambda)
            self._sys.pimd_setup(C, freqs, c1, c2)
            self._nm = (freqs, C)
        else:
            self._params.routine = _lib.ROUTINE[routine]
        # initial velocities: MB at kT (COM momentum removed), or the caller's
        rng = np.random.default_rng(seed)
        if initial_velocity is not None:
            self.setVelocity(initial_velocity)
        else:
            kTP = self._kT * self._n_bead
            p = rng.standard_normal((self._n_bead, self._natoms, 3)) * np.sqrt(self._mass * kTP)[None, :, None]
            if not self._is_pimd:
                p -= p.mean(1, keepdims=True)
            if self._mobile is not None:
                p *= self._mobile[None, :, None]
            self._sys.set_velocities(p / self._mass[None, :, None])
        self._is_initial_state = True
        self._error_code = 0
        labels = {('NVE', False): 'NVE run (velocity-Verlet integrator)',
                  ('NVT', False): 'NVT run (Nose-Hoover thermostat)',
                  ('NVT_langevin', False): 'NVT Langevin run (BAOAB Langevin thermostat)',
                  ('NVT', True): 'NVT PIMD run (PILE-L Langevin thermostat)',
                  ('NPT', False): 'NPT run (Nose-Hoover thermostat/barostat)',
                  ('NPT', True): 'NPT PIMD run (PILE-L thermostat, Langevin barostat)'}
        info = []
        if temperature is not None and routine != 'NVE':
            info.append('temperature=%g K' % temperature)
        if routine == 'NPT':
            info.append('pressure=%g bar' % self._pressure)
        if self._is_pimd:
            info += ['n_beads=%d' % self._n_bead, 'pimd_lambda=%g' % self._pimd_lambda]
        print('# Initialized %s with %d atoms%s' % (labels[(routine, self._is_pimd)], self._natoms,
                                                   (' (%s)' % ', '.join(info)) if info else ''))

    # ---------------------------------------------------------------- internals
    def _allocate_nlist(self, knbr=None):
        rc_skin = self._rcut + self._neighbor_skin
        self._sys.nlist_allocate(rc_skin, self._neighbor_buffer_ratio, knbr)
        print('# Neighborlist allocated with size %s, rcut_plus_skin = %g, buffer_ratio = %g'
              % (np.array(self._sys.knbr) - 1, rc_skin, self._neighbor_buffer_ratio))

    def _state_pos(self):
        return self._sys.get_positions()

    def _report_values(self):
        e, f, v, _ = self._sys.compute(self._precision)
        ke = self._sys.kinetic_energy()
        P = self._n_bead
        pe = float(e.mean()) if self._is_pimd else float(e[0])
        ke_tot = float(ke.sum())
        N = self._natoms
        # temperature over the mobile degrees of freedom (deepmd_jax/md.py:478)
        dof_scale = N / int(self._mobile.sum()) if self._mobile is not None else 1
        temp = 2 * ke_tot / (3 * N * P) / TEMP_UNIT_CONVERSION / P * dof_scale
        rep = {'Temperature': temp, 'KE': ke_tot / P ** 2, 'PE': pe}
        if self._routine == 'NVE':
            rep['Invariant'] = ke_tot + pe
        elif self._routine == 'NVT_langevin':
            rep['Invariant'] = ke_tot + pe - float(self._sys.thermostat_energy()[0])
        elif self._routine == 'NVT' and not self._is_pimd:
            rep['Invariant'] = ke_tot + pe + float(self._sys.thermostat_energy()[0])
        elif self._routine == 'NVT' and self._pimd_print_invariant:
            rep['Invariant'] = self._pimd_invariant(e, ke_tot, self._box)
        elif self._routine == 'NPT':
            # deepmd_jax/md.py:523-538: pressure, moving box lengths, invariant
            box = self.getBox()
            # iso_pressure of the (bead-averaged) energy (deepmd_jax/md_utils.py:139-159)
            rep['Pressure'] = (N * self._kT + float(np.trace(v.mean(0))) / 3.0) / cell_volume(box) / PRESS_UNIT_CONVERSION
            for a in sorted(a for g in self._couple_axes for a in g):
                rep['box_' + 'xyz'[a]] = float(box[a])
            if not self._is_pimd:
                # before the first step the barostat is at rest: its part is P V alone
                te = float(self._sys.thermostat_energy()[0])
                try:
                    self._sys.dev.npt_state()
                except _lib.DPMDError:
                    te += self._params.pressure * float(np.prod(box))
                rep['Invariant'] = ke_tot + pe + te
            elif self._pimd_print_invariant:
                # deepmd_jax/md.py:540-557: the barostat part KE_box + P V - box work is kept
                # on the device
                rep['Invariant'] = self._pimd_invariant(e, ke_tot, box) + self._barostat_terms(box)
        if self._deviation:
            rep['Model_Devi'] = self._model_deviation()
        self._last_force = f
        self._last_virial = v
        self._last_energy = e
        return rep

    def _model_deviation(self):
        """max over atoms and components of the standard deviation of the deviation models'
        forces at the current positions (deepmd_jax/md.py:559-567)."""
        pos = self._sys.get_positions()
        box = self.getBox()
        forces = []
        for ds, rc in self._deviation:
            if self._routine == 'NPT':
                ds.dev.set_box(box)
                ds.set_positions(pos)
                ds.nlist_allocate(rc + self._neighbor_skin, self._neighbor_buffer_ratio)
            else:
                ds.set_positions(pos)
                if ds.nlist_update():
                    ds.nlist_allocate(rc + self._neighbor_skin, self._neighbor_buffer_ratio)
            forces.append(ds.compute(self._precision)[1][0])
        return float(np.std(np.stack(forces), axis=0).max())

    def _barostat_terms(self, box):
        """KE_box + P V - box work of path-integral NPT, kept on the device; P V before the
        first step (box at rest)."""
        try:
            return float(self._sys.dev.npt_state()[43])
        except _lib.DPMDError:
            return self._params.pressure * float(np.prod(box))

    def _pimd_invariant(self, e_beads, ke_tot, box):
        """(H_rp - PILE-L work) / P with H_rp = KE + sum of bead energies + spring energy
        (deepmd_jax/md.py:512-521)."""
        P = self._n_bead
        omega_p = P * self._kT / HBAR
        h_rp = ke_tot + float(np.sum(e_beads)) + spring_energy(self._sys.get_positions(), self._mass, box, omega_p)
        return (h_rp - float(self._sys.thermostat_energy()[0])) / P

    def _print_report(self):
        rep = self._report_values()
        keys = list(rep.keys())
        rate = self._debug and self._reference_nlist  # NbrUpdateRate column (deepmd_jax/md.py:583-592)
        if self._is_initial_state:
            print(' '.join(['%-8s' % 'Step'] + ['%-12s' % k for k in keys] + ['%-6s' % 'Time']
                           + (['%-6s' % 'NbrUpdateRate'] if rate else [])))
        now = time.time()
        row = [self.step] + [rep[k] for k in keys] + [now - self._tic_between_report]
        if rate:
            row.append(self._sys.dev.nlist_update_rate())
        self.log.append(row)
        print('%-8d ' % row[0] + ' '.join('%-12.3f' % x for x in row[1:len(keys) + 1])
              + ' %-6.3f' % row[len(keys) + 1] + (' %-6.3f' % row[-1] if rate else ''))
        self._tic_between_report = now

    def _initialize_run(self, steps, writer=None, label=None):
        print('# Running %d steps%s...' % (steps, (' and dumping to %s' % label) if writer else ''))
        traj_length = steps + int(self._is_initial_state)
        self._offset = self.step - int(self._is_initial_state)
        shape = (self._n_bead, self._natoms, 3) if self._is_pimd else (self._natoms, 3)
        if writer is None:
            self._position_trajectory = np.zeros((traj_length,) + shape)
            self._velocity_trajectory = np.zeros((traj_length,) + shape)
            self._box_trajectory = np.zeros((traj_length,) + self._box.shape)
            if self._is_pimd:
                self._centroid_trajectory = np.zeros((traj_length, self._natoms, 3))
            if self._is_initial_state:
                self._position_trajectory[0] = self.getPosition()
                self._velocity_trajectory[0] = self.getVelocity()
                self._box_trajectory[0] = self.getBox()
                if self._is_pimd:
                    self._centroid_trajectory[0] = self.getCentroidPosition()
        else:
            # frames stream to the dump files; the device writes each chunk into one reused buffer
            n = self._step_chunk_size
            self._chunk_pos = np.empty((n, self._n_bead, self._natoms, 3))
            self._chunk_vel = np.empty((n, self._n_bead, self._natoms, 3))
            if self._is_initial_state and self.step % self._dump_interval == 0:
                self._put_frame(writer, self._sys.get_positions(), self._sys.get_velocities(), self.getBox(),
                                self.step)
        self._tic_of_this_run = time.time()
        self._tic_between_report = time.time()
        self._error_code = 0
        self._print_report()
        self._is_initial_state = False

    def _resolve_overflow(self):
        """deepmd_jax/md.py:644-676.  Bit 1 (NPT: the box shrank past what the list's skin
        covers): widen the skin by 0.4 A and rebuild the list on the current box.  Bit 2:
        reallocate, growing the buffer ratio when bit 4 says an overflow was already resolved
        in this run.  Bit 4 then stays set until the next run() (md.py:870)."""
        flag4 = False
        if self._error_code & 1:
            self._neighbor_skin += 0.4
            box = np.asarray(self.getBox(), np.float64).reshape(-1)
            if box.size == 3 and box.min() > 2 * (self._rcut + self._neighbor_skin) * 1.02:
                self._allocate_nlist()
            else:
                raise NotImplementedError('NPT box shrank to %s: below 2*(rcut+skin)*1.02 = %.3f A with the widened '
                                          'skin, outside the neighbour-list regime this engine runs NPT in'
                                          % (box, 2 * (self._rcut + self._neighbor_skin) * 1.02))
        elif self._error_code & 2:
            if self._error_code & 4:
                self._neighbor_buffer_ratio += 0.05
            self._allocate_nlist()
            flag4 = True
        print('# Successfully resolved overflow code', self._error_code)
        self._error_code = 4 if flag4 else 0

    def _put_frame(self, writer, pos, vel, box, step):
        """Queue one frame ((P,N,3) positions and velocities) for the dump files, as float32
        copies (the reference's default frame dtype)."""
        kinds = {d[0] for d in writer.dumps}
        p32 = np.array(pos if self._is_pimd else pos[0], np.float32) if 'position' in kinds else None
        v32 = np.array(vel if self._is_pimd else vel[0], np.float32) if 'velocity' in kinds else None
        c32 = np.array(centroid(pos, box), np.float32) if 'centroid' in kinds else None
        writer.put(p32, v32, c32, np.array(box, np.float32), step)

    def run(self, steps, dump_interval=None, dump_prefix=None, dump_content='positions', dump_mode='overwrite'):
        """Run `steps` steps; returns the trajectory dict, or None when frames stream to XYZ files
        (deepmd_jax/md.py:874-964)."""
        dumps, symbols, self._dump_interval = prepare_dumps(self._type_idx, self._type_symbols, self._n_bead,
                                                            self.report_interval, dump_interval, dump_prefix,
                                                            dump_content, dump_mode)
        writer = DumpWriter(dumps, symbols) if dumps else None
        try:
            return self._run(steps, writer, target_label(dumps, dump_prefix) if dumps else None)
        finally:
            if writer is not None:
                writer.close()
                self._chunk_pos = self._chunk_vel = None

    def _run(self, steps, writer, label):
        self._initialize_run(steps, writer, label)
        remaining = steps
        while remaining > 0:
            nxt = min(self.report_interval - self.step % self.report_interval, self._step_chunk_size, remaining)
            lo = self.step - self._offset
            # the device writes the chunk's frames straight into the trajectory arrays (or the
            # reused chunk buffer when streaming to dump files)
            shp = (nxt, self._n_bead, self._natoms, 3)
            if writer is None:
                out = (self._position_trajectory[lo:lo + nxt].reshape(shp),
                       self._velocity_trajectory[lo:lo + nxt].reshape(shp))
            else:
                out = (self._chunk_pos[:nxt], self._chunk_vel[:nxt])
            tp, tv, err = self._sys.md_run(self._params, nxt, record=True, out=out)
            if err & 16:
                print('# Warning: Nan or Inf encountered in simulation. Terminating.')
                self._error_code |= 16
                remaining = 0
            if err & 3:
                self._error_code |= err & 3
                self._resolve_overflow()
                continue
            boxes = self._sys.dev.box_trajectory(nxt)[:, 0] if self._routine == 'NPT' else None
            if writer is not None:
                for k in range(nxt):
                    if (self.step + k + 1) % self._dump_interval == 0:
                        self._put_frame(writer, tp[k], tv[k], self._box if boxes is None else boxes[k],
                                        self.step + k + 1)
            else:
                if self._is_pimd:
                    for k in range(nxt):
                        self._centroid_trajectory[lo + k] = centroid(tp[k], self._box if boxes is None else boxes[k])
                self._box_trajectory[lo:lo + nxt] = self._box if boxes is None else boxes
            self.step += nxt
            remaining -= nxt
            if self.step % self.report_interval == 0 or remaining == 0:
                self._print_report()
        print_run_profile(steps, time.time() - self._tic_of_this_run, self._natoms * self._n_bead, self._dt)
        if writer is not None:
            return None
        traj = {'position': self._position_trajectory, 'velocity': self._velocity_trajectory,
                'box': self._box_trajectory}
        if self._is_pimd:
            traj['centroid'] = self._centroid_trajectory
        return traj

    # ------------------------------------------------------------ getters/setters
    def getPosition(self):
        p = self._state_pos()
        return p if self._is_pimd else p[0]

    def getCentroidPosition(self):
        p = self._state_pos()
        return centroid(p, self.getBox()) if self._is_pimd else p[0]

    def setPosition(self, position, box=None):
        """Set positions; under NPT an isotropically scaled box may come with them
        (deepmd_jax/md.py:985-1015: it holds for the forces until the barostat moves the box
        again, which starts from the barostat's own state, as there)."""
        if box is not None:
            if self._routine != 'NPT':
                raise ValueError('Box can only be changed in NPT simulations.')
            b = np.asarray(box, np.float64)
            b = np.full(3, float(b.reshape(()))) if b.size == 1 else b
            cur = self.getBox()
            if b.shape != cur.shape:
                raise ValueError('Box must have the same shape as the initial box.')
            if not np.allclose(b, cur * (b[0] / cur[0]), rtol=1e-4, atol=1e-6):
                raise ValueError('Only isotropic box fluctuations are allowed in the current implementation.')
        position = np.asarray(position, np.float64)
        if self._is_pimd and position.ndim == 2:
            position = np.broadcast_to(position, (self._n_bead,) + position.shape)
        expect = (self._n_bead, self._natoms, 3) if self._is_pimd else (self._natoms, 3)
        if position.shape != expect:
            raise ValueError('Position must have the same shape as the initial position, or you have to create a '
                             'new Simulation instance.')
        if box is not None:
            self._sys.dev.set_box(b)
        self._sys.set_positions(position.reshape(self._n_bead, self._natoms, 3))
        if self._sys.nlist_update():
            self._allocate_nlist()

    def getVelocity(self):
        v = self._sys.get_velocities()
        return v if self._is_pimd else v[0]

    def setVelocity(self, velocity):
        v = np.asarray(velocity, np.float64)
        if self._is_pimd and v.ndim == 2:
            v = np.broadcast_to(v, (self._n_bead,) + v.shape)
        if self._mobile is not None:
            v = v * self._mobile[:, None]
        self._sys.set_velocities(v.reshape(self._n_bead, self._natoms, 3))

    def remove_com_motion(self):
        """Remove the centre-of-mass velocity (deepmd_jax/md.py:743-757): over all beads and atoms
        under PIMD, else over the mobile atoms, subtracted from every atom."""
        v = self._sys.get_velocities()
        w = np.broadcast_to(self._mass[None, :, None], v.shape[:2] + (1,))
        if self._mobile is not None:
            w = w * self._mobile[None, :, None]
        self._sys.set_velocities(v - (v * w).sum((0, 1)) / w.sum())

    def setRandomVelocity(self, temperature, remove_com_motion=True, seed=None):
        rng = np.random.default_rng(seed)
        kT = temperature * TEMP_UNIT_CONVERSION
        v = rng.standard_normal((self._n_bead, self._natoms, 3)) * np.sqrt(kT * self._n_bead / self._mass)[None, :, None]
        self.setVelocity(v if self._is_pimd else v[0])
        if remove_com_motion:
            self.remove_com_motion()

    def getBox(self):
        if self._routine == 'NPT':
            return self._sys.dev.get_boxes()[0]
        return self._box.copy()

    def getEnergy(self):
        e, _, _, _ = self._sys.compute(self._precision)
        return float(e.mean())

    def getForce(self):
        _, f, _, _ = self._sys.compute(self._precision)
        return f if self._is_pimd else f[0]

    def getVirial(self):
        _, _, v, _ = self._sys.compute(self._precision)
        return v if self._is_pimd else v[0]

    def getPressure(self):
        """Isotropic pressure (bar): (N kT - dU/deps / 3) / V, deepmd_jax/md_utils.py:139-159."""
        if self._temperature is None:
            raise ValueError('getPressure needs a temperature')
        _, _, v, _ = self._sys.compute(self._precision)
        dUdeps = -float(np.trace(v.mean(0)))
        V = cell_volume(self.getBox())
        return (self._natoms * self._kT - dUdeps / 3.0) / V / PRESS_UNIT_CONVERSION

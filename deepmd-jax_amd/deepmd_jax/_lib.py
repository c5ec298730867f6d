"""ctypes binding of libdpmd_hip.so (C ABI declared in include/dpmd.h).

The library is the only compute path: importing this module fails loudly when the
shared object is missing, and every call raises on a non-zero status.
"""
import ctypes
import os

import numpy as np

MAXT = 4
# arithmetic of the embedding / fitting GEMMs (DPMD_PREC_*): FP32, BF16, FP16, and the mixed
# forms with BF16 embedding GEMMs and split-BF16 ('bf16x3') or FP32 ('bf16-fit32') fitting GEMMs
PREC = {'fp32': 0, 'bf16': 1, 'fp16': 2, 'bf16x3': 3, 'bf16-fit32': 4}
ROUTINE = {'NVE': 0, 'NVT': 1, 'NVT_langevin': 2, 'PIMD': 3, 'NPT': 4, 'PIMD_NPT': 5}

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libdpmd_hip.so')

c_i32, c_f64, c_vp = ctypes.c_int32, ctypes.c_double, ctypes.c_void_p


class ModelDesc(ctypes.Structure):
    _fields_ = [('ntypes', c_i32), ('embed_widths', c_i32 * 3), ('fit_widths', c_i32 * 3), ('axis', c_i32),
                ('rcut', c_f64), ('sr_mean', c_f64 * MAXT), ('sr_std', c_f64 * MAXT), ('nnbrs', c_f64),
                ('ebias', c_f64 * MAXT), ('out_norm', c_f64), ('tbias', c_vp),
                ('emb_kernel', (c_vp * 3) * (MAXT * MAXT)), ('emb_bias', (c_vp * 3) * (MAXT * MAXT)),
                ('fit_kernel', (c_vp * 4) * MAXT), ('fit_bias', (c_vp * 4) * MAXT), ('fit_dt', (c_vp * 3) * MAXT),
                ('tab_ngrids', c_i32), ('tab_srmin', ctypes.c_float * (MAXT * MAXT)),
                ('tab_srmax', ctypes.c_float * (MAXT * MAXT)), ('tab_coeff', c_vp * (MAXT * MAXT)),
                ('atomic', c_i32), ('nsel_mask', c_i32)]


class MDParams(ctypes.Structure):
    _fields_ = [('routine', c_i32), ('precision', c_i32), ('dt', c_f64), ('kT', c_f64), ('gamma', c_f64),
                ('tau', c_f64), ('chain_length', c_i32), ('record_every', c_i32), ('seed', ctypes.c_uint64),
                ('pimd_lambda', c_f64), ('pressure', c_f64), ('tau_p', c_f64), ('chain_length_p', c_i32),
                ('couple', c_i32 * 3), ('remove_com', c_i32), ('chain_steps', c_i32), ('sy_steps', c_i32),
                ('chain_steps_p', c_i32), ('sy_steps_p', c_i32)]


class DPMDError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libdpmd_hip.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DPMDError('%s not found: build it with `python deepmd-jax_amd/build.py`' % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sig = {
        'dpmd_last_error': (ctypes.c_char_p, []),
        'dpmd_version': (c_i32, []),
        'dpmd_model_create': (c_i32, [P(ModelDesc), c_i32, P(c_vp)]),
        'dpmd_model_destroy': (c_i32, [c_vp]),
        'dpmd_system_create': (c_i32, [c_vp, c_i32, c_i32, c_vp, P(c_vp)]),
        'dpmd_system_create_ragged': (c_i32, [c_vp, c_i32, c_vp, c_vp, P(c_vp)]),
        'dpmd_system_reshape': (c_i32, [c_vp, c_i32, c_vp, c_vp]),
        'dpmd_system_destroy': (c_i32, [c_vp]),
        'dpmd_system_set_box': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_boxes': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_cells': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_images': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_image_grid': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_rng': (c_i32, [c_vp, c_vp, c_vp]),
        'dpmd_system_get_rng': (c_i32, [c_vp, c_vp, c_vp]),
        'dpmd_system_set_positions': (c_i32, [c_vp, c_vp]),
        'dpmd_system_get_positions': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_velocities': (c_i32, [c_vp, c_vp]),
        'dpmd_system_get_velocities': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_masses': (c_i32, [c_vp, c_vp]),
        'dpmd_nlist_allocate': (c_i32, [c_vp, c_f64, c_f64, c_vp, c_vp]),
        'dpmd_nlist_update': (c_i32, [c_vp, c_vp]),
        'dpmd_nlist_export': (c_i32, [c_vp, c_i32, c_vp]),
        'dpmd_nlist_capacity': (c_i32, [c_vp, c_vp]),
        'dpmd_compute': (c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp]),
        'dpmd_compute_noalloc': (c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp]),
        'dpmd_system_set_deterministic': (c_i32, [c_vp, c_i32]),
        'dpmd_nlist_overflow': (c_i32, [c_vp, c_vp]),
        'dpmd_md_run': (c_i32, [c_vp, P(MDParams), c_i32, c_vp, c_vp, c_vp]),
        'dpmd_pimd_setup': (c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
        'dpmd_kinetic_energy': (c_i32, [c_vp, c_vp]),
        'dpmd_thermostat_energy': (c_i32, [c_vp, c_vp]),
        'dpmd_thermostat_state_get': (c_i32, [c_vp, c_vp]),
        'dpmd_thermostat_state_set': (c_i32, [c_vp, c_vp]),
        'dpmd_timing_enable': (c_i32, [c_vp, c_i32]),
        'dpmd_timing_get': (c_i32, [c_vp, c_i32, c_vp, c_vp]),
        'dpmd_pair_count': (c_i32, [c_vp, c_vp]),
        'dpmd_synchronize': (c_i32, [c_vp]),
        'dpmd_system_get_boxes': (c_i32, [c_vp, c_vp]),
        'dpmd_md_box_trajectory': (c_i32, [c_vp, c_vp]),
        'dpmd_npt_state_get': (c_i32, [c_vp, c_vp]),
        'dpmd_nlist_rebuilds': (c_i32, [c_vp, c_vp]),
        'dpmd_system_set_fixed': (c_i32, [c_vp, c_vp]),
        'dpmd_pppm_compute': (c_i32, [c_i32, c_vp, c_f64, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
        'dpmd_system_set_dplr': (c_i32, [c_vp, c_vp, c_vp, c_vp, c_f64, c_vp]),
        'dpmd_dplr_get': (c_i32, [c_vp, c_vp, c_vp]),
        'dpmd_atomic_predict': (c_i32, [c_vp, c_vp]),
        'dpmd_nlist_update_rate': (c_i32, [c_vp, c_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().dpmd_last_error().decode(errors='replace')
        if rc == -1:
            raise ValueError(msg)
        if rc == -4:
            raise NotImplementedError(msg)
        raise DPMDError('libdpmd_hip error %d: %s' % (rc, msg))


def ptr(a):
    return a.ctypes.data_as(c_vp) if a is not None else None


def exported_symbols():
    """Names of the C-ABI entry points bound here (all declared in include/dpmd.h)."""
    lib()
    return ['dpmd_last_error', 'dpmd_version', 'dpmd_model_create', 'dpmd_model_destroy', 'dpmd_system_create',
            'dpmd_system_create_ragged', 'dpmd_system_reshape', 'dpmd_system_destroy', 'dpmd_system_set_box',
            'dpmd_system_set_boxes', 'dpmd_system_set_cells', 'dpmd_system_set_images',
            'dpmd_system_set_image_grid', 'dpmd_system_set_rng', 'dpmd_system_get_rng',
            'dpmd_system_set_positions', 'dpmd_system_get_positions',
            'dpmd_system_set_velocities', 'dpmd_system_get_velocities', 'dpmd_system_set_masses',
            'dpmd_nlist_allocate', 'dpmd_nlist_update', 'dpmd_nlist_export', 'dpmd_nlist_capacity', 'dpmd_compute',
            'dpmd_md_run', 'dpmd_pimd_setup', 'dpmd_kinetic_energy', 'dpmd_thermostat_energy',
            'dpmd_thermostat_state_get', 'dpmd_thermostat_state_set', 'dpmd_timing_enable',
            'dpmd_timing_get', 'dpmd_pair_count', 'dpmd_synchronize', 'dpmd_system_get_boxes',
            'dpmd_md_box_trajectory', 'dpmd_npt_state_get', 'dpmd_nlist_rebuilds', 'dpmd_system_set_fixed',
            'dpmd_pppm_compute', 'dpmd_system_set_dplr', 'dpmd_dplr_get', 'dpmd_atomic_predict',
            'dpmd_nlist_update_rate', 'dpmd_compute_noalloc', 'dpmd_system_set_deterministic',
            'dpmd_nlist_overflow']


def pppm_compute(box3, beta, grid, xyz, q, device=0):
    """PPPM energy (eV), forces (n, 3) and virial (3, 3) of point charges q (e) at xyz (Angstrom)
    in an orthorhombic box on the GPU (get_p3mlr_fn, deepmd_jax/utils.py:177-222)."""
    b = np.ascontiguousarray(np.asarray(box3, np.float64).reshape(3))
    g = np.ascontiguousarray(np.asarray(grid, np.int32).reshape(3))
    x = np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1, 3))
    qq = np.ascontiguousarray(np.asarray(q, np.float64).reshape(-1))
    if len(qq) != len(x):
        raise ValueError('xyz and q disagree in length')
    e = np.zeros(1)
    f = np.zeros_like(x)
    w = np.zeros((3, 3))
    check(lib().dpmd_pppm_compute(int(device), ptr(b), float(beta), ptr(g), len(x), ptr(x), ptr(qq), ptr(e), ptr(f),
                                  ptr(w)))
    return float(e[0]), f, w


class DeviceModel:
    """Model weights resident on one GPU (replaces load_model's device_put)."""

    def __init__(self, params, variables, device=0):
        P = variables['params']
        d = ModelDesc()
        nt = int(params['ntypes'])
        atomic = bool(params.get('atomic', False))
        if params.get('use_mp', False):
            raise NotImplementedError('only the se_a model (no message passing) is supported')
        kind = params.get('type', 'atomic')
        if atomic and kind not in ('atomic', 'atomic_t2'):
            raise NotImplementedError("atomic models of type %r are not supported" % kind)
        if atomic and params.get('is_compressed', False):
            raise NotImplementedError('compressed atomic models are not supported')
        if not params.get('use_2nd', True):
            raise NotImplementedError('use_2nd=False is not supported')
        if list(params.get('valid_types', range(nt))) != list(range(nt)):
            raise NotImplementedError('models with valid_types != range(ntypes) are not supported')
        ew = list(params['embed_widths'])
        fw = list(params['fit_widths'])
        if len(ew) != 3 or len(fw) != 3:
            raise NotImplementedError('kernels implement 3 embedding layers and 3 fitting layers')
        d.ntypes = nt
        for i in range(3):
            d.embed_widths[i] = int(ew[i])
            d.fit_widths[i] = int(fw[i])
        d.axis = int(params['axis'])
        d.rcut = float(params['rcut'])
        for t in range(nt):
            d.sr_mean[t] = float(np.asarray(params['sr_mean'])[t])
            d.sr_std[t] = float(np.asarray(params['sr_std'])[t])
            eb = params.get('Ebias')
            d.ebias[t] = 0.0 if eb is None else float(np.asarray(eb).reshape(-1)[t])
        d.nnbrs = float(params['Nnbrs'])
        d.out_norm = float(params.get('out_norm', 1.0))
        keep = []

        def arr(x):
            a = np.ascontiguousarray(np.asarray(x, np.float32))
            keep.append(a)
            return a.ctypes.data

        d.tbias = arr(P['Tbias'])
        compressed = bool(params.get('is_compressed', False))
        if atomic:
            # nets of centre type i live at (position of i in nsel) * ntypes + j; centre types
            # outside nsel carry no output and get the first selected type's nets
            nsel = [int(t) for t in params['nsel']]
            pos = {t: k for k, t in enumerate(nsel)}
            d.atomic = 1 if kind == 'atomic' else 2
            d.nsel_mask = sum(1 << t for t in nsel)
            P = dict(P)
            src = {}
            for i in range(nt):
                ci = pos.get(i, 0)
                for j in range(nt):
                    src['embedding_net_%d' % (i * nt + j)] = P['embedding_net_%d' % (ci * nt + j)]
                src['fitting_net_%d' % i] = P['fitting_net_%d' % ci]
            P.update(src)
        for k in range(nt * nt):
            net = P.get('embedding_net_%d' % k)
            if net is None:
                if not compressed:
                    raise ValueError('embedding_net_%d missing from the model variables' % k)
                continue
            for l in range(3):
                d.emb_kernel[k][l] = arr(net['Dense_%d' % l]['kernel'])
                d.emb_bias[k][l] = arr(net['Dense_%d' % l]['bias'])
        if compressed:
            # tabulated nets of compress_model (deepmd_jax/utils.py:250-297)
            cvs = variables.get('compress_var')
            if cvs is None:
                raise ValueError("model.params['is_compressed'] is set but variables has no 'compress_var'")
            ngrids = None
            for k in range(nt * nt):
                cv = cvs['embedding_net_%d' % k]
                c = np.asarray(cv['poly_coeff'], np.float32)
                if c.ndim != 3 or c.shape[1] != 6 or c.shape[2] != int(ew[-1]):
                    raise ValueError('poly_coeff of embedding_net_%d has shape %s, expected (Ngrids, 6, %d)'
                                     % (k, c.shape, int(ew[-1])))
                if ngrids is None:
                    ngrids = c.shape[0]
                elif c.shape[0] != ngrids:
                    raise NotImplementedError('tabulated nets with different interval counts')
                d.tab_srmin[k] = float(np.float32(cv['srmin']))
                d.tab_srmax[k] = float(np.float32(cv['srmax']))
                d.tab_coeff[k] = arr(c)
            d.tab_ngrids = int(ngrids)
        for t in range(nt):
            net = P['fitting_net_%d' % t]
            for l in range(3 if atomic else 4):
                d.fit_kernel[t][l] = arr(net['Dense_%d' % l]['kernel'])
                b = net['Dense_%d' % l].get('bias')
                d.fit_bias[t][l] = arr(b) if b is not None else None
            for l in (1, 2):
                d.fit_dt[t][l] = arr(net['dt%d' % l]) if ('dt%d' % l) in net else None
        h = c_vp()
        check(lib().dpmd_model_create(ctypes.byref(d), int(device), ctypes.byref(h)))
        self.handle = h
        self.params = params
        self.ntypes = nt
        self.nsel_mask = int(d.nsel_mask)
        self.atomic = int(d.atomic)
        self.rcut = float(params['rcut'])

    def __del__(self):
        if getattr(self, 'handle', None) and _lib is not None:
            _lib.dpmd_model_destroy(self.handle)
            self.handle = None


class DeviceSystem:
    """Atoms resident on the GPU in the library's type-sorted layout.

    Two forms: ``DeviceSystem(model, type_idx, nreplica)`` holds nreplica copies of one
    composition (replicas, ring-polymer beads; arrays shaped (nreplica, natoms, ...)), and
    ``DeviceSystem(model, segments=[type_idx_0, type_idx_1, ...])`` holds independent
    systems of uneven size (arrays flat over all atoms, segments concatenated).
    """

    def __init__(self, model, type_idx=None, nreplica=1, segments=None):
        self.model = model
        h = c_vp()
        if segments is None:
            self.ragged = False
            self.type_idx = np.ascontiguousarray(np.asarray(type_idx, np.int32))
            self.natoms = len(self.type_idx)
            self.nrep = int(nreplica)
            self.seg_natoms = [self.natoms] * self.nrep
            check(lib().dpmd_system_create(model.handle, self.natoms, self.nrep, ptr(self.type_idx),
                                           ctypes.byref(h)))
        else:
            self.handle = None
            self._layout(segments)
            check(lib().dpmd_system_create_ragged(model.handle, self.nrep, ptr(np.asarray(self.seg_natoms, np.int32)),
                                                  ptr(self.type_idx), ctypes.byref(h)))
        self.handle = h
        self.knbr = None

    def _layout(self, segments):
        segs = [np.asarray(t, np.int32).reshape(-1) for t in segments]
        if not segs:
            raise ValueError('need at least one segment')
        self.ragged = True
        self.seg_natoms = [len(t) for t in segs]
        self.type_idx = np.ascontiguousarray(np.concatenate(segs))
        self.natoms = len(self.type_idx)
        self.nrep = len(segs)
        self.seg_offsets = np.concatenate([[0], np.cumsum(self.seg_natoms)]).astype(np.int64)

    def reshape(self, segments):
        """Lay the system out for new segments, reusing device buffers (all state is reset)."""
        self._layout(segments)
        check(lib().dpmd_system_reshape(self.handle, self.nrep, ptr(np.asarray(self.seg_natoms, np.int32)),
                                        ptr(self.type_idx)))
        self.knbr = None

    def split(self, arr):
        """Per-segment views of a flat (natoms_total, ...) array of a ragged system."""
        return [arr[self.seg_offsets[k]:self.seg_offsets[k + 1]] for k in range(self.nrep)]

    def _shape(self, *tail):
        return ((self.natoms,) if self.ragged else (self.nrep, self.natoms)) + tail

    def __del__(self):
        if getattr(self, 'handle', None) and _lib is not None:
            _lib.dpmd_system_destroy(self.handle)
            self.handle = None

    def set_box(self, box3):
        """Box lengths (3,) or a cell (3, 3) (rows: lattice vectors) for every replica."""
        b = np.asarray(box3, np.float64)
        if b.size == 9:
            self.set_cells(np.broadcast_to(b.reshape(1, 3, 3), (self.nrep, 3, 3)))
            return
        b = np.ascontiguousarray(b.reshape(3))
        check(lib().dpmd_system_set_box(self.handle, ptr(b)))

    def set_cells(self, cells):
        """Per-replica cells (nrep, 3, 3), rows the lattice vectors; a general cell drops the
        neighbour list."""
        c = np.ascontiguousarray(np.asarray(cells, np.float64).reshape(self.nrep, 3, 3))
        check(lib().dpmd_system_set_cells(self.handle, ptr(c)))

    def set_boxes(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, np.float64).reshape(self.nrep, 3))
        check(lib().dpmd_system_set_boxes(self.handle, ptr(b)))

    def set_images(self, nimg):
        n = np.ascontiguousarray(np.asarray(nimg, np.int32).reshape(self.nrep))
        check(lib().dpmd_system_set_images(self.handle, ptr(n)))

    def set_image_grid(self, k):
        """Images per axis (3,) or (nrep, 3) of supercell replicas."""
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32).reshape(-1, 3), (self.nrep, 3)), np.int32)
        check(lib().dpmd_system_set_image_grid(self.handle, ptr(kk)))

    def set_rng(self, seeds=None, steps=None):
        sd = None if seeds is None else np.ascontiguousarray(np.asarray(seeds, np.uint64).reshape(self.nrep))
        st = None if steps is None else np.ascontiguousarray(np.asarray(steps, np.uint64).reshape(self.nrep))
        check(lib().dpmd_system_set_rng(self.handle, ptr(sd), ptr(st)))

    def get_rng(self):
        sd = np.zeros(self.nrep, np.uint64)
        st = np.zeros(self.nrep, np.uint64)
        check(lib().dpmd_system_get_rng(self.handle, ptr(sd), ptr(st)))
        return sd, st

    def set_positions(self, pos):
        p = np.ascontiguousarray(np.asarray(pos, np.float64).reshape(self._shape(3)))
        check(lib().dpmd_system_set_positions(self.handle, ptr(p)))

    def get_positions(self):
        p = np.empty(self._shape(3))
        check(lib().dpmd_system_get_positions(self.handle, ptr(p)))
        return p

    def set_velocities(self, vel):
        v = np.ascontiguousarray(np.asarray(vel, np.float64).reshape(self._shape(3)))
        check(lib().dpmd_system_set_velocities(self.handle, ptr(v)))

    def get_velocities(self):
        v = np.empty(self._shape(3))
        check(lib().dpmd_system_get_velocities(self.handle, ptr(v)))
        return v

    def set_masses(self, mass):
        """Per-atom masses: (natoms) of one replica for the replica form, all atoms otherwise."""
        m = np.asarray(mass, np.float64).reshape(-1)
        if not self.ragged:
            m = np.tile(m.reshape(self.natoms), self.nrep)
        m = np.ascontiguousarray(m)
        check(lib().dpmd_system_set_masses(self.handle, ptr(m)))

    def nlist_allocate(self, rc_skin, buffer_ratio=1.2, knbr_override=None):
        out = np.zeros(MAXT, np.int32)
        ko = None if knbr_override is None else np.ascontiguousarray(np.asarray(knbr_override, np.int32))
        check(lib().dpmd_nlist_allocate(self.handle, float(rc_skin), float(buffer_ratio), ptr(ko), ptr(out)))
        self.knbr = [int(k) for k in out[:self.model.ntypes]]
        self.rc_skin_last = float(rc_skin)
        return self.knbr

    def nlist_overflow(self):
        """True when the last build truncated a row at its capacity."""
        ov = np.zeros(1, np.int32)
        check(lib().dpmd_nlist_overflow(self.handle, ptr(ov)))
        return bool(ov[0])

    def nlist_update(self):
        ov = np.zeros(1, np.int32)
        check(lib().dpmd_nlist_update(self.handle, ptr(ov)))
        return bool(ov[0])

    def nlist_export(self, j):
        cap = self.knbr[j] - 1
        out = np.empty(self._shape(max(cap, 1)), np.int32)
        check(lib().dpmd_nlist_export(self.handle, int(j), ptr(out)))
        return out[..., :cap]

    def compute(self, precision='fp32', want_atom=False):
        e = np.zeros(self.nrep)
        f = np.zeros(self._shape(3))
        v = np.zeros((self.nrep, 9))
        ea = np.zeros(self._shape()) if want_atom else None
        check(lib().dpmd_compute(self.handle, PREC[precision], ptr(e), ptr(f), ptr(v), ptr(ea)))
        return e, f, v.reshape(self.nrep, 3, 3), ea

    def set_deterministic(self, on=True):
        """Bit-reproducible forces: pair gradients gathered per atom in a fixed order instead
        of FP32 atomics (dpmd_system_set_deterministic)."""
        check(lib().dpmd_system_set_deterministic(self.handle, 1 if on else 0))

    def compute_noalloc(self, precision='fp32'):
        """dpmd_compute_noalloc: energy, forces and virial with the work buffers in the
        library's per-GPU scratch array (no device allocation in the call)."""
        e = np.zeros(self.nrep)
        f = np.zeros(self._shape(3))
        v = np.zeros((self.nrep, 9))
        check(lib().dpmd_compute_noalloc(self.handle, PREC[precision], ptr(e), ptr(f), ptr(v)))
        return e, f, v.reshape(self.nrep, 3, 3)

    def md_run(self, params, nsteps, record=True, out=None):
        """Run nsteps device steps; with record, the per-step positions and velocities in
        caller order, written into `out` = (pos, vel) when given (C-contiguous float64 arrays
        of shape (nsteps,) + the system's (.., 3) shape, e.g. slices of a trajectory)."""
        shape = (nsteps,) + self._shape(3)
        if record and out is not None:
            tp, tv = out
            for a in (tp, tv):
                if a.shape != shape or a.dtype != np.float64 or not a.flags.c_contiguous:
                    raise ValueError('md_run out arrays must be C-contiguous float64 of shape %s' % (shape,))
        else:
            tp = np.empty(shape) if record else None
            tv = np.empty(shape) if record else None
        err = np.zeros(1, np.uint32)
        check(lib().dpmd_md_run(self.handle, ctypes.byref(params), int(nsteps), ptr(tp), ptr(tv), ptr(err)))
        return tp, tv, int(err[0])

    def get_boxes(self):
        """Current box lengths (nrep, 3); they move under NPT."""
        b = np.zeros((self.nrep, 3))
        check(lib().dpmd_system_get_boxes(self.handle, ptr(b)))
        return b

    def box_trajectory(self, nsteps):
        """Per-step boxes (nsteps, nrep, 3) of the last recorded md_run chunk."""
        b = np.zeros((nsteps, self.nrep, 3))
        check(lib().dpmd_md_box_trajectory(self.handle, ptr(b)))
        return b

    def set_dplr(self, wc_model, q_atoms, q_wc, beta, grid):
        """DPLR long-range term: wc_model a DeviceModel of an atomic (Wannier-centroid) model,
        q_atoms / q_wc charges per type (ntypes), PPPM beta and grid (3,); None removes it."""
        if wc_model is None:
            check(lib().dpmd_system_set_dplr(self.handle, None, None, None, 0.0, None))
            self._wc_model = None
            return
        qa = np.ascontiguousarray(np.asarray(q_atoms, np.float64).reshape(-1))
        qw = np.ascontiguousarray(np.asarray(q_wc, np.float64).reshape(-1))
        g = np.ascontiguousarray(np.asarray(grid, np.int32).reshape(3))
        check(lib().dpmd_system_set_dplr(self.handle, wc_model.handle, ptr(qa), ptr(qw), float(beta), ptr(g)))
        self._wc_model = wc_model  # keeps the centroid model alive while the system uses it
        self._n_wc = int(sum(np.isin(self.type_idx, [t for t in range(wc_model.ntypes)
                                                     if int(wc_model.nsel_mask) >> t & 1])))

    def atomic_predict(self):
        """Outputs of the system's atomic model: (nrep, M, 3) vectors or (nrep, M, 9) symmetric
        tensors for the M atoms of the selected types per replica, in caller order."""
        ncomp = 3 if self.model.atomic == 1 else 9
        sel = [t for t in range(self.model.ntypes) if self.model.nsel_mask >> t & 1]
        m = int(np.isin(self.type_idx, sel).sum()) * (1 if self.ragged else self.nrep)
        out = np.zeros(max(m, 1) * ncomp)
        check(lib().dpmd_atomic_predict(self.handle, ptr(out)))
        out = out[:m * ncomp]
        return out.reshape(-1, ncomp) if self.ragged else out.reshape(self.nrep, -1, ncomp)

    def dplr_get(self):
        """(long-range energy per replica (nrep,), centroids) of the last evaluation: centroids
        (nrep, M, 3) for the M atoms of the selected types per replica in caller order (ragged
        systems: (sum M, 3) segment by segment)."""
        m = self._n_wc * (1 if self.ragged else self.nrep)
        e = np.zeros(self.nrep)
        c = np.zeros((max(m, 1), 3))
        check(lib().dpmd_dplr_get(self.handle, ptr(e), ptr(c)))
        c = c[:m]
        return e, (c if self.ragged else c.reshape(self.nrep, -1, 3))

    def set_fixed(self, fixed):
        """Fixed-atom mask in caller order (flat over all atoms of all replicas), or None."""
        f = None if fixed is None else np.ascontiguousarray(np.asarray(fixed).reshape(-1), np.int32)
        check(lib().dpmd_system_set_fixed(self.handle, ptr(f)))

    def nlist_update_rate(self):
        r = np.zeros(1)
        check(lib().dpmd_nlist_update_rate(self.handle, ptr(r)))
        return float(r[0])

    def nlist_rebuilds(self):
        """Drift-triggered neighbour-list rebuilds inside md_run since the last allocation."""
        c = ctypes.c_int64()
        check(lib().dpmd_nlist_rebuilds(self.handle, ctypes.byref(c)))
        return int(c.value)

    def npt_state(self):
        """The 64-double NPT barostat state (log box scales, box momenta and masses, dU/deps,
        reference box, barostat chain; layout in md.hip)."""
        st = np.zeros(64)
        check(lib().dpmd_npt_state_get(self.handle, ptr(st)))
        return st

    def pimd_setup(self, C, freqs, c1, c2):
        a = [np.ascontiguousarray(np.asarray(x, np.float64)) for x in (C, freqs, c1, c2)]
        check(lib().dpmd_pimd_setup(self.handle, *[ptr(x) for x in a]))

    def kinetic_energy(self):
        ke = np.zeros(self.nrep)
        check(lib().dpmd_kinetic_energy(self.handle, ptr(ke)))
        return ke

    def thermostat_energy(self):
        e = np.zeros(self.nrep)
        check(lib().dpmd_thermostat_energy(self.handle, ptr(e)))
        return e

    def thermostat_state(self, state=None):
        """Get (state=None) or set the 64-double per-segment thermostat state."""
        if state is None:
            st = np.zeros((self.nrep, 64))
            check(lib().dpmd_thermostat_state_get(self.handle, ptr(st)))
            return st
        st = np.ascontiguousarray(np.asarray(state, np.float64).reshape(self.nrep, 64))
        check(lib().dpmd_thermostat_state_set(self.handle, ptr(st)))
        return None

    def timing(self, on=None, which=None):
        if on is not None:
            check(lib().dpmd_timing_enable(self.handle, 1 if on else 0))
            return None
        ms = np.zeros(1)
        cnt = np.zeros(1, np.int64)
        check(lib().dpmd_timing_get(self.handle, int(which), ptr(ms), ptr(cnt)))
        return float(ms[0]), int(cnt[0])

    def pair_count(self):
        p = np.zeros(1)
        check(lib().dpmd_pair_count(self.handle, ptr(p)))
        return float(p[0])

    def synchronize(self):
        check(lib().dpmd_synchronize(self.handle))

"""Frame evaluation on the GPU engine.

Boxes (orthorhombic lengths, or general cells whose rows are the lattice vectors) with a
perpendicular width shorter than 2*rcut, which deepmd-jax evaluates on its dense
lattice-image path (deepmd_jax/utils.py:66-78, deepmd_jax/data.py:347-364), are evaluated
here as a k1 x k2 x k3 periodic supercell with every k_d*w_d >= 2*rcut: each atom's images
have identical environments, so E = E_super / prod(k), the force on an atom is the force on
its first image and the virial is W_super / prod(k).
"""
import numpy as np

from . import _lib


def canonical_cell(box):
    """(3,3) cell, rows the lattice vectors, from a scalar, (3,), (9,) or (3,3) box."""
    b = np.asarray(box, np.float64)
    if b.size == 1:
        return np.eye(3) * float(b.reshape(()))
    if b.size == 3:
        return np.diag(b.reshape(3))
    if b.size == 9:
        return b.reshape(3, 3).copy()
    raise ValueError('box must be a scalar, (3,), (9,) or (3,3)')


def canonical_box(box):
    """(3,) box lengths of an orthorhombic box (deepmd_jax/md.py:208-215 keeps those as (3,)),
    the (3,3) cell of a general one."""
    c = canonical_cell(box)
    if np.any(c - np.diag(np.diag(c))):
        return c
    return np.diag(c).copy()


def cell_matrix(box):
    b = np.asarray(box, np.float64)
    return np.diag(b) if b.shape == (3,) else b.reshape(3, 3)


def cell_volume(box):
    b = np.asarray(box, np.float64)
    return float(np.prod(b)) if b.shape == (3,) else abs(float(np.linalg.det(b.reshape(3, 3))))


def cell_widths(box):
    """Distances between opposite faces: the box lengths, or 1/|column d of H^-1| of a cell."""
    b = np.asarray(box, np.float64)
    if b.shape == (3,):
        return b.copy()
    return 1.0 / np.linalg.norm(np.linalg.inv(b.reshape(3, 3)), axis=0)


def supercell_factors(box, rcut):
    return np.maximum(1, np.ceil(2.0 * rcut / cell_widths(box) * (1 + 1e-12))).astype(int)


def replicate(coord, box, type_idx, k):
    """Coordinates of the k[0] x k[1] x k[2] supercell (images shifted by whole lattice
    vectors); image 0 holds the original atoms.  Returns (coords, supercell box, types)."""
    shifts = np.stack(np.meshgrid(np.arange(k[0]), np.arange(k[1]), np.arange(k[2]), indexing='ij'), -1).reshape(-1, 3)
    b = np.asarray(box, np.float64)
    c = (coord[None, :, :] + (shifts @ cell_matrix(b))[:, None, :]).reshape(-1, 3)
    sbox = b * k if b.shape == (3,) else np.asarray(k, np.float64)[:, None] * b.reshape(3, 3)
    return c, sbox, np.tile(type_idx, len(shifts))


def dplr_setup(params, device=0):
    """(centroid DeviceModel, charge per type of the atoms, of the centroids, beta, resolution) of a
    DPLR model's parameters (deepmd_jax/train.py:305-313), or None for other models."""
    if params.get('type') != 'dplr':
        return None
    wc_model, wc_var = params['dplr_wannier_model_and_variables']
    wp = wc_model.params
    nt = int(params['ntypes'])
    nsel = [int(t) for t in wp['nsel']]
    q_wc = np.asarray(params['dplr_q_wc'], np.float64).reshape(-1)
    if len(q_wc) != len(nsel):
        raise ValueError('dplr_q_wc must correspond to atomic_sel of the Wannier model.')
    qw = np.zeros(nt)
    qw[nsel] = q_wc
    qa = np.asarray(params['dplr_q_atoms'], np.float64).reshape(-1)[:nt]
    return (_lib.DeviceModel(wp, wc_var, device), qa, qw, float(params.get('dplr_beta', 0.4)),
            float(params.get('dplr_resolution', 5)))


def dplr_grid(box3, beta, resolution):
    """PPPM grid of get_p3mlr_grid_size (deepmd_jax/utils.py:174-175)."""
    return (np.asarray(box3, np.float64) * beta * resolution).astype(int)


class Evaluator:
    """Energies, forces and virials of frames for one model (box: (3,) lengths or a (3,3)
    cell).  A DPLR model adds its long-range term to every frame (orthorhombic boxes,
    deepmd_jax/train.py:796-797)."""

    def __init__(self, params, variables, device=0, model=None):
        self.params = params
        self.model = model if model is not None else _lib.DeviceModel(params, variables, device)
        self.rcut = float(params['rcut'])
        self._systems = {}
        self.dplr = dplr_setup(params, device)

    def _system(self, type_idx, nrep):
        key = (tuple(np.asarray(type_idx).tolist()), nrep)
        s = self._systems.get(key)
        if s is None:
            s = _lib.DeviceSystem(self.model, type_idx, nrep)
            self._systems[key] = s
        return s

    def batch_key(self, box):
        """Frames with equal keys can share one batch (one launch, a box per frame): the same
        images per axis and, for DPLR, the same PPPM grid."""
        b = canonical_box(box)
        k = tuple(int(x) for x in supercell_factors(b, self.rcut))
        if self.dplr is not None and b.shape == (3,):
            _, _, _, beta, res = self.dplr
            return k, tuple(int(x) for x in dplr_grid(b, beta, res))
        return k

    def _load(self, coords, boxes, type_idx):
        """A system holding the frames (supercells of images for narrow boxes, all with the
        images per axis of the first frame), a cell per frame, positions set."""
        coords = np.asarray(coords, np.float64)
        boxes = [canonical_box(b) for b in boxes]
        type_idx = np.asarray(type_idx)
        k = supercell_factors(boxes[0], self.rcut)
        nk = int(np.prod(k))
        if nk > 1:
            cs = [replicate(c, b, type_idx, k) for c, b in zip(coords, boxes)]
            sc, sboxes, stype = np.stack([c[0] for c in cs]), [c[1] for c in cs], cs[0][2]
        else:
            sc, sboxes, stype = coords, boxes, type_idx
        sys = self._system(stype, len(coords))
        sys.set_cells(np.stack([cell_matrix(b) for b in sboxes]))
        sys.set_positions(sc)
        return sys, boxes, k, nk

    def evaluate_batch(self, coords, box3, type_idx, precision='fp32'):
        """coords (F,N,3) sharing one box -> (E (F,), F (F,N,3), virial (F,3,3))."""
        return self.evaluate_frames(coords, [box3] * len(coords), type_idx, precision)

    def evaluate_frames(self, coords, boxes, type_idx, precision='fp32'):
        """coords (F,N,3), one box per frame (frames of one batch_key) -> (E (F,), F (F,N,3),
        virial (F,3,3)); the frames run as the segments of one system."""
        N = np.asarray(coords).shape[1]
        sys, boxes, k, nk = self._load(coords, boxes, type_idx)
        if self.dplr is not None:
            if boxes[0].shape != (3,):
                raise ValueError('For "dplr" currently only orthorhombic boxes are supported.')
            # each frame's PPPM on its cell (a narrow box as its supercell of images: the
            # short-range model sees the images, the long-range term the cell)
            wdm, qa, qw, beta, res = self.dplr
            if nk > 1:
                sys.set_image_grid(k)
            sys.nlist_allocate(self.rcut, 1.2)
            sys.set_dplr(wdm, qa, qw, beta, dplr_grid(boxes[0], beta, res))
            e, f, v, _ = sys.compute(precision)
            return e, f[:, :N], v
        sys.nlist_allocate(self.rcut, 1.2)
        e, f, v, _ = sys.compute(precision)
        return e / nk, f[:, :N], v / nk

    def evaluate_atomic_batch(self, coords, box3, type_idx):
        """Atomic model outputs of frames sharing one box: (F, M, 3) vectors or (F, M, 9) tensors for
        the M atoms of the selected types, in caller order."""
        return self.evaluate_atomic_frames(coords, [box3] * len(coords), type_idx)

    def evaluate_atomic_frames(self, coords, boxes, type_idx):
        """evaluate_atomic_batch with one box per frame (frames of one batch_key)."""
        sys, _, _, _ = self._load(coords, boxes, type_idx)
        sys.nlist_allocate(self.rcut, 1.2)
        sel = [int(t) for t in self.params['nsel']]
        m = int(np.isin(np.asarray(type_idx), sel).sum())
        return sys.atomic_predict()[:, :m]  # image 0 holds the original atoms

    def evaluate_frame(self, coord, box3, type_idx, precision='fp32'):
        e, f, v = self.evaluate_batch(np.asarray(coord)[None], box3, type_idx, precision)
        return e[0], f[0], v[0]


def energy_forces_stress(evaluator, positions, cell, type_idx, precision='fp32'):
    """Energy, forces and stress of one frame as the reference's ASE calculator returns them
    (deepmd_jax/ase_calc.py:79-143): stress = (1/V) dU/d(strain) in Voigt order
    (xx, yy, zz, yz, xz, xy), eV/A^3 -- the negative of jax_md.quantity.stress.  With the
    engine's virial W = -dU/d(strain) that is -W / V."""
    box3 = canonical_box(cell)
    e, f, w = evaluator.evaluate_frame(positions, box3, type_idx, precision)
    s = -np.asarray(w) / cell_volume(box3)
    return float(e), f, np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])

"""model.pkl I/O (deepmd_jax/utils.py:225-237) with a restricted unpickler.

A deepmd-jax model.pkl is a pickle of {'model': DPModel(params), 'variables':
{'params': {...}}}.  Loading resolves only an allowlist of globals: numpy array
reconstruction, this package's DPModel, and the jax / flax names such a file
refers to, which are mapped to plain-numpy stand-ins.  Anything else is refused.
"""
import pickle

import numpy as np

from .dpmodel import DPModel


def _jax_reconstruct_array(fun, args, arr_state, aval_state):
    arr = fun(*args)
    if arr_state is not None:
        arr.__setstate__(arr_state)
    return np.asarray(arr)


class _Stub:
    """Stand-in for flax/jax objects embedded in a pickled module's state."""

    def __init__(self, *a, **k):
        pass

    def __setstate__(self, state):
        self.state = state


_ALLOWED = {
    ('numpy', 'ndarray'), ('numpy', 'dtype'), ('numpy.core.multiarray', '_reconstruct'),
    ('numpy._core.multiarray', '_reconstruct'), ('numpy.core.multiarray', 'scalar'),
    ('numpy._core.multiarray', 'scalar'), ('builtins', 'slice'), ('builtins', 'range'),
    ('collections', 'OrderedDict'), ('copyreg', '_reconstructor'), ('builtins', 'object'),
}


class _Unpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _ALLOWED:
            return super().find_class(module, name)
        if module in ('deepmd_jax.dpmodel',) and name == 'DPModel':
            return DPModel
        if module.startswith('jax') and name == '_reconstruct_array':
            return _jax_reconstruct_array
        if module.startswith('flax') or module.startswith('jax') or module.startswith('deepmd_jax'):
            return _Stub
        raise pickle.UnpicklingError('global %s.%s is not allowed in a model file' % (module, name))


def _to_numpy(tree):
    if isinstance(tree, dict):
        return {k: _to_numpy(v) for k, v in tree.items()}
    if isinstance(tree, (list, tuple)):
        return type(tree)(_to_numpy(v) for v in tree)
    if hasattr(tree, '__array__') and not isinstance(tree, np.ndarray):
        return np.asarray(tree)
    return tree


def load_model(path, replicate=True):
    """Return (DPModel, variables) with variables as numpy arrays."""
    with open(path, 'rb') as f:
        m = _Unpickler(f).load()
    print("# Model loaded from '%s'." % path)
    model = m['model']
    if not isinstance(model, DPModel):
        raise ValueError('model.pkl does not hold a DPModel')
    return model, _to_numpy(m['variables'])


def save_model(path, model, variables):
    with open(path, 'wb') as f:
        pickle.dump({'model': model, 'variables': _to_numpy(variables)}, f)
    print("# Model saved to '%s'." % path)


# ---------------------------------------------------------------------------------
# Model compression (deepmd_jax/utils.py:250-297): each type-pair embedding net is
# replaced by a piecewise quintic in the normalised switch value s~, one polynomial per
# interval of a uniform grid over [s~(r >= rcut), s~(rmin)], matching the net's value,
# first and second derivative at every grid node (C2 quintic Hermite).  The kernels then
# evaluate the table (deepmd_jax/utils.py:126-151) instead of the MLP.

def sr(r, rcut):
    """Switched 1/r of deepmd_jax/utils.py:29-32 (scalar or numpy)."""
    t = r / rcut
    return (r > 1e-14) * (t < 1) / (r + 1e-15) * (1 - 3 * t ** 2 + 2 * t ** 3)


def embedding_net_jet(net, x):
    """Value, first and second derivative in x of an uncompressed embedding net
    (deepmd_jax/utils.py:107-125: tanh layers, repeat / mean-pool residual), FP64.
    x (M,) -> three (M, W) arrays."""
    x = np.asarray(x, np.float64)
    f, d1, d2 = x[:, None], np.ones((len(x), 1)), np.zeros((len(x), 1))
    li = 0
    while 'Dense_%d' % li in net:
        K = np.asarray(net['Dense_%d' % li]['kernel'], np.float64)
        b = np.asarray(net['Dense_%d' % li]['bias'], np.float64)
        z, z1, z2 = f @ K + b, d1 @ K, d2 @ K
        y = np.tanh(z)
        g = 1.0 - y * y
        y1 = g * z1
        y2 = g * z2 - 2.0 * y * y1 * z1
        if li > 0:
            win, wout = f.shape[1], y.shape[1]
            if wout % win == 0:
                Z = wout // win
                y, y1, y2 = (y + np.repeat(f, Z, axis=1), y1 + np.repeat(d1, Z, axis=1),
                             y2 + np.repeat(d2, Z, axis=1))
            else:
                Z = win // wout
                pool = lambda a: a.reshape(a.shape[0], Z, wout).mean(1)
                y, y1, y2 = y + pool(f), y1 + pool(d1), y2 + pool(d2)
        f, d1, d2 = y, y1, y2
        li += 1
    return f, d1, d2


def quintic_hermite(nodes, f0, f1, f2):
    """Per-interval power-basis coefficients (Ngrids, 6, W) of the C2 quintic through
    value / slope / curvature at consecutive nodes; row c holds the coefficient of
    (x - node_k)^(5-c), the layout of deepmd_jax/utils.py:286 (PPoly.c transposed)."""
    h = np.diff(nodes)[:, None]
    a0, a1, a2 = f0[:-1], f1[:-1], 0.5 * f2[:-1]
    d0 = f0[1:] - (a0 + h * (a1 + h * a2))
    d1 = f1[1:] - (a1 + 2.0 * h * a2)
    d2 = f2[1:] - 2.0 * a2
    c3 = (20.0 * d0 - 8.0 * h * d1 + h * h * d2) / (2.0 * h ** 3)
    c4 = (-30.0 * d0 + 14.0 * h * d1 - 2.0 * h * h * d2) / (2.0 * h ** 4)
    c5 = (12.0 * d0 - 6.0 * h * d1 + h * h * d2) / (2.0 * h ** 5)
    return np.stack([c5, c4, c3, a2, a1, a0], axis=1)


def compress_model(model, variables, Ngrids=1024, rmin=0.6):
    """Tabulate the embedding nets of an se_a energy model (deepmd_jax/utils.py:250-297).

    Sets model.params['is_compressed'] and adds variables['compress_var'][name] =
    {'srmin', 'srmax', 'poly_coeff' (Ngrids, 6, W)} (float32) for every embedding net
    name = 'embedding_net_{i*ntypes+j}', whose s~ range comes from centre type i.
    Prints the mean / max error of value, first and second derivative on a 10x finer grid."""
    params = model.params
    if params.get('use_mp', False) or params.get('atomic', False):
        raise NotImplementedError('compression of message-passing or atomic models is not supported')
    model.params['is_compressed'] = True
    Y = len(params['sr_mean'])
    mean = np.asarray(params['sr_mean'], np.float64)
    std = np.asarray(params['sr_std'], np.float64)
    lo_hi = (np.array([0.0, sr(rmin, params['rcut'])])[None] - mean[:, None]) / std[:, None]
    P = variables['params']
    variables['compress_var'] = {}
    errs = [[], [], []]
    for k in range(Y * Y):
        name = 'embedding_net_%d' % k
        s0, s1 = lo_hi[k // Y]
        nodes = np.linspace(s0, s1, Ngrids + 1)
        f0, f1, f2 = embedding_net_jet(P[name], nodes)
        coeff = quintic_hermite(nodes, f0, f1, f2)
        cv = {'srmin': np.float32(s0), 'srmax': np.float32(s1), 'poly_coeff': coeff.astype(np.float32)}
        variables['compress_var'][name] = cv
        xs = np.linspace(s0, s1 * (1 - 1e-6), Ngrids * 10 + 1)
        t0, t1, t2 = embedding_net_jet(P[name], xs)
        p0, p1, p2 = tabulated_jet(cv, xs)
        for e, a, b in zip(errs, (t0, t1, t2), (p0, p1, p2)):
            e.append((a - b).ravel())
    e0, e1, e2 = (np.concatenate(e) for e in errs)
    print('# Model Compressed: %d embedding nets, %d intervals' % (Y * Y, Ngrids))
    print('# Compression (0,1,2)-order error: Mean = (%.1e,%.1e,%.1e), Max = (%.1e,%.1e,%.1e)'
          % (e0.mean(), e1.mean(), e2.mean(), e0.max(), e1.max(), e2.max()))
    return model, variables


def tabulated_jet(cv, x):
    """Value and first two derivatives of a tabulated net at s~ = x (FP64 arithmetic on the
    stored coefficients; the interval rule of deepmd_jax/utils.py:130-135)."""
    srmin, srmax = np.float32(cv['srmin']), np.float32(cv['srmax'])
    c = np.asarray(cv['poly_coeff'], np.float64)
    n = c.shape[0]
    delta = np.float32((srmax - srmin) / np.float32(n))
    xr = np.asarray(x, np.float64) - float(srmin)
    idx = np.trunc(xr / float(delta)).astype(np.int64)
    x0 = (xr - idx * float(delta))[:, None]
    cc = c[np.clip(idx, 0, n - 1)]
    v = d1 = d2 = 0.0
    for i in range(6):
        v = v * x0 + cc[:, i]
    for i in range(5):
        d1 = d1 * x0 + (5 - i) * cc[:, i]
    for i in range(4):
        d2 = d2 * x0 + (5 - i) * (4 - i) * cc[:, i]
    return v, d1, d2


def get_p3mlr_grid_size(box3, beta, resolution=5):
    """PPPM grid of a box, (box * beta * resolution) truncated (deepmd_jax/utils.py:174-176)."""
    return tuple(int(m) for m in (np.asarray(box3, np.float64) * beta * resolution).astype(int))


def get_p3mlr_fn(box3_ref, beta, M=None, resolution=5, device=0):
    """PPPM long-range energy on the GPU (deepmd_jax/utils.py:178-222): returns
    p3mlr_fn(coord_N3, q_N, box3) -> energy (eV) for charges (e) at coord (A) in the orthorhombic
    box3, TSC assignment on the grid frozen from box3_ref (or M).  The reference differentiates
    its function with jax; here p3mlr_fn.energy_force_virial(coord_N3, q_N, box3) returns the
    energy, -dE/dx (N, 3) and the virial -dE/dstrain (3, 3) directly."""
    from ._lib import pppm_compute
    grid = get_p3mlr_grid_size(box3_ref, beta, resolution) if M is None else tuple(int(m) for m in M)

    def energy_force_virial(coord_N3, q_N, box3):
        return pppm_compute(box3, beta, grid, coord_N3, q_N, device)

    def p3mlr_fn(coord_N3, q_N, box3):
        return energy_force_virial(coord_N3, q_N, box3)[0]

    p3mlr_fn.energy_force_virial = energy_force_virial
    p3mlr_fn.grid = grid
    return p3mlr_fn


def periodic_replicate(n_copy, coord, box, type_idx=None, force=None, velocity=None):
    """Replicate a periodic system n_copy times per direction (deepmd_jax/utils.py:299-327):
    returns [coord, box] plus type_idx, force, velocity when given.  As there, copy i along
    direction d is shifted by i times column d of the (3,3) box (the box lengths of an
    orthorhombic one), directions nest with the last outermost, and the returned box is
    diag(n_copy) @ box (or n_copy * box for lengths)."""
    n_copy = (np.asarray(n_copy) * np.ones(3)).astype(int)
    box = np.array(box, np.float64)
    coord = np.asarray(coord, np.float64)
    if box.size <= 3:
        return_box = n_copy * box
        cell = np.diag(box * np.ones(3))
    else:
        cell = box.reshape(3, 3)
        return_box = np.diag(n_copy) @ cell
    for d in range(3):
        coord = np.concatenate([(coord + i * cell[:, d])[None] for i in range(n_copy[d])]).reshape(-1, 3)
    ret = [coord, return_box]
    n = int(n_copy.prod())
    if type_idx is not None:
        ret.append(np.tile(type_idx, n))
    if force is not None:
        ret.append(np.concatenate([force] * n))
    if velocity is not None:
        ret.append(np.concatenate([velocity] * n))
    return ret


def dplr_charges(type_idx, q_atoms, q_wc, nsel, ntypes):
    """Per-atom charges (N,) and per-centroid charges (M,) in caller order over the atoms of the
    nsel types (deepmd_jax/utils.py:355-366)."""
    type_idx = np.asarray(type_idx)
    nsel = np.asarray(list(nsel))
    pos = np.full(ntypes, -1, dtype=int)
    pos[nsel] = np.arange(len(nsel))
    sel = np.isin(type_idx, nsel)
    return np.asarray(q_atoms)[type_idx], np.asarray(q_wc)[pos[type_idx[sel]]]

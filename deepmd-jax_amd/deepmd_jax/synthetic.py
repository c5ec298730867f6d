"""Synthetic water boxes and the seeded random-init se_a model.

The water generator builds liquid-density water: molecules on a g^3 grid
(g = ceil(n^(1/3))), O on a site, the two H atoms at O + R.(+-0.9572 sin 52.26 deg,
0.9572 cos 52.26 deg, 0) with R a Haar-random rotation, every atom jittered
uniformly, positions wrapped into [0, L).  Atom order is O,H,H per molecule and
the type index is 0 for O and 1 for H.

The model follows the parameter layout of deepmd-jax's `DPModel`
(deepmd_jax/dpmodel.py:30-100, names as flax assigns them:
`embedding_net_{i*Y+j}/Dense_{0,1,2}`, `fitting_net_{t}/Dense_{0..3}`, `dt1`, `dt2`,
`Tbias`) and its `params` dict (deepmd_jax/train.py:288-304).  Normalisation
statistics follow `DatasetLeaf._get_stats` (deepmd_jax/data.py:123-147).
"""
import numpy as np

AVOGADRO = 6.02214076e23
WATER_MOLAR_MASS = 18.015
OH_BOND = 0.9572
HOH_ANGLE_DEG = 104.52
MASSES = (15.9994, 1.0078)


def water_box_length(n_molecules, density=1.0):
    """Cubic box edge (Angstrom) holding n water molecules at `density` g/cm^3."""
    vol_cm3 = n_molecules * WATER_MOLAR_MASS / AVOGADRO / density
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def _haar_rotations(rng, n):
    """n Haar-random rotations; draws the same normals as n successive (3,3) draws."""
    q, r = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    neg = np.linalg.det(q) < 0
    q[neg, :, 0] = -q[neg, :, 0]
    return q


def water_box(n_molecules, seed=0, jitter=0.1, density=1.0, permute_sites=False):
    """Return (coord (3n,3) float64, box (3,) float64, type_idx (3n,) int).

    permute_sites: take the first n sites of a seeded permutation of the g^3
    grid (used for request streams where n is not a perfect cube)."""
    rng = np.random.default_rng(seed)
    L = water_box_length(n_molecules, density)
    g = int(np.ceil(n_molecules ** (1.0 / 3.0) - 1e-9))
    a = L / g
    grid = np.stack(np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing='ij'), -1).reshape(-1, 3)
    if permute_sites:
        grid = grid[rng.permutation(len(grid))[:n_molecules]]
    else:
        grid = grid[:n_molecules]
    sites = (grid + 0.5) * a
    half = np.deg2rad(HOH_ANGLE_DEG / 2)
    local = np.array([[OH_BOND * np.sin(half), OH_BOND * np.cos(half), 0.0],
                      [-OH_BOND * np.sin(half), OH_BOND * np.cos(half), 0.0]])
    R = _haar_rotations(rng, n_molecules)
    coord = np.empty((n_molecules, 3, 3))
    coord[:, 0] = sites
    coord[:, 1:] = sites[:, None, :] + np.matmul(local[None], R.transpose(0, 2, 1))
    coord = coord.reshape(3 * n_molecules, 3)
    coord += rng.uniform(-jitter, jitter, coord.shape)
    coord = np.mod(coord, L)
    type_idx = np.tile(np.array([0, 1, 1]), n_molecules)
    return coord, np.full(3, L), type_idx


def maxwell_boltzmann(type_idx, temperature, seed=0, masses=MASSES):
    """Velocities (Angstrom/fs) at `temperature` K, centre-of-mass momentum removed."""
    from .constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    rng = np.random.default_rng([seed, 1])
    m = np.asarray(masses)[type_idx] * MASS_UNIT_CONVERSION
    kT = temperature * TEMP_UNIT_CONVERSION
    p = rng.standard_normal((len(type_idx), 3)) * np.sqrt(m * kT)[:, None]
    p -= p.mean(0, keepdims=True)
    return p / m[:, None]


def _switch(r, rcut):
    t = r / rcut
    return (r > 1e-14) * (t < 1) / (r + 1e-15) * (1 - 3 * t ** 2 + 2 * t ** 3)


def sr_stats(coord, box, type_idx, rcut, ntypes=2):
    """sr_mean, sr_std, Nnbrs over one frame, as DatasetLeaf._get_stats
    (deepmd_jax/data.py:123-147) computes them: every (centre, neighbour) pair
    including the zero self term, per centre type, minimum image (L >= 2 rcut)."""
    coord = np.asarray(coord, np.float64)
    L = np.asarray(box, np.float64).reshape(-1)[:3]
    if L.min() < 2 * rcut:
        raise ValueError('sr_stats needs a box of at least 2*rcut')
    order = np.argsort(type_idx, kind='stable')
    c = coord[order]
    t = np.asarray(type_idx)[order]
    d = c[None, :, :] - c[:, None, :]
    d = np.mod(d - L / 2, L) - L / 2
    r = np.linalg.norm(np.where(np.abs(d) > 1e-15, d, 1e-15), axis=-1)
    s = _switch(r, rcut)
    mean, std = [], []
    for i in range(ntypes):
        si = s[t == i]
        cnt = (si > 1e-15).sum()
        m1 = si.sum() / cnt
        mean.append(m1)
        std.append(np.sqrt((si ** 2).sum() / cnt - m1 ** 2))
    nnbrs = (s > 0).sum(1).mean() + 1
    return np.array(mean), np.array(std), float(nnbrs)


def init_model(seed=0, coord=None, box=None, type_idx=None, rcut=6.0,
               embed_widths=(25, 50, 100), fit_widths=(240, 240, 240), axis=16,
               ntypes=2, fit_dt=0.1):
    """Seeded random-init se_a model: returns (params dict, variables dict).

    Kernels ~ N(0, 1/fan_in), biases ~ N(0, 1), final fitting bias 0,
    fitting dt = fit_dt, Tbias = 0, Ebias = 0.  Normalisation statistics come
    from the given frame (default: the 64-molecule box, seed 0)."""
    if coord is None:
        coord, box, type_idx = water_box(64, seed=0)
    mean, std, nnbrs = sr_stats(coord, box, type_idx, rcut, ntypes)
    rng = np.random.default_rng([seed, 7])
    f32 = np.float32

    def dense(fan_in, fan_out, bias=True):
        k = (rng.standard_normal((fan_in, fan_out)) / np.sqrt(fan_in)).astype(f32)
        b = rng.standard_normal(fan_out).astype(f32) if bias else np.zeros(fan_out, f32)
        return {'kernel': k, 'bias': b}

    P = {}
    for k in range(ntypes * ntypes):
        net, prev = {}, 1
        for li, w in enumerate(embed_widths):
            net['Dense_%d' % li] = dense(prev, w)
            prev = w
        P['embedding_net_%d' % k] = net
    P['Tbias'] = np.zeros(embed_widths[-1], f32)
    for t in range(ntypes):
        net, prev = {}, axis * embed_widths[-1]
        for li, w in enumerate(fit_widths):
            net['Dense_%d' % li] = dense(prev, w)
            if li > 0 and fit_widths[li] == fit_widths[li - 1]:
                net['dt%d' % li] = np.full(w, fit_dt, f32)
            prev = w
        net['Dense_%d' % len(fit_widths)] = dense(prev, 1, bias=False)
        P['fitting_net_%d' % t] = net
    params = {
        'type': 'energy', 'atomic_data_prefix': None,
        'embed_widths': list(embed_widths), 'embedMP_widths': None,
        'fit_widths': list(fit_widths), 'axis': axis,
        'Ebias': np.zeros(ntypes, f32), 'rcut': rcut, 'use_2nd': True,
        'use_mp': False, 'atomic': False, 'hybrid': False, 'nsel': None,
        'out_norm': 1.0, 'ntypes': ntypes, 'valid_types': np.arange(ntypes),
        'sr_mean': mean, 'sr_std': std, 'Nnbrs': nnbrs,
    }
    return params, {'params': P}


def init_wannier_model(seed=0, coord=None, box=None, type_idx=None, rcut=6.0, embed_widths=(25, 50, 100),
                       fit_widths=(100, 100, 100), axis=16, ntypes=2, nsel=(0,), fit_dt=0.1, out_norm=0.05,
                       kind='atomic'):
    """Seeded random-init atomic vector model (type 'atomic', a Wannier-centroid model as
    deepmd_jax/train.py:121-150 builds it): embedding nets for the centre types in nsel
    (net k = position in nsel * ntypes + neighbour type), one fitting net of hidden layers
    per nsel type with fit_widths[-1] == embed_widths[-1], outputs scaled by out_norm (the
    displacements stay a fraction of an Angstrom).  kind='atomic_t2' gives the symmetric-tensor
    (polarizability) head instead of the vector."""
    if fit_widths[-1] != embed_widths[-1]:
        raise ValueError('For atomic models, embed_widths[-1] must equal fit_widths[-1].')
    params, var = init_model(seed, coord, box, type_idx, rcut, embed_widths, fit_widths, axis, ntypes, fit_dt)
    P = var['params']
    out = {'Tbias': P['Tbias']}
    for ci, i in enumerate(nsel):
        for j in range(ntypes):
            out['embedding_net_%d' % (ci * ntypes + j)] = P['embedding_net_%d' % (i * ntypes + j)]
        net = dict(P['fitting_net_%d' % i])
        net.pop('Dense_%d' % len(fit_widths))
        out['fitting_net_%d' % ci] = net
    params.update({'type': kind, 'atomic': True, 'nsel': list(nsel), 'out_norm': float(out_norm),
                   'atomic_data_prefix': 'atomic_dipole' if kind == 'atomic' else 'atomic_polarizability'})
    return params, {'params': out}

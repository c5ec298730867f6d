"""DPLR oracle pins (CPU): PPPM energy (deepmd_jax/utils.py:174-222), the atomic Wannier-centroid
model (deepmd_jax/dpmodel.py:83-98) and the combined DPLR energy (deepmd_jax/md.py:404-415).

- PPPM converges to the direct Ewald reciprocal-space sum as the grid is refined (relative
  error 7e-4 at resolution 3 down to below 1e-5 at 12): a dropped or wrong factor of the
  influence function (TSC window, its alias sum, 2 pi / V, the Gaussian screening) breaks it.
- The centroid model is a vector: for an isolated cluster its outputs rotate with the cluster,
  and it is invariant under translation and equivariant under permutation of equal atoms.
- DPLR forces are the gradient of the DPLR energy (central differences).
"""
import numpy as np
import pytest
import torch

from oracle import dp_oracle as O


def test_pppm_converges_to_ewald():
    rng = np.random.default_rng(0)
    L = np.array([10., 11., 12.])
    x = rng.uniform(0, 1, (40, 3)) * L
    q = rng.standard_normal(40)
    q -= q.mean()
    ew = O.ewald_reciprocal(x, q, L, 0.4, 10)
    errs = []
    for res in (3, 5, 12):
        e = O.p3m_energy(torch.as_tensor(x), q, L, 0.4, O.p3m_grid(L, 0.4, res)).item()
        errs.append(abs(e - ew) / abs(ew))
    assert errs[0] < 2e-3 and errs[1] < 3e-4 and errs[2] < 1e-5, errs
    assert errs[0] > errs[1] > errs[2]
    # periodic: shifting every charge by a lattice vector leaves the energy unchanged
    M = O.p3m_grid(L, 0.4, 5)
    e0 = O.p3m_energy(torch.as_tensor(x), q, L, 0.4, M).item()
    e1 = O.p3m_energy(torch.as_tensor(x + L * np.array([1, -2, 3])), q, L, 0.4, M).item()
    assert abs(e1 - e0) < 1e-10 * abs(e0)


def _cluster(seed=0):
    from deepmd_jax import synthetic as syn
    coord, _, types = syn.water_box(8, seed=seed)
    coord = coord - coord.mean(0)
    return coord, np.full(3, 40.0), types


def test_wannier_model_is_a_vector():
    from deepmd_jax import synthetic as syn
    params, var = syn.init_wannier_model(3)
    coord, box, types = _cluster()
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    f = lambda x: O.dp_atomic(params, var, torch.as_tensor(x + 20.0), box, types, lattice=lat).detach().numpy()
    w0 = f(coord)
    assert w0.shape == ((types == 0).sum(), 3) and np.abs(w0).max() > 1e-4
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    np.testing.assert_allclose(f(coord @ Q.T), w0 @ Q.T, atol=1e-10)
    np.testing.assert_allclose(f(coord + np.array([0.3, -0.2, 0.1])), w0, atol=1e-10)
    # swapping two oxygens (with their hydrogens' order kept) swaps their rows
    o = np.where(types == 0)[0]
    perm = np.arange(len(types))
    perm[[o[0], o[1]]] = perm[[o[1], o[0]]]
    w1 = f(coord[perm])
    np.testing.assert_allclose(w1[[1, 0]], w0[[0, 1]], atol=1e-10)
    # zero hidden output -> zero displacement
    var0 = {'params': dict(var['params'])}
    net = {}
    for k, v in var0['params']['fitting_net_0'].items():
        net[k] = ({'kernel': np.zeros_like(v['kernel']), 'bias': np.zeros_like(v['bias'])}
                  if k.startswith('Dense_') else np.zeros_like(v))
    var0['params']['fitting_net_0'] = net
    assert np.abs(O.dp_atomic(params, var0, torch.as_tensor(coord + 20), box, types, lattice=lat).numpy()).max() == 0


def test_dplr_forces_are_the_energy_gradient(model0):
    from deepmd_jax import synthetic as syn
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    coord, box, types = syn.water_box(27, seed=9)
    M = O.p3m_grid(box, 0.4, 5)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    args = (params, var, wp, wv)
    q_atoms, q_wc = [6.0, 1.0], [-8.0]
    E, Elr, F, wc = O.dplr_energy_force(*args, coord, box, types, q_atoms, q_wc, 0.4, M, lattice=lat)
    assert wc.shape == ((types == 0).sum(), 3)
    assert abs(Elr) > 1e-3
    rng = np.random.default_rng(2)
    for i in rng.choice(len(types), 4, replace=False):
        for d in range(3):
            h = 1e-5
            xp, xm = coord.copy(), coord.copy()
            xp[i, d] += h
            xm[i, d] -= h
            ep = O.dplr_energy_force(*args, xp, box, types, q_atoms, q_wc, 0.4, M, lattice=lat)[0]
            em = O.dplr_energy_force(*args, xm, box, types, q_atoms, q_wc, 0.4, M, lattice=lat)[0]
            assert abs(-(ep - em) / (2 * h) - F[i, d]) < 1e-6 * max(1.0, abs(F[i, d])), (i, d)


def test_tensor_model_is_a_symmetric_tensor():
    """atomic_t2 head (deepmd_jax/dpmodel.py:90-99): symmetric 3x3 per selected atom that rotates as
    R A R^T with an isolated cluster."""
    from deepmd_jax import synthetic as syn
    params, var = syn.init_wannier_model(5, kind='atomic_t2')
    coord, box, types = _cluster(1)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    f = lambda x: O.dp_atomic(params, var, torch.as_tensor(x + 20.0), box, types, lattice=lat).detach().numpy()
    a0 = f(coord).reshape(-1, 3, 3)
    np.testing.assert_allclose(a0, a0.transpose(0, 2, 1), atol=1e-12)
    assert np.abs(a0).max() > 1e-4
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    a1 = f(coord @ Q.T).reshape(-1, 3, 3)
    np.testing.assert_allclose(a1, Q @ a0 @ Q.T, atol=1e-10)


def test_p3mlr_grid_size_and_host_helpers():
    """get_p3mlr_grid_size (truncation of box * beta * resolution), dplr_charges and
    periodic_replicate (deepmd_jax/utils.py:174-176,299-327,355-366) on closed cases."""
    from deepmd_jax.evaluator import replicate
    from deepmd_jax.utils import dplr_charges, get_p3mlr_grid_size, periodic_replicate
    assert get_p3mlr_grid_size(np.array([12.42, 9.99, 20.0]), 0.4, 5) == (24, 19, 40)
    qa, qw = dplr_charges([1, 0, 1, 0, 0], [6.0, 1.0], [-8.0], [0], 2)
    np.testing.assert_array_equal(qa, [1, 6, 1, 6, 6])
    np.testing.assert_array_equal(qw, [-8, -8, -8])
    rng = np.random.default_rng(0)
    c = rng.uniform(0, 5, (4, 3))
    t = np.array([0, 1, 1, 0])
    H = np.array([[5.0, 0, 0], [1.0, 6.0, 0], [0.5, -1.0, 7.0]])
    for box in (np.array([5.0, 6.0, 7.0]), H):
        out_c, out_b, out_t, out_f = periodic_replicate([2, 1, 3], c, box, t, force=np.ones((4, 3)))
        assert out_c.shape == (24, 3) and out_f.shape == (24, 3)
        np.testing.assert_array_equal(out_t, np.tile(t, 6))
        cell = np.diag(box) if box.shape == (3,) else box
        # every block of 4 atoms is the cell shifted by i0 col0 + i1 col1 + i2 col2, last direction outermost
        for i2 in range(3):
            for i0 in range(2):
                blk = out_c[(i2 * 2 + i0) * 4:(i2 * 2 + i0 + 1) * 4]
                np.testing.assert_allclose(blk, c + i0 * cell[:, 0] + i2 * cell[:, 2])
        np.testing.assert_allclose(out_b, np.array([2, 1, 3]) * box if box.shape == (3,) else np.diag([2, 1, 3]) @ H)
    # orthorhombic: the same set of images as the evaluator's supercell
    a = periodic_replicate([2, 2, 1], c, np.array([5.0, 6.0, 7.0]))[0]
    b = replicate(c, np.array([5.0, 6.0, 7.0]), t, np.array([2, 2, 1]))[0]
    np.testing.assert_allclose(np.sort(a, axis=0), np.sort(b, axis=0))

"""Pins of the CPU oracle (oracle/dp_oracle.py) by closed forms, invariances, an
independent library routine (scipy's periodic KD-tree), brute force, autograd
against central finite differences and golden fixtures."""
import os

import numpy as np
import pytest
import torch

from oracle import dp_oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_switch_closed_form():
    rc = 6.0
    r = torch.tensor([0.0, 1e-15, 1.0, 3.0, 5.999, 6.0, 7.0], dtype=torch.float64)
    s = O.sr(r, rc)
    t = r / rc
    expect = (1 - 3 * t ** 2 + 2 * t ** 3) / (r + 1e-15)
    assert s[0] == 0 and s[1] == 0          # r <= 1e-14 is excluded
    assert abs(s[3] - 0.5 / 3.0) < 1e-15    # t = 1/2: 1 - 3/4 + 1/4 = 1/2
    assert torch.allclose(s[2:5], expect[2:5])
    assert s[5] == 0 and s[6] == 0         # t >= 1 is cut


def test_tensor_3to6_layout():
    x = torch.tensor([[2.0], [3.0], [5.0]], dtype=torch.float64)  # components along axis 0
    y = O.tensor_3to6(x, axis=0, bias=1.0).reshape(-1)
    r2 = 2 ** 0.5
    # roll by one along the component axis: partner of x is z, of y is x, of z is y
    assert torch.allclose(y, torch.tensor([3.0, 8.0, 24.0, r2 * 10, r2 * 6, r2 * 15], dtype=torch.float64))


def test_shift_minimum_image():
    L = torch.tensor([10.0, 10.0, 10.0], dtype=torch.float64)
    d = torch.tensor([[4.9, -4.9, 6.0], [-6.0, 14.0, 0.0]], dtype=torch.float64)
    s = O.shift_ortho(d, L)
    assert torch.allclose(s, torch.tensor([[4.9, -4.9, -4.0], [4.0, 4.0, 0.0]], dtype=torch.float64))


def _small(nmol=32, seed=0):
    from deepmd_jax import synthetic as syn
    return syn.water_box(nmol, seed=seed)


def test_zero_weight_nets_give_tanh_bias():
    """All embedding weights 0 and biases b: h1 = tanh(b0), h2 = tanh(b1) + repeat(h1), ..."""
    b0, b1, b2 = np.full(2, 0.3), np.full(4, -0.2), np.full(8, 0.1)
    net = {'Dense_0': {'kernel': np.zeros((1, 2)), 'bias': b0},
           'Dense_1': {'kernel': np.zeros((2, 4)), 'bias': b1},
           'Dense_2': {'kernel': np.zeros((4, 8)), 'bias': b2}}
    s = torch.randn(1, 5, 1, dtype=torch.float64)
    red = torch.eye(5, dtype=torch.float64)[None]
    out = O.embedding_net(net, s, red)[0, 0]
    h1 = np.tanh(0.3)
    h2 = np.tanh(-0.2) + h1
    h3 = np.tanh(0.1) + h2
    assert torch.allclose(out, torch.full((8,), h3, dtype=torch.float64))


def test_descriptor_reduces_to_outer_product(model0):
    """With T3 = T6 = 0 the descriptor is G[a, w] = TW[w] TW[a] (dpmodel.py:75-78)."""
    params, var = model0
    A, W = params['axis'], params['embed_widths'][-1]
    T = torch.zeros(1, 10, W, dtype=torch.float64)
    T[0, 0] = torch.randn(W, dtype=torch.float64)
    TW = T[:, 0]
    T3, T6 = T[:, 1:4], T[:, 4:]
    G = TW[:, None] * TW[:, :A, None] + (T3[:, :, None] * T3[:, :, :A, None]).sum(1)
    G2 = O.tensor_3to6(T3[:, :, A:2 * A], axis=1) + T6[:, :, A:2 * A]
    G = G + (G2[..., None] * T6[:, :, None]).sum(1)
    assert torch.allclose(G[0], torch.outer(TW[0, :A], TW[0]))


def test_invariances(model0):
    params, var = model0
    coord, box, types = _small(64, seed=0)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    E0, F0, W0 = O.energy_force_virial(params, var, coord, box, types, lattice=lat)
    # translation
    E1, F1, _ = O.energy_force_virial(params, var, coord + np.array([1.3, -2.1, 0.7]), box, types, lattice=lat)
    assert abs(E1 - E0) < 1e-9 and torch.allclose(F1, F0, atol=1e-9)
    # permutation within a type (swap two H atoms and two O atoms)
    perm = np.arange(len(types))
    perm[[1, 2]] = perm[[2, 1]]
    perm[[0, 3]] = perm[[3, 0]]
    E2, F2, _ = O.energy_force_virial(params, var, coord[perm], box, types[perm], lattice=lat)
    assert abs(E2 - E0) < 1e-9 and torch.allclose(F2, F0[perm], atol=1e-9)
    # momentum conservation and virial symmetry
    assert torch.allclose(F0.sum(0), torch.zeros(3, dtype=torch.float64), atol=1e-10)
    assert torch.allclose(W0, W0.T, atol=1e-9)


def test_rotation_invariance_isolated_cluster(model0):
    """A molecule cluster in a large box: E invariant, forces co-rotate."""
    params, var = model0
    coord, box, types = _small(8, seed=4)
    c = coord - coord.mean(0)
    big = np.full(3, 40.0)
    q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))
    lat = O.lattice_candidate(np.diag(big), params['rcut'])
    E0, F0, _ = O.energy_force_virial(params, var, c + 20, big, types, lattice=lat)
    E1, F1, _ = O.energy_force_virial(params, var, c @ q.T + 20, big, types, lattice=lat)
    assert abs(E1 - E0) < 1e-9
    assert torch.allclose(F1, F0 @ torch.as_tensor(q.T), atol=1e-9)


def test_forces_match_central_differences(model0):
    params, var = model0
    coord, box, types = _small(32, seed=6)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    _, F, _ = O.energy_force_virial(params, var, coord, box, types, lattice=lat)
    h = 1e-4
    rng = np.random.default_rng(0)
    for a in rng.choice(len(types), 6, replace=False):
        for d in range(3):
            cp, cm = coord.copy(), coord.copy()
            cp[a, d] += h
            cm[a, d] -= h
            ep = O.dp_energy(params, var, torch.as_tensor(cp), box, types, lattice=lat).item()
            em = O.dp_energy(params, var, torch.as_tensor(cm), box, types, lattice=lat).item()
            assert abs(-(ep - em) / (2 * h) - F[a, d].item()) < 1e-6


def test_virial_trace_is_strain_derivative(model0):
    """tr(virial) = -dU/d eps under x -> (1+eps) x, box -> (1+eps) box (md_utils.py:139-159)."""
    params, var = model0
    coord, box, types = _small(64, seed=7)
    _, _, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    h = 1e-5
    ep = O.dp_energy(params, var, torch.as_tensor(coord * (1 + h)), box * (1 + h), types,
                     lattice=O.lattice_candidate(np.diag(box * (1 + h)), params['rcut'])).item()
    em = O.dp_energy(params, var, torch.as_tensor(coord * (1 - h)), box * (1 - h), types,
                     lattice=O.lattice_candidate(np.diag(box * (1 - h)), params['rcut'])).item()
    assert abs(np.trace(W.numpy()) + (ep - em) / (2 * h)) < 1e-5 * abs(np.trace(W.numpy())) + 1e-7


def test_lattice_path_equals_nlist_path(model0):
    params, var = model0
    coord, box, types = _small(64, seed=8)
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 0.2)
    E0, F0, W0 = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E1, F1, W1 = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
    assert abs(E1 - E0) < 1e-10 and torch.allclose(F0, F1, atol=1e-10) and torch.allclose(W0, W1, atol=1e-9)


def test_small_box_lattice_images_vs_supercell(model0):
    """Box below 2*rcut: the lattice-image path equals the supercell energy / 8."""
    params, var = model0
    coord, box, types = _small(32, seed=9)
    assert box[0] < 2 * params['rcut']
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    assert lat[1] >= 2
    E0, F0, _ = O.energy_force_virial(params, var, coord, box, types, lattice=lat)
    sh = np.stack(np.meshgrid(*[np.arange(2)] * 3, indexing='ij'), -1).reshape(-1, 3)
    sc = (coord[None] + (sh * box)[:, None]).reshape(-1, 3)
    st = np.tile(types, 8)
    E1, F1, _ = O.energy_force_virial(params, var, sc, box * 2, st, lattice='auto')
    assert abs(E1 / 8 - E0) < 1e-9 * abs(E0)
    assert torch.allclose(F1[:len(types)], F0, atol=1e-9)


def test_nlist_against_scipy_kdtree():
    """Neighbour sets within rc+skin equal scipy's periodic cKDTree (away from the
    float32 cutoff boundary, where the oracle follows jax_md's float32 test)."""
    from scipy.spatial import cKDTree
    coord, box, types = _small(216, seed=2)
    rc = 7.0
    c, rows, border = O.neighbor_sets(coord, box, types, rc)
    tree = cKDTree(c.astype(np.float64), boxsize=np.asarray(box, np.float32).astype(np.float64) + 1e-9)
    for i in range(0, len(c), 3):
        inner = set(tree.query_ball_point(c[i].astype(np.float64), rc * (1 - 1e-6))) - {i}
        outer = set(tree.query_ball_point(c[i].astype(np.float64), rc * (1 + 1e-6))) - {i}
        assert inner <= set(rows[i].tolist()) <= outer
        assert len(outer) - len(inner) <= 2


def test_knbr_rule():
    assert O.knbr_from_counts([56, 112, 0]) == [68, 135, 1]
    assert O.knbr_from_counts([10], 1.3) == [1 + int(10 * 1.3 + 10 * (1.3 - 1.2))]


def test_typed_nlist_format():
    coord, box, types = _small(64, seed=3)
    nb, ov, knbr = O.typed_nlist(coord, box, types, 6.2)
    N = len(types)
    assert not ov
    for i in range(2):
        for j in range(2):
            a = nb[i][j]
            assert a.shape[1] == knbr[j] - 1
            valid = a < N
            # ascending valid entries then padding N
            for row in a:
                v = row[row < N]
                assert np.all(np.diff(v) > 0)
                assert np.all(row[len(v):] == N)


def test_golden_fixture(model0):
    """Regression pin: oracle E/F/virial on the 64-molecule box (tests/golden)."""
    g = np.load(os.path.join(GOLDEN, 'water64_seed0.npz'))
    params, var = model0
    coord, box, types = _small(64, seed=0)
    np.testing.assert_array_equal(coord, g['coord'])
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    assert abs(E.item() - g['energy']) < 1e-10
    np.testing.assert_allclose(F.numpy(), g['force'], atol=1e-10)
    np.testing.assert_allclose(W.numpy(), g['virial'], atol=1e-9)
    np.testing.assert_allclose(params['sr_mean'], g['sr_mean'], rtol=1e-14)
    np.testing.assert_allclose(params['sr_std'], g['sr_std'], rtol=1e-14)


def test_ring_modes_closed_form():
    """Normal-mode frequencies 2 P kT/hbar sin(pi k / P) (as multiset) and C orthogonal."""
    P, kT = 8, 0.025
    f, C = O.ring_modes(P, kT)
    hbar = 6.582119569e-1
    expect = np.sort(2 * P * kT / hbar * np.sin(np.pi * np.arange(P) / P))
    np.testing.assert_allclose(np.sort(f), expect, atol=1e-12)
    np.testing.assert_allclose(C.T @ C, np.eye(P), atol=1e-12)
    assert f[0] == 0


def test_free_ring_matches_harmonic_solution():
    """With no physical force the free-ring step is the exact harmonic ring solution:
    energy 0.5 p^2/m + 0.5 m w_P^2 sum (x_b - x_{b-1})^2 is conserved exactly and
    two half steps equal one full step."""
    P, kT = 6, 0.03
    f, C = O.ring_modes(P, kT)
    rng = np.random.default_rng(0)
    m = np.array([1.5, 16.0])
    x = rng.standard_normal((P, 2, 3)) * 0.1 + 5.0
    p = rng.standard_normal((P, 2, 3)) * 0.05
    box = np.full(3, 50.0)
    wP = P * kT / 6.582119569e-1

    def H(x, p):
        d = x - np.roll(x, 1, axis=0)
        return 0.5 * (p ** 2 / m[None, :, None]).sum() + 0.5 * wP ** 2 * (m[None, :, None] * d ** 2).sum()
    x1, p1 = O.free_ring(x, p, m, C, f, 3.0, box)
    assert abs(H(x1, p1) - H(x, p)) < 1e-10 * H(x, p)
    xa, pa = O.free_ring(*O.free_ring(x, p, m, C, f, 1.5, box), m, C, f, 1.5, box)
    np.testing.assert_allclose(xa, x1, atol=1e-12)
    np.testing.assert_allclose(pa, p1, atol=1e-12)


def test_velocity_verlet_harmonic_oscillator():
    """VV on a harmonic oscillator: energy error stays bounded (1%) over 1000 steps and
    integrating back with -dt returns to the start."""
    k, m, dt = 2.0, np.array([3.0]), 0.05
    ff = lambda x: -k * x
    x, p = np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 0.5, 0.0]])
    F = ff(x)
    xs, ps = x.copy(), p.copy()
    for _ in range(1000):
        x, p, F = O.vv_step(x, p, m, ff, dt, F)
    E0 = 0.5 * (ps ** 2).sum() / m[0] + 0.5 * k * (xs ** 2).sum()
    E1 = 0.5 * (p ** 2).sum() / m[0] + 0.5 * k * (x ** 2).sum()
    assert abs(E1 - E0) < 1e-2 * E0
    for _ in range(1000):
        x, p, F = O.vv_step(x, p, m, ff, -dt, F)
    np.testing.assert_allclose(x, xs, atol=1e-9)
    np.testing.assert_allclose(p, ps, atol=1e-9)


def test_local_fd_matches_autograd(model0):
    """The local-cluster central differences reproduce periodic autograd forces (box
    larger than 4 rcut + 1, so a 2 rcut + 0.5 cluster around an atom has no image ambiguity)."""
    params, var = model0
    coord, box, types = _small(1000, seed=13)
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'])
    _, F, _ = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
    atoms = [0, 5, 1001]
    fd = O.local_fd_forces(params, var, coord, box, types, atoms)
    np.testing.assert_allclose(fd, F.numpy()[atoms], atol=2e-6)

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
    tb = np.random.default_rng(2).standard_normal(W)
    G = O.descriptor_from_T(T, tb, A)
    TW = T[0, 0] + torch.as_tensor(tb)
    assert torch.allclose(G[0], torch.outer(TW[:A], TW))


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


# ---------------------------------------------------------------------------------
# Hand-computed pins of the forward conventions of deepmd_jax/utils.py:114-124 (repeat
# residual), :161-163 (fitting dt residual where widths are equal), dpmodel.py:38-39
# (normalisation by the centre type), :50-51 (net index i*Y + j, centre type outer) and
# :74-78 (second-order term on T[:, :, A:2A]).  Every expected value below is written
# out from the formulas, never by calling the oracle's own helpers.

def test_embedding_repeat_residual_hand_values():
    """1 -> 2 -> 4 -> 8 net with distinct parameters: the residual repeats each element in
    place ([a, a, b, b, ...], jnp.repeat), not a tile ([a, b, a, b, ...])."""
    k0, b0 = np.array([[0.7, -1.3]]), np.array([0.2, -0.4])
    k1 = np.array([[0.5, -0.2, 0.1, 0.9], [-0.6, 0.3, 0.8, -0.1]])
    b1 = np.array([0.05, -0.3, 0.6, -0.9])
    k2 = np.arange(32, dtype=np.float64).reshape(4, 8) / 40.0 - 0.4
    b2 = np.linspace(-0.5, 0.5, 8)
    net = {'Dense_0': {'kernel': k0, 'bias': b0}, 'Dense_1': {'kernel': k1, 'bias': b1},
           'Dense_2': {'kernel': k2, 'bias': b2}}
    s = 0.37
    h1 = np.tanh([0.7 * s + 0.2, -1.3 * s - 0.4])
    z1 = [0.5 * h1[0] - 0.6 * h1[1] + 0.05, -0.2 * h1[0] + 0.3 * h1[1] - 0.3,
          0.1 * h1[0] + 0.8 * h1[1] + 0.6, 0.9 * h1[0] - 0.1 * h1[1] - 0.9]
    h2 = np.tanh(z1) + np.array([h1[0], h1[0], h1[1], h1[1]])
    z2 = [sum(h2[r] * k2[r, c] for r in range(4)) + b2[c] for c in range(8)]
    h3 = np.tanh(z2) + np.array([h2[0], h2[0], h2[1], h2[1], h2[2], h2[2], h2[3], h2[3]])
    out = O.embedding_net(net, torch.tensor([[[s]]], dtype=torch.float64),
                          torch.ones(1, 1, 1, dtype=torch.float64))[0, 0].numpy()
    np.testing.assert_allclose(out, h3, rtol=0, atol=1e-14)
    # zero kernels: the layers reduce to tanh(bias) plus the repeated parent
    zn = {k: {'kernel': np.zeros_like(v['kernel']), 'bias': v['bias']} for k, v in net.items()}
    g1 = np.tanh(b0)
    g2 = np.tanh(b1) + np.array([g1[0], g1[0], g1[1], g1[1]])
    g3 = np.tanh(b2) + np.array([g2[0], g2[0], g2[1], g2[1], g2[2], g2[2], g2[3], g2[3]])
    out = O.embedding_net(zn, torch.tensor([[[s]]], dtype=torch.float64),
                          torch.ones(1, 1, 1, dtype=torch.float64))[0, 0].numpy()
    np.testing.assert_allclose(out, g3, rtol=0, atol=1e-14)


def test_fitting_dt_residual_hand_values():
    """Equal-width hidden layers add x_prev after scaling by dt; a width change has no
    residual (utils.py:158-163).  Zero hidden kernels make every layer tanh(b)."""
    b0, b1, b2 = np.array([0.3, -0.8, 1.1]), np.array([-0.5, 0.2, 0.9]), np.array([0.7, -1.2, 0.05])
    dt1, dt2 = np.array([0.1, 0.25, -0.4]), np.array([0.6, -0.05, 0.3])
    k3, b3 = np.array([[1.5], [-0.7], [0.25]]), np.array([0.125])
    net = {'Dense_0': {'kernel': np.zeros((5, 3)), 'bias': b0},
           'Dense_1': {'kernel': np.zeros((3, 3)), 'bias': b1}, 'dt1': dt1,
           'Dense_2': {'kernel': np.zeros((3, 3)), 'bias': b2}, 'dt2': dt2,
           'Dense_3': {'kernel': k3, 'bias': b3}}
    x = torch.tensor([[0.4, -1.0, 2.0, 0.0, 3.0]], dtype=torch.float64)
    y1 = np.tanh(b0)
    y2 = np.tanh(b1) * dt1 + y1
    y3 = np.tanh(b2) * dt2 + y2
    e = 1.5 * y3[0] - 0.7 * y3[1] + 0.25 * y3[2] + 0.125
    assert abs(O.fitting_net(net, x)[0, 0].item() - e) < 1e-14
    # 5 -> 3 -> 2 -> 1: the width change has no residual and no dt
    net2 = {'Dense_0': {'kernel': np.zeros((5, 3)), 'bias': b0},
            'Dense_1': {'kernel': np.array([[1.0, 0.0], [0.0, 2.0], [0.5, -1.0]]), 'bias': np.array([0.1, -0.2])},
            'Dense_2': {'kernel': np.array([[2.0], [-3.0]]), 'bias': np.array([0.0])}}
    z = np.tanh([y1[0] + 0.5 * y1[2] + 0.1, 2.0 * y1[1] - y1[2] - 0.2])
    assert abs(O.fitting_net(net2, x)[0, 0].item() - (2.0 * z[0] - 3.0 * z[1])) < 1e-14


def test_descriptor_hand_entries():
    """G of a tiny T (W = 4, A = 2) with every block nonzero, entries written out:
    G[a,w] = TW_w TW_a + sum_c T3[c,w] T3[c,a] + sum_k (3to6(T3[:,A+a]) + T6[k,A+a]) T6[k,w],
    3to6(v) = (v0^2, v1^2, v2^2, sqrt2 v0 v2, sqrt2 v1 v0, sqrt2 v2 v1)."""
    T = torch.tensor([[[1, 2, 0, -1],                  # T0
                       [0, 1, 1, 2], [1, 0, -1, 1], [2, 1, 0, 0],   # T3 rows x, y, z
                       [1, 0, 0, 1], [0, 1, 1, 0], [0, 0, 1, 1],    # T6 rows 0..2
                       [1, 1, 0, 0], [0, 0, 0, 1], [1, 0, 1, 0]]],  # T6 rows 3..5
                     dtype=torch.float64)
    Tbias = np.array([0.5, 0.0, -1.0, 0.0])
    G = O.descriptor_from_T(T, Tbias, 2)[0].numpy()
    r2 = 2 ** 0.5
    # TW = (1.5, 2, -1, -1)
    # a = 0: T3[:, 0] = (0, 1, 2); T3[:, A+0] = T3[:, 2] = (1, -1, 0);
    #        3to6 = (1, 1, 0, 0, -r2, 0); T6[:, 2] = (0, 1, 1, 0, 0, 1) -> v = (1, 2, 1, 0, -r2, 1)
    # a = 1: T3[:, 1] = (1, 0, 1); T3[:, 3] = (2, 1, 0); 3to6 = (4, 1, 0, 0, 2 r2, 0);
    #        T6[:, 3] = (1, 0, 1, 0, 1, 0) -> v = (5, 1, 1, 0, 1 + 2 r2, 0)
    # T6 columns w: (1,0,0,1,0,1), (0,1,0,1,0,0), (0,1,1,0,0,1), (1,0,1,0,1,0)
    expect = np.array([
        [1.5 * 1.5 + 5 + (1 + 0 + 0 + 0 + 0 + 1),            # w=0: T3 dot (0,1,2)(0,1,2)=5
         2 * 1.5 + (1 * 0 + 0 * 1 + 1 * 2) + (0 + 2 + 0 + 0 + 0 + 0),
         -1 * 1.5 + (1 * 0 + (-1) * 1 + 0 * 2) + (0 + 2 + 1 + 0 + 0 + 1),
         -1 * 1.5 + (2 * 0 + 1 * 1 + 0 * 2) + (1 + 0 + 1 + 0 - r2 + 0)],
        [1.5 * 2 + (0 * 1 + 1 * 0 + 2 * 1) + (5 + 0 + 0 + 0 + 0 + 0),
         2 * 2 + (1 + 0 + 1) + (0 + 1 + 0 + 0 + 0 + 0),
         -1 * 2 + (1 * 1 + (-1) * 0 + 0 * 1) + (0 + 1 + 1 + 0 + 0 + 0),
         -1 * 2 + (2 * 1 + 1 * 0 + 0 * 1) + (5 + 0 + 1 + 0 + (1 + 2 * r2) + 0)]])
    np.testing.assert_allclose(G, expect, rtol=0, atol=1e-13)


def _pair_model(Y=2, W=8, A=2):
    """A tiny two-type model whose only nonzero embedding nets are (O centre, H neighbour)
    and (H centre, O neighbour), distinct per-type statistics, Tbias, Nnbrs and fitting nets."""
    rng = np.random.default_rng(11)

    def dense(i, o, scale=1.0):
        return {'kernel': rng.standard_normal((i, o)) * scale / np.sqrt(i), 'bias': rng.standard_normal(o)}

    def zero_net():
        return {'Dense_0': {'kernel': np.zeros((1, 2)), 'bias': np.zeros(2)},
                'Dense_1': {'kernel': np.zeros((2, 4)), 'bias': np.zeros(4)},
                'Dense_2': {'kernel': np.zeros((4, 8)), 'bias': np.zeros(8)}}
    P = {'embedding_net_%d' % k: zero_net() for k in range(Y * Y)}
    for k in (1, 2):
        P['embedding_net_%d' % k] = {'Dense_0': dense(1, 2), 'Dense_1': dense(2, 4), 'Dense_2': dense(4, 8)}
    P['Tbias'] = np.array([0.3, -0.2, 0.1, 0.0, 0.25, -0.15, 0.05, 0.4])
    for t in range(Y):
        P['fitting_net_%d' % t] = {'Dense_0': dense(A * W, 3, 4.0), 'Dense_1': dense(3, 1)}
    params = {'type': 'energy', 'embed_widths': [2, 4, 8], 'fit_widths': [3], 'axis': A, 'rcut': 6.0,
              'use_2nd': True, 'atomic': False, 'ntypes': Y, 'out_norm': 1.7, 'Ebias': np.array([0.5, -0.25]),
              'sr_mean': np.array([0.11, 0.37]), 'sr_std': np.array([0.53, 1.9]), 'Nnbrs': 3.7}
    return params, {'params': P}


def _net_hand(net, s):
    """1 -> 2 -> 4 -> 8 embedding net on a scalar, residual parents written out."""
    d0, d1, d2 = net['Dense_0'], net['Dense_1'], net['Dense_2']
    h1 = np.tanh(s * d0['kernel'][0] + d0['bias'])
    h2 = np.tanh(h1 @ d1['kernel'] + d1['bias']) + h1[[0, 0, 1, 1]]
    return np.tanh(h2 @ d2['kernel'] + d2['bias']) + h2[[0, 0, 1, 1, 2, 2, 3, 3]]


def test_single_pair_energy_closed_form():
    """One O and one H, 2.3 A apart along a general direction in a 30 A box.  Centre i of
    type t sees one neighbour: s = sr(r), s_hat = s / std_t, s_til = (s - mean_t) / std_t
    (the centre's statistics), h = net_{t*Y + other}(s_til), T = R_X h^T / Nnbrs.  With
    c = 1/Nnbrs and R0^2 + |R3|^2 = 4 s_hat^2, sum_k 3to6(R3)_k R6_k = 6 s_hat^3,
    sum_k R6_k^2 = 6 s_hat^2 for any unit direction:
      G[a,w] = (c s_hat h_w + tb_w)(c s_hat h_a + tb_a) + 3 c^2 s_hat^2 h_a h_w
               + 6 c^3 s_hat^3 h_{A+a}^2 h_w + 6 c^2 s_hat^2 h_{A+a} h_w,
    and E = out_norm * sum_t (fit_t(G_t) + Ebias_t)."""
    params, var = _pair_model()
    P = var['params']
    A, Y, rc = params['axis'], params['ntypes'], params['rcut']
    u = np.array([0.48, -0.6, 0.64])
    u /= np.linalg.norm(u)
    r = 2.3
    coord = np.array([[12.0, 14.0, 15.0], [12.0, 14.0, 15.0] + r * u])
    types = np.array([0, 1])
    box = np.full(3, 30.0)
    t_ = r / rc
    s = (1 - 3 * t_ ** 2 + 2 * t_ ** 3) / (r + 1e-15)
    c = 1.0 / params['Nnbrs']
    tb = P['Tbias']
    E = 0.0
    for t in range(Y):
        sh = s / params['sr_std'][t]
        st = (s - params['sr_mean'][t]) / params['sr_std'][t]
        h = _net_hand(P['embedding_net_%d' % (t * Y + (1 - t))], st)
        G = np.empty((A, 8))
        for a in range(A):
            for w in range(8):
                G[a, w] = ((c * sh * h[w] + tb[w]) * (c * sh * h[a] + tb[a]) + 3 * c * c * sh * sh * h[a] * h[w]
                           + 6 * c ** 3 * sh ** 3 * h[A + a] ** 2 * h[w] + 6 * c * c * sh * sh * h[A + a] * h[w])
        f = P['fitting_net_%d' % t]
        y = np.tanh(G.reshape(-1) @ f['Dense_0']['kernel'] + f['Dense_0']['bias'])
        E += (y @ f['Dense_1']['kernel'])[0] + f['Dense_1']['bias'][0] + params['Ebias'][t]
    E *= params['out_norm']
    lat = O.lattice_candidate(np.diag(box), rc)
    for order in ([0, 1], [1, 0]):          # atom order in the frame does not matter
        E0 = O.dp_energy(params, var, torch.as_tensor(coord[order]), box, types[order], lattice=lat).item()
        assert abs(E0 - E) < 1e-12 * max(1.0, abs(E)), (E0, E)


@pytest.mark.parametrize('rc_skin,factor', [(7.0, 1.0), (6.3, 1.0), (7.0, 1.004)])
def test_nlist_fast_path_equals_brute_force(rc_skin, factor):
    """The KD-tree candidate path of neighbor_sets gives the brute force's FP32 sets, and
    typed_nlist's rows equal a per-row restatement of get_nm (md_utils.py:431-459)."""
    coord, box, types = _small(512, seed=5)
    nbox = box * factor
    x = coord * factor
    ref_box = None if factor == 1.0 else box
    _, r1, _ = O.neighbor_sets(x, nbox, types, rc_skin, reference_box=ref_box, fast=False)
    _, r2, _ = O.neighbor_sets(x, nbox, types, rc_skin, reference_box=ref_box, fast=True)
    assert all(np.array_equal(a, b) for a, b in zip(r1, r2))
    nb, _, knbr = O.typed_nlist(x, nbox, types, rc_skin, reference_box=ref_box)
    N = len(types)
    off = np.concatenate([[0], np.cumsum(np.bincount(types))])
    for i in range(2):
        for j in range(2):
            for a in range(off[i], off[i + 1]):
                r = r1[a][(r1[a] >= off[j]) & (r1[a] < off[j + 1])] - off[j]
                row = np.full(knbr[j] - 1, N)
                row[:min(len(r), knbr[j] - 1)] = r[:knbr[j] - 1]
                np.testing.assert_array_equal(nb[i][j][a - off[i]], row)


def test_canonicalize_wraps_in_fp64_first():
    """An unwrapped coordinate canonicalises to the FP32 value of its wrapped FP64 position,
    as the reference's FP32 positions (kept wrapped every step) do."""
    box = np.full(3, 24.83)
    x = np.array([[3.123456789, 12.5, 0.001]])
    for k in (-3, -1, 1, 4):
        u = x + k * box
        np.testing.assert_array_equal(O.canonicalize_f32(u, box), O.canonicalize_f32(x, box))

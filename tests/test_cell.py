"""General (triclinic) cells: oracle pins and host logic (CPU).

The reference evaluates (3,3) boxes with the fractional minimum image (shift,
deepmd_jax/utils.py:20-27) on its dense lattice-image path (deepmd_jax/utils.py:66-78,
lattice candidates of deepmd_jax/data.py:347-364).  The oracle's general-cell path is pinned
by properties a plausible mistake breaks: a rotated orthorhombic box is a general cell whose
energy, forces and virial are the rotated ones; a cell written with other lattice vectors of
the same lattice gives the same result; the neighbour-list form equals the lattice form when
every width exceeds 2 rcut; forces and the full virial match central differences of the
energy (the strain moves atoms and lattice vectors alike); a k1 x k2 x k3 supercell holds
prod(k) times the energy of the cell.
"""
import numpy as np
import pytest
import torch

from oracle import dp_oracle as O


def skewed(coord, box, tilt=(0.22, -0.17, 0.13)):
    """Fractional coordinates of an orthorhombic frame mapped into the cell with rows
    (Lx,0,0), (t0 Ly,Ly,0), (t1 Lz,t2 Lz,Lz)."""
    L = np.asarray(box, np.float64)
    H = np.array([[L[0], 0, 0], [tilt[0] * L[1], L[1], 0], [tilt[1] * L[2], tilt[2] * L[2], L[2]]])
    return (np.asarray(coord) / L) @ H, H


def _rot():
    from scipy.spatial.transform import Rotation
    return Rotation.from_euler('xyz', [0.3, -0.5, 1.1]).as_matrix()


def test_rotated_box_is_a_general_cell(model0):
    from deepmd_jax import synthetic as syn
    params, var = model0
    coord, box, types = syn.water_box(32, seed=3)
    E0, F0, W0 = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    R = _rot()
    E1, F1, W1 = O.energy_force_virial(params, var, coord @ R.T, np.diag(box) @ R.T, types, lattice='auto')
    assert abs(E1.item() - E0.item()) < 1e-10
    np.testing.assert_allclose(F1.numpy(), F0.numpy() @ R.T, rtol=0, atol=1e-10)
    np.testing.assert_allclose(W1.numpy(), R @ W0.numpy() @ R.T, rtol=0, atol=1e-10)


def test_other_lattice_vectors_of_the_same_lattice(model0):
    from deepmd_jax import synthetic as syn
    params, var = model0
    coord, box, types = syn.water_box(32, seed=4)
    L = box[0]
    E0, F0, W0 = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    for H in (np.array([[L, 0, 0], [L, L, 0], [0, 0, L]]), np.array([[L, 0, 0], [0, L, 0], [-L, 2 * L, L]])):
        E1, F1, W1 = O.energy_force_virial(params, var, coord, H, types, lattice='auto')
        assert abs(E1.item() - E0.item()) < 1e-10
        np.testing.assert_allclose(F1.numpy(), F0.numpy(), rtol=0, atol=1e-10)
        np.testing.assert_allclose(W1.numpy(), W0.numpy(), rtol=0, atol=1e-10)


def test_nlist_form_equals_lattice_form_in_a_wide_cell(model0):
    from deepmd_jax import synthetic as syn
    params, var = model0
    c0, box, types = syn.water_box(128, seed=5)
    coord, H = skewed(c0, box)
    from deepmd_jax.evaluator import cell_widths
    assert cell_widths(H).min() > 2 * params['rcut']
    nb, _, _ = O.typed_nlist(coord, H, types, params['rcut'])
    E0, F0, W0 = O.energy_force_virial(params, var, coord, H, types, nbrs_nm=nb)
    E1, F1, W1 = O.energy_force_virial(params, var, coord, H, types, lattice='auto')
    assert abs(E1.item() - E0.item()) < 1e-9
    np.testing.assert_allclose(F1.numpy(), F0.numpy(), rtol=0, atol=1e-10)
    np.testing.assert_allclose(W1.numpy(), W0.numpy(), rtol=0, atol=1e-9)


def test_forces_and_full_virial_against_finite_differences(model0):
    from deepmd_jax import synthetic as syn
    params, var = model0
    c0, box, types = syn.water_box(16, seed=6)
    coord, H = skewed(c0, box)
    lat = O.lattice_candidate(H, params['rcut'])

    def energy(x, cell):
        return O.energy_force_virial(params, var, x, cell, types, lattice=O.lattice_candidate(cell, params['rcut']))[0].item()
    E, F, W = O.energy_force_virial(params, var, coord, H, types, lattice=lat)
    h = 1e-5
    for a in (0, 7, 20):
        for d in range(3):
            xp, xm = coord.copy(), coord.copy()
            xp[a, d] += h
            xm[a, d] -= h
            fd = -(energy(xp, H) - energy(xm, H)) / (2 * h)
            assert abs(fd - F.numpy()[a, d]) < 1e-6 * max(1.0, np.abs(F.numpy()).max()), (a, d, fd, F.numpy()[a, d])
    # row vectors: x -> x (I + e)^T, lattice vectors alike; W = -dE/de
    for a in range(3):
        for b in range(3):
            e = np.zeros((3, 3))
            e[a, b] = h
            ep = energy(coord @ (np.eye(3) + e).T, H @ (np.eye(3) + e).T)
            em = energy(coord @ (np.eye(3) - e).T, H @ (np.eye(3) - e).T)
            fd = -(ep - em) / (2 * h)
            assert abs(fd - W.numpy()[a, b]) < 1e-6 * max(1.0, np.abs(W.numpy()).max()), (a, b, fd, W.numpy()[a, b])


def test_supercell_of_a_general_cell(model0):
    """replicate() of a narrow general cell: the supercell's energy is prod(k) times the
    cell's (both on the lattice form), the first image carries the cell's forces."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.evaluator import cell_widths, replicate, supercell_factors
    params, var = model0
    c0, box, types = syn.water_box(8, seed=7)
    coord, H = skewed(c0, box)
    k = supercell_factors(H, params['rcut'])
    assert (k > 1).all()
    sc, sH, st = replicate(coord, H, types, k)
    assert (cell_widths(sH) >= 2 * params['rcut']).all()
    np.testing.assert_allclose(sH, k[:, None] * H)
    E0, F0, W0 = O.energy_force_virial(params, var, coord, H, types, lattice='auto')
    nb, _, _ = O.typed_nlist(sc, sH, st, params['rcut'])
    E1, F1, W1 = O.energy_force_virial(params, var, sc, sH, st, nbrs_nm=nb)
    n = int(np.prod(k))
    assert abs(E1.item() / n - E0.item()) < 1e-9
    np.testing.assert_allclose(F1.numpy()[:len(types)], F0.numpy(), rtol=0, atol=1e-9)
    np.testing.assert_allclose(W1.numpy() / n, W0.numpy(), rtol=0, atol=1e-9)


def test_cell_widths_closed_form():
    from deepmd_jax.evaluator import canonical_box, cell_volume, cell_widths
    rng = np.random.default_rng(0)
    H = np.diag([10.0, 12.0, 9.0]) + rng.uniform(-3, 3, (3, 3)) * (1 - np.eye(3))
    V = abs(np.linalg.det(H))
    a, b, c = H
    want = [V / np.linalg.norm(np.cross(b, c)), V / np.linalg.norm(np.cross(c, a)), V / np.linalg.norm(np.cross(a, b))]
    np.testing.assert_allclose(cell_widths(H), want, rtol=1e-12)
    assert cell_volume(H) == pytest.approx(V)
    np.testing.assert_array_equal(cell_widths(np.array([3.0, 4.0, 5.0])), [3, 4, 5])
    assert canonical_box(np.diag([3.0, 4.0, 5.0])).shape == (3,)
    assert canonical_box(H).shape == (3, 3)


def test_unwrap_beads_general_cell():
    rng = np.random.default_rng(1)
    H = np.array([[9.0, 0, 0], [2.5, 8.0, 0], [-1.5, 3.0, 10.0]])
    x0 = rng.uniform(0, 5, (4, 6, 3))
    x0 = x0[0:1] + 0.05 * (x0 - x0[0:1])  # beads close to bead 0
    n = rng.integers(-2, 3, (4, 6, 3))
    n[0] = 0
    x = x0 + n @ H
    np.testing.assert_allclose(O.unwrap_beads(x, H), x0, atol=1e-12)
    from deepmd_jax.pimd import unwrap_across_beads
    np.testing.assert_allclose(unwrap_across_beads(x, H), x0, atol=1e-12)


def test_evaluate_box_shapes():
    from deepmd_jax.train import _boxes
    b = _boxes(np.array([[1.0, 2.0, 3.0]]), 1)
    np.testing.assert_array_equal(b[0], np.diag([1.0, 2.0, 3.0]))
    H = np.array([[9.0, 0, 0], [2.5, 8.0, 0], [-1.5, 3.0, 10.0]])
    np.testing.assert_array_equal(_boxes(H.reshape(1, 9), 1)[0], H)
    np.testing.assert_array_equal(_boxes(H[None], 1)[0], H)
    with pytest.raises(ValueError):
        _boxes(np.zeros((1, 4)), 1)

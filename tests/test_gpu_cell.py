"""General (triclinic) cells on the device against the FP64 oracle.

Cells with rows (Lx,0,0), (0.22 Ly,Ly,0), (-0.17 Lz,0.13 Lz,Lz) hold water frames mapped
through fractional coordinates.  Tolerances are those of the orthorhombic parity tests:
energy 1e-5 eV/atom and 1e-5 relative, forces 1e-4 * max(1, max |F|), virial 1e-4 * max |W|;
trajectories within 1e-5 (A, A/fs) of the oracle's.  Neighbour sets equal the oracle's FP64
sets except for pairs within 1e-4 A of the cutoff (FP32 fractional coordinates).
"""
import contextlib
import io

import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def skewed(coord, box, tilt=(0.22, -0.17, 0.13)):
    L = np.asarray(box, np.float64)
    H = np.array([[L[0], 0, 0], [tilt[0] * L[1], L[1], 0], [tilt[1] * L[2], tilt[2] * L[2], L[2]]])
    return (np.asarray(coord) / L) @ H, H


def _system(dm, coord, H, types, rc_skin, nrep=1, m=None):
    from deepmd_jax import _lib
    s = _lib.DeviceSystem(dm, types, nrep)
    s.set_box(H)
    s.set_masses(np.ones(len(types)) if m is None else m)
    s.set_positions(np.broadcast_to(coord, (nrep,) + coord.shape[-2:]) if coord.ndim == 2 else coord)
    s.nlist_allocate(rc_skin, 1.2)
    return s


def _check(e, f, v, E, F, W, N):
    assert abs(e - E) / N <= 1e-5 and abs(e - E) / abs(E) <= 1e-5, (e, E)
    assert np.abs(f - F).max() <= 1e-4 * max(1.0, np.abs(F).max()), np.abs(f - F).max()
    assert np.abs(v - W).max() <= 1e-4 * np.abs(W).max(), (v, W)


def test_cell_neighbour_sets_and_energy_force_virial(device_model, model0):
    from deepmd_jax import synthetic as syn
    print('case cell efv', flush=True)
    params, var = model0
    c0, box, types = syn.water_box(216, seed=21)
    coord, H = skewed(c0, box)
    coord = coord + np.array([-3.0, 40.0, 7.5])  # unwrapped input
    rc_skin = 6.5
    s = _system(device_model, coord, H, types, rc_skin)
    # neighbour sets (type-sorted indices) against the FP64 brute force
    _, rows = O.neighbor_sets_cell(coord, H, types, rc_skin)
    _, rows_in = O.neighbor_sets_cell(coord, H, types, rc_skin - 1e-4)
    _, rows_out = O.neighbor_sets_cell(coord, H, types, rc_skin + 1e-4)
    off = np.concatenate([[0], np.cumsum(np.bincount(types))])
    got = [s.nlist_export(j)[0] for j in range(2)]
    N = len(types)
    for a in range(N):
        mine = set()
        for j in range(2):
            mine |= {int(v) + off[j] for v in got[j][a] if v < N}
        assert set(rows_in[a].tolist()) <= mine <= set(rows_out[a].tolist()), a
    nb, _, _ = O.typed_nlist(coord, H, types, params['rcut'])
    E, F, W = O.energy_force_virial(params, var, coord, H, types, nbrs_nm=nb)
    e, f, v, _ = s.compute('fp32')
    _check(e[0], f[0], v[0], E.item(), F.numpy(), W.numpy(), N)


def test_rotated_box_matches_orthorhombic(device_model, model0):
    from scipy.spatial.transform import Rotation
    from deepmd_jax import synthetic as syn
    print('case cell rotated', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=22)
    R = Rotation.from_euler('xyz', [0.3, -0.5, 1.1]).as_matrix()
    e0, f0, v0, _ = _system(device_model, coord, box, types, params['rcut']).compute('fp32')
    e1, f1, v1, _ = _system(device_model, coord @ R.T, np.diag(box) @ R.T, types, params['rcut']).compute('fp32')
    _check(e1[0], f1[0], v1[0], e0[0], f0[0] @ R.T, R @ v0[0] @ R.T, len(types))


def test_cell_nve_against_oracle(device_model, model0):
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case cell nve', flush=True)
    params, var = model0
    c0, box, types = syn.water_box(128, seed=23)
    coord, H = skewed(c0, box)
    vel = syn.maxwell_boltzmann(types, 600.0, seed=23)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _system(device_model, coord, H, types, 6.2, m=m)
    s.set_velocities(vel)
    p = _lib.MDParams()
    p.routine, p.precision, p.dt = 0, 0, 0.5
    nsteps = 10
    tp, tv, err = s.md_run(p, nsteps)
    assert err == 0
    x, pm = coord.copy(), vel * m[:, None]

    def force(xx):
        nb, _, _ = O.typed_nlist(xx, H, types, params['rcut'])
        return O.energy_force_virial(params, var, xx, H, types, nbrs_nm=nb)[1].numpy()
    F = force(x)
    for _ in range(nsteps):
        pm += 0.25 * F
        x += 0.5 * pm / m[:, None]
        F = force(x)
        pm += 0.25 * F
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5


def test_cell_pimd_against_oracle(device_model, model0):
    """Ring-polymer BAOAB without noise in a general cell; two beads start one lattice
    vector away from bead 0's image, which the free-ring step unwraps."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.pimd import normal_mode_transform
    print('case cell pimd', flush=True)
    params, var = model0
    P, dt = 4, 0.25
    c0, box, types = syn.water_box(128, seed=24)
    coord, H = skewed(c0, box)
    rng = np.random.default_rng(24)
    x = coord[None] + rng.uniform(-0.05, 0.05, (P,) + coord.shape)
    x[1] += H[1]
    x[3] -= H[0] + H[2]
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT = 300 * TEMP_UNIT_CONVERSION
    pm = rng.standard_normal(x.shape) * np.sqrt(m * kT * P)[None, :, None]
    freqs, C = normal_mode_transform(P, kT)
    s = _system(device_model, x, H, types, 6.2, nrep=P, m=m)
    s.set_velocities(pm / m[None, :, None])
    s.pimd_setup(C, freqs, np.ones(P), np.zeros(P))
    p = _lib.MDParams()
    p.routine, p.dt = 3, dt
    tp, tv, err = s.md_run(p, 3)
    assert err == 0

    def forces(xx):
        out = []
        for b in range(P):
            nb, _, _ = O.typed_nlist(xx[b], H, types, params['rcut'])
            out.append(O.energy_force_virial(params, var, xx[b], H, types, nbrs_nm=nb)[1].numpy())
        return np.stack(out)
    F = forces(x)
    xo, po = x.copy(), pm.copy()
    for _ in range(3):
        xo, po, F = O.pimd_step(xo, po, m, forces, dt, C, freqs, np.ones(P), np.zeros(P), F, H)
    assert np.abs(tp[-1] - xo).max() < 1e-5
    assert np.abs(tv[-1] - po / m[None, :, None]).max() < 1e-5


def test_evaluate_general_cells(model_pkl, model0):
    """evaluate() with (F,3,3) boxes: a wide general cell and a narrow one (supercell of
    images), each against the oracle's lattice form."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.train import evaluate
    print('case cell evaluate', flush=True)
    params, var = model0
    for nmol, seed in ((128, 25), (16, 26)):
        frames = [skewed(*syn.water_box(nmol, seed=seed + k)[:2]) for k in range(2)]
        types = syn.water_box(nmol, seed=seed)[2]
        coord = np.stack([f[0] for f in frames])
        boxes = np.stack([f[1] for f in frames])
        with contextlib.redirect_stdout(io.StringIO()):
            out = evaluate(model_pkl, coord, boxes, types, batch_size=2)
        for k in range(2):
            E, F, _ = O.energy_force_virial(params, var, coord[k], boxes[k], types, lattice='auto')
            assert abs(out['energy'][k] - E.item()) / len(types) <= 1e-5, (nmol, k)
            assert np.abs(out['force'][k] - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())


def test_simulation_general_cell(model_pkl, model0):
    """Simulation with a (3,3) box: NVT (Nose-Hoover) conserves its invariant, positions
    stay finite; NPT and DPLR refuse general cells as the reference does."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case cell simulation', flush=True)
    c0, box, types = syn.water_box(128, seed=27)
    coord, H = skewed(c0, box)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(model_pkl, H, types, syn.MASSES, 'NVT', 0.25, coord, temperature=300.0,
                         initial_velocity=syn.maxwell_boltzmann(types, 300.0, seed=27), report_interval=10)
        traj = sim.run(60)
    log = np.array(sim.log)
    inv = log[:, 4]
    assert np.isfinite(log).all()
    assert traj['box'].shape[1:] == (3, 3)
    assert np.abs(inv - inv[0]).max() / len(types) < 2e-4, inv
    with pytest.raises(NotImplementedError):
        with contextlib.redirect_stdout(io.StringIO()):
            Simulation(model_pkl, H, types, syn.MASSES, 'NPT', 0.25, coord, temperature=300.0, pressure=1.0)


@pytest.mark.parametrize('nmol', [16, 128])
def test_ase_stress_general_cell(device_model, model0, nmol):
    """The ASE calculator's stress (deepmd_jax/ase_calc.py:79-143, which evaluates every cell
    on the general path) for a general cell: (1/V) dU/d(strain) in Voigt order within 1e-4
    of its largest component; 16 H2O runs as a supercell of images."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.evaluator import Evaluator, energy_forces_stress
    print('case cell ase', nmol, flush=True)
    params, var = model0
    c0, box, types = syn.water_box(nmol, seed=28)
    coord, H = skewed(c0, box)
    E, F, W = O.energy_force_virial(params, var, coord, H, types, lattice='auto')
    S = -W.numpy() / abs(np.linalg.det(H))
    ref = np.array([S[0, 0], S[1, 1], S[2, 2], S[1, 2], S[0, 2], S[0, 1]])
    e, f, s = energy_forces_stress(Evaluator(params, var, model=device_model), coord, H, types)
    assert abs(e - E.item()) / len(types) <= 1e-5
    assert np.abs(f - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())
    assert np.abs(s - ref).max() <= 1e-4 * np.abs(ref).max(), (s, ref)


def test_evaluate_frames_with_their_own_boxes(model_pkl, model0):
    """evaluate() batches consecutive frames with different boxes (same images per axis) as
    the segments of one system, each with its own cell: three scaled orthorhombic frames and
    two general cells, against the oracle frame by frame."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.train import evaluate
    print('case cell frames', flush=True)
    params, var = model0
    c0, box, types = syn.water_box(64, seed=29)
    coords, boxes = [], []
    for s in (1.0, 1.01, 0.995):
        coords.append(c0 * s)
        boxes.append(np.diag(box * s))
    for tilt in ((0.1, 0.0, -0.05), (0.0, 0.12, 0.07)):
        c, H = skewed(c0 * 1.3, box * 1.3, tilt)
        coords.append(c)
        boxes.append(H)
    coords, boxes = np.stack(coords), np.stack(boxes)
    with contextlib.redirect_stdout(io.StringIO()):
        out = evaluate(model_pkl, coords, boxes, types, batch_size=4)
    for k in range(len(coords)):
        E, F, _ = O.energy_force_virial(params, var, coords[k], boxes[k], types, lattice='auto')
        assert abs(out['energy'][k] - E.item()) / len(types) <= 1e-5, k
        assert np.abs(out['force'][k] - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max()), k


def test_pimd_simulation_general_cell(model_pkl):
    """Path-integral NVT (4 beads) in a general cell through Simulation: the PIMD invariant
    (spring energy over beads unwrapped through the cell, deepmd_jax/md.py:512-521) holds, the
    centroid trajectory is the bead mean after the fractional unwrap."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    from deepmd_jax.pimd import centroid
    print('case cell pimd simulation', flush=True)
    c0, box, types = syn.water_box(128, seed=30)
    coord, H = skewed(c0, box)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(model_pkl, H, types, syn.MASSES, 'NVT', 0.25, coord, temperature=300.0, n_bead=4,
                         pimd_print_invariant=True, report_interval=10, seed=3)
        traj = sim.run(40)
    log = np.array(sim.log)
    inv = log[:, 4]
    assert np.isfinite(log).all()
    assert np.abs(inv - inv[0]).max() / len(types) < 2e-3, inv
    np.testing.assert_allclose(traj['centroid'][-1], centroid(traj['position'][-1], H), atol=1e-9)

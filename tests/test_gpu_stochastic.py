"""Stochastic integrators on the device against the FP64 oracle, noise included: the BAOAB
Langevin step (deepmd_jax/routine.py:46-107) and the PILE-L ring-polymer step
(deepmd_jax/pimd.py:92-111,183-207).  The oracle restates the engine's Philox4x32-10 /
Box-Muller stream (pinned by its known-answer vectors in tests/test_stochastic.py), keyed as
the device keys it, so every O step is compared term by term: a wrong c1, c2, sqrt(m), key,
counter or atom index moves the trajectory by ~1e-4 A per step."""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _sorted_perm(types):
    return np.argsort(np.asarray(types), kind='stable')


def test_baoab_against_oracle(device_model, model0):
    """A 128-molecule request box (config-4 shape: first n sites of a seeded permutation),
    NVT Langevin at 300 K, gamma = 1/100 fs^-1, dt 0.5 fs: 12 steps, then 8 more in a second
    call (the counter continues), segment seed 5 and step counter 3 set through set_rng."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    print('case baoab', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(128, seed=21, permute_sites=True)
    coord = coord.astype(np.float32).astype(np.float64)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=4)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT, gamma, dt, seed = 300 * TEMP_UNIT_CONVERSION, 0.01, 0.5, 77
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.set_rng(seeds=[5], steps=[3])
    s.nlist_allocate(7.0, 1.2)
    p = _lib.MDParams()
    p.routine, p.dt, p.kT, p.gamma, p.seed = 2, dt, kT, gamma, seed
    tp1, tv1, err1 = s.md_run(p, 12)
    tp2, tv2, err2 = s.md_run(p, 8)
    assert err1 == 0 and err2 == 0
    work_dev = s.thermostat_energy()[0]
    assert s.get_rng()[1][0] == 3 + 20

    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    ff = lambda x: O.energy_force_virial(params, var, x, box, types, lattice=lat)[1].numpy()
    perm = _sorted_perm(types)
    x, pm = coord.copy(), vel * m[:, None]
    F = ff(x)
    work = 0.0
    traj = []
    for k in range(20):
        xi = np.empty((len(types), 3))
        xi[perm] = O.langevin_noise(seed, 5, 3 + k, len(types))
        x, pm, F, dke = O.nvt_langevin_step(x, pm, m, ff, dt, F, gamma, kT, xi)
        work += dke
        traj.append((x.copy(), pm / m[:, None]))
    tp = np.concatenate([tp1, tp2])
    tv = np.concatenate([tv1, tv2])
    for k in (0, 11, 12, 19):
        assert np.abs(tp[k] - traj[k][0]).max() < 1e-5, k
        assert np.abs(tv[k] - traj[k][1]).max() < 1e-5 * max(1.0, np.abs(traj[k][1]).max()), k
    assert abs(work_dev - work) < 1e-4 * max(1.0, abs(work)), (work_dev, work)
    # the noise is not negligible: the same run without it ends elsewhere
    xd, pd, Fd = coord.copy(), vel * m[:, None], ff(coord)
    for k in range(20):
        xd, pd, Fd, _ = O.nvt_langevin_step(xd, pd, m, ff, dt, Fd, gamma, kT, np.zeros((len(types), 3)))
    assert np.abs(xd - traj[-1][0]).max() > 1e-3


def test_pile_l_against_oracle(device_model, model0):
    """Ring-polymer BAOAB with PILE-L at 300 K (tau_t 100 fs, lambda 1, gamma_k = 2 omega_k),
    4 beads x 64 H2O, dt 0.25 fs: 6 + 4 steps in two calls (the counter continues)."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.pimd import normal_mode_transform, pile_l_factors
    print('case pile_l', flush=True)
    params, var = model0
    P, dt, seed = 4, 0.25, 31
    coord, box, types = syn.water_box(64, seed=3)
    rng = np.random.default_rng(5)
    x = (coord[None] + rng.uniform(-0.05, 0.05, (P,) + coord.shape)).astype(np.float32).astype(np.float64)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT = 300 * TEMP_UNIT_CONVERSION
    pm = rng.standard_normal(x.shape) * np.sqrt(m * kT * P)[None, :, None]
    freqs, C = normal_mode_transform(P, kT)
    _, c1, c2 = pile_l_factors(freqs, dt, kT, P, 100.0, 1.0)
    assert np.all(c1 < 1) and np.all(c2 > 0)
    s = _lib.DeviceSystem(device_model, types, P)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(x)
    s.set_velocities(pm / m[None, :, None])
    s.nlist_allocate(6.2, 1.2)
    s.pimd_setup(C, freqs, c1, c2)
    p = _lib.MDParams()
    p.routine, p.dt, p.seed = 3, dt, seed
    tp1, tv1, err1 = s.md_run(p, 6)
    tp2, tv2, err2 = s.md_run(p, 4)
    assert err1 == 0 and err2 == 0
    lat = O.lattice_candidate(np.diag(box), params['rcut'])

    def forces(xx):
        return np.stack([O.energy_force_virial(params, var, xx[b], box, types, lattice=lat)[1].numpy()
                         for b in range(P)])
    perm = _sorted_perm(types)
    N = len(types)
    F = forces(x)
    xo, po = x.copy(), pm.copy()
    traj = []
    for k in range(10):
        z = np.empty((P, N, 3))
        z[:, perm] = O.pimd_noise(seed, k, N, P)
        xo, po, F = O.pimd_step(xo, po, m, forces, dt, C, freqs, c1, c2, F, box, noise=z)
        traj.append((xo.copy(), po / m[None, :, None]))
    tp = np.concatenate([tp1, tp2])
    tv = np.concatenate([tv1, tv2])
    for k in (0, 5, 6, 9):
        assert np.abs(tp[k] - traj[k][0]).max() < 1e-5, k
        assert np.abs(tv[k] - traj[k][1]).max() < 1e-5 * max(1.0, np.abs(traj[k][1]).max()), k

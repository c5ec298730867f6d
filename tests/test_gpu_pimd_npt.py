"""Path-integral NPT (deepmd_jax/pimd.py:297-518) on the device against the FP64 oracle.

Trajectory parity: 4 beads x 216 H2O (L = 18.6 A), 10 steps with the stochastic terms off
(PILE-L c1 = 1, c2 = 0; box friction 0): boxes within 1e-7 relative, bead positions and
velocities within 1e-5 (A, A/fs), the centroid-virial dU/deps within 1e-4 relative.
Simulation surface: routine='NPT' with n_bead > 1 and pimd_print_invariant; the reported
invariant (deepmd_jax/md.py:540-557) stays within 2e-3 eV/atom over 100 thermostatted steps.
"""
import contextlib
import io

import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu

PRESS = 6.241509e-7  # bar -> eV/A^3
KT300 = 300 * 8.617333e-5


@pytest.mark.parametrize('couple', [((0, 1, 2),), ((0, 1), (2,))])
def test_pimd_npt_trajectory_vs_oracle(device_model, model0, couple):
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    from deepmd_jax.md import parse_couple_axes
    print('case pimd npt trajectory', couple, flush=True)
    params, var = model0
    P, steps, dt = 4, 10, 0.5
    coord, box, types = syn.water_box(216, seed=21)
    N = len(types)
    rng = np.random.default_rng(21)
    xb = coord[None] + rng.normal(0, 0.02, (P, N, 3))
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    pb = rng.standard_normal((P, N, 3)) * np.sqrt(m * KT300 * P)[None, :, None]
    freqs, C = O.ring_modes(P, KT300)
    s = _lib.DeviceSystem(device_model, types, P)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(xb)
    s.set_velocities(pb / m[None, :, None])
    s.pimd_setup(C, freqs, np.ones(P), np.zeros(P))
    s.nlist_allocate(6.5, 1.2)
    p = _lib.MDParams()
    p.routine, p.precision, p.dt, p.kT, p.gamma = 5, 0, dt, KT300, 0.0
    p.pressure, p.tau_p = 2.0e4 * PRESS, 200.0
    _, gid = parse_couple_axes(couple)
    for a in range(3):
        p.couple[a] = int(gid[a])
    tp, tv, err = s.md_run(p, steps)
    assert err == 0
    boxes = s.box_trajectory(steps)[:, 0]
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 1.0)

    def ff(x, b):
        _, F, W = O.energy_force_virial(params, var, x, b, types, nbrs_nm=nb)
        return F.numpy(), -np.diag(W.numpy())
    o = O.PIMDNPTOracle(xb, pb, m, box, ff, dt, KT300, p.pressure, C, freqs, np.ones(P), np.zeros(P), 0.0,
                        p.tau_p, couple_axes=couple)
    for k in range(steps):
        o.step()
        assert np.abs(boxes[k] - o.box).max() <= 1e-7 * o.box.max(), (k, boxes[k], o.box)
    assert np.abs(boxes[-1] - box).max() > 1e-5  # the box moved
    assert np.abs(tp[-1] - o.x).max() < 1e-5
    assert np.abs(tv[-1] - o.p / m[None, :, None]).max() < 1e-5
    st = s.npt_state()
    G = len(couple)
    np.testing.assert_allclose(st[9:9 + G], o.dUdV, rtol=1e-4, atol=1e-4 * np.abs(o.dUdV).max())
    np.testing.assert_allclose(st[3:3 + G], o.Pb, rtol=1e-4, atol=1e-6 * np.abs(o.Mb).max() ** 0.5)


def test_pimd_npt_simulation_invariant(model_pkl):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case pimd npt simulation', flush=True)
    coord, box, types = syn.water_box(216, seed=22)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(model_pkl, box, types, syn.MASSES, 'NPT', 0.25, coord, temperature=300.0, pressure=1.0e4,
                         n_bead=4, tau_p=200.0, pimd_print_invariant=True, report_interval=20, seed=5)
        traj = sim.run(100)
    keys = ['Step', 'Temperature', 'KE', 'PE', 'Pressure', 'box_x', 'box_y', 'box_z', 'Invariant']
    log = np.array(sim.log)
    assert log.shape[1] == len(keys) + 1
    inv = log[:, keys.index('Invariant')]
    assert np.isfinite(log).all()
    assert np.abs(inv - inv[0]).max() / len(types) < 2e-3, inv
    assert np.abs(traj['box'][-1] - box).max() > 1e-5
    assert traj['position'].shape == (101, 4, len(types), 3) and traj['centroid'].shape == (101, len(types), 3)
    np.testing.assert_allclose(traj['box'][-1], sim.getBox(), rtol=0, atol=0)
    # NVT PIMD reports its invariant too when asked
    with contextlib.redirect_stdout(io.StringIO()):
        sim2 = Simulation(model_pkl, box, types, syn.MASSES, 'NVT', 0.25, coord, temperature=300.0, n_bead=4,
                          pimd_print_invariant=True, report_interval=20, seed=6)
        sim2.run(60)
    inv2 = np.array(sim2.log)[:, 4]
    assert np.abs(inv2 - inv2[0]).max() / len(types) < 2e-3, inv2

"""GPU NPT (deepmd_jax/routine.py:161-456) through the C ABI against the FP64 oracle.

Trajectory parity: 20 steps of 216 H2O (L = 18.6 A) with the device NPT step (FP32 forces,
FP64 state and dU/deps from the FP64-reduced virial) against the oracle's NPT step with
FP64 forces: positions and velocities within 1e-5, boxes within 1e-7 relative.
Conservation: PE + KE + thermostat/barostat terms over 300 steps within 1e-3 eV/atom
(for G > 1 groups the chain head enters as G kT xi_0, see tests/test_npt.py).
"""
import contextlib
import io

import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu

PRESS = 6.241509e-7  # bar -> eV/A^3
KT300 = 300 * 8.617333e-5


def _npt_params(couple, pressure_bar=2.0e4, tau_p=200.0):
    from deepmd_jax import _lib
    from deepmd_jax.md import parse_couple_axes
    p = _lib.MDParams()
    p.routine, p.precision, p.dt, p.kT, p.tau, p.chain_length = 4, 0, 0.5, KT300, 100.0, 3
    p.pressure, p.tau_p, p.chain_length_p = pressure_bar * PRESS, tau_p, 3
    _, gid = parse_couple_axes(couple)
    for a in range(3):
        p.couple[a] = int(gid[a])
    return p


def _system(dm, coord, vel, m, box, types):
    from deepmd_jax import _lib
    s = _lib.DeviceSystem(dm, types, 1)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.nlist_allocate(6.5, 1.2)
    return s


@pytest.mark.parametrize('couple', [((0, 1, 2),), ((0, 1), (2,))])
def test_npt_trajectory_vs_oracle(device_model, model0, couple):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case npt trajectory', couple, flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=7)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=7)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _system(device_model, coord, vel, m, box, types)
    p = _npt_params(couple)
    tp, tv, err = s.md_run(p, 20)
    assert err == 0
    boxes = s.box_trajectory(20)[:, 0]
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 1.0)

    def force_fn(x, b):
        _, F, W = O.energy_force_virial(params, var, x, b, types, nbrs_nm=nb)
        return F.numpy(), -np.diag(W.numpy())
    o = O.NPTOracle(coord, vel * m[:, None], m, box, force_fn, 0.5, KT300, p.pressure, 100.0, p.tau_p, 3, 3,
                    couple_axes=couple)
    for k in range(20):
        o.step()
        assert np.abs(boxes[k] - o.box).max() <= 1e-7 * o.box.max(), (k, boxes[k], o.box)
    assert np.abs(boxes[-1] - box).max() > 1e-5  # the box moved
    assert np.abs(tp[-1, 0] - o.x).max() < 1e-5
    assert np.abs(tv[-1, 0] - o.p / m[:, None]).max() < 1e-5
    np.testing.assert_allclose(s.get_boxes()[0], boxes[-1], rtol=0, atol=0)


@pytest.mark.parametrize('couple', [((0, 1, 2),), ((0, 1), (2,)), ((0,),)])
def test_npt_invariant_conserved_on_device(device_model, couple):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case npt invariant', couple, flush=True)
    coord, box, types = syn.water_box(216, seed=8)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=8)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _system(device_model, coord, vel, m, box, types)
    p = _npt_params(couple)
    G = len(couple)

    def H():
        e, _, _, _ = s.compute('fp32')
        return e[0] + s.kinetic_energy()[0] + s.thermostat_energy()[0] + (G - 1) * KT300 * s.npt_state()[16]
    rc_skin = [6.5]

    def run(n):
        # the compressing box can overflow the neighbour rows (bit 2) or shrink past what the
        # list's skin covers (bit 1): widen the skin by 0.4 A on bit 1, reallocate on the
        # current box and re-run the chunk from its entry state, as Simulation.run does
        # (deepmd_jax/md.py:644-676)
        for _ in range(6):
            _, _, err = s.md_run(p, n, record=False)
            if not err & 3:
                assert err == 0
                return
            if err & 1:
                rc_skin[0] += 0.4
            s.nlist_allocate(rc_skin[0], 1.3)
        raise AssertionError('neighbour list kept overflowing')
    run(1)
    h0, b0 = H(), s.get_boxes()[0]
    for _ in range(3):
        run(100)
    h1, b1 = H(), s.get_boxes()[0]
    assert abs(h1 - h0) / len(types) < 1e-3, (h0, h1)
    moving = [a for g in couple for a in g]
    assert all(abs(b1[a] - b0[a]) > 1e-4 for a in moving)
    assert all(b1[a] == b0[a] for a in range(3) if a not in moving)


def test_npt_simulation_surface(model0, device_model):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    print('case npt simulation', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=9, permute_sites=True)
    mk = dict(model=(DPModel(params), var, device_model), report_interval=10)
    with contextlib.redirect_stdout(io.StringIO()):
        with pytest.raises(ValueError):
            Simulation(None, box, types, syn.MASSES, 'NPT', 0.5, coord, temperature=300.0, **mk)
        with pytest.raises(NotImplementedError):
            c2, b2, t2 = syn.water_box(64, seed=0)
            Simulation(None, b2, t2, syn.MASSES, 'NPT', 0.5, c2, temperature=300.0, pressure=1.0, **mk)
        sim = Simulation(None, box, types, syn.MASSES, 'NPT', 0.5, coord, temperature=300.0, pressure=2.0e4,
                         tau_p=200.0, couple_axes=((0, 1), (2,)), **mk)
        traj = sim.run(25)
    assert traj['box'].shape == (26, 3)
    np.testing.assert_allclose(traj['box'][0], box, rtol=0, atol=1e-12)
    np.testing.assert_allclose(traj['box'][-1], sim.getBox(), rtol=0, atol=0)
    assert abs(traj['box'][-1, 2] - box[2]) > 1e-6 and traj['box'][-1, 0] == traj['box'][-1, 1]
    assert np.isfinite(sim.getPressure())
    # step, Temperature KE PE Pressure box_x box_y box_z Invariant, time
    assert len(sim.log) >= 3 and len(sim.log[0]) == 1 + 8 + 1
    # the step-0 invariant carries P V of the box at rest (no jump at the first step)
    inv = np.array(sim.log)[:, 8]
    assert np.abs(inv - inv[0]).max() / len(types) < 1e-3, inv
    # setPosition with an isotropically scaled box (deepmd_jax/md.py:985-1015)
    b0, x0 = sim.getBox(), sim.getPosition()
    with pytest.raises(ValueError):
        sim.setPosition(x0, box=b0 * np.array([1.0, 1.0, 1.01]))
    sim.setPosition(x0 * 1.005, box=b0 * 1.005)
    np.testing.assert_allclose(sim.getBox(), b0 * 1.005, rtol=1e-12)
    assert np.isfinite(sim.getEnergy())

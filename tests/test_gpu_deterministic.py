"""Deterministic forces (dpmd_system_set_deterministic): pair gradients written once by skin
slot and gathered per atom in row order through a reverse index rebuilt with the list,
instead of FP32 atomics.  Two evaluations of the same positions give bit-identical forces,
the forces agree with the atomics path and the FP64 oracle to the FP32 tolerance, and two MD
runs from the same state (with drift-triggered rebuilds) give bit-identical trajectories."""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _system(device_model, coord, box, types, rc_skin, masses=None, det=True):
    from deepmd_jax import _lib
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)) if masses is None else masses)
    s.set_positions(coord)
    s.nlist_allocate(rc_skin, 1.2)
    s.set_deterministic(det)
    return s


@pytest.mark.parametrize('nmol,prec', [(512, 'fp32'), (512, 'bf16'), (4096, 'fp32')])
def test_deterministic_forces_bit_identical(device_model, model0, nmol, prec):
    from deepmd_jax import synthetic as syn
    print('case deterministic', nmol, prec, flush=True)
    params, var = model0
    coord, box, types = syn.water_box(nmol, seed=2)
    coord = np.asarray(coord, np.float32).astype(np.float64)
    s = _system(device_model, coord, box, types, 7.0)
    e1, f1, v1, _ = s.compute(prec)
    e2, f2, v2, _ = s.compute(prec)
    assert np.array_equal(f1, f2) and np.array_equal(e1, e2) and np.array_equal(v1, v2)
    a = _system(device_model, coord, box, types, 7.0, det=False)
    ea, fa, va, _ = a.compute(prec)
    fs = max(1.0, np.abs(fa).max())
    assert np.abs(f1 - fa).max() <= 1e-5 * fs
    assert abs(e1[0] - ea[0]) <= 1e-9 * abs(ea[0])
    if prec == 'fp32' and nmol == 512:
        nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'])
        E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
        assert np.abs(f1[0] - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())
    # switching back to atomics and on again keeps the reverse index consistent
    s.set_deterministic(False)
    s.set_deterministic(True)
    e3, f3, _, _ = s.compute(prec)
    assert np.array_equal(f1, f3)


def test_deterministic_md_trajectories_bit_identical(device_model):
    """40 NVE steps (dt 1 fs, skin 0.5 A: several drift-triggered rebuilds, each rebuilding the
    reverse index on the device) twice from the same state: identical trajectories."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case deterministic md', flush=True)
    coord, box, types = syn.water_box(4096, seed=3)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    vel = syn.maxwell_boltzmann(types, 600.0, seed=3)
    trajs, rebuilds = [], []
    for _ in range(2):
        s = _system(device_model, coord, box, types, 6.5, masses=m)
        s.nlist_allocate(6.5, 1.5)
        s.set_velocities(vel)
        p = _lib.MDParams()
        p.routine, p.dt = 0, 1.0
        tp, tv, err = s.md_run(p, 40)
        assert err == 0
        trajs.append((tp, tv))
        rebuilds.append(s.nlist_rebuilds())
    assert rebuilds[0] >= 2 and rebuilds[0] == rebuilds[1]
    assert np.array_equal(trajs[0][0], trajs[1][0]) and np.array_equal(trajs[0][1], trajs[1][1])


def test_simulation_deterministic_option(device_model, model0):
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    print('case deterministic simulation', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=6, permute_sites=True)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=6)
    trajs = []
    for _ in range(2):
        with contextlib.redirect_stdout(io.StringIO()):
            sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=vel,
                             report_interval=10 ** 6, neighbor_skin=1.0, deterministic=True,
                             model=(DPModel(params), var, device_model))
            trajs.append(sim.run(15))
    assert np.array_equal(trajs[0]['position'], trajs[1]['position'])
    assert np.array_equal(trajs[0]['velocity'], trajs[1]['velocity'])

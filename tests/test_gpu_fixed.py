"""GPU parity of fixed atoms, Nose-Hoover NVT and per-step centre-of-mass removal.

Device FP64 integrators with FP32 forces against the FP64 oracle (tests/test_fixed.py pins
the oracle).  Tolerances: forces 1e-4 * max(1, max |F|) eV/A, energies 1e-5 eV/atom,
positions and velocities after 20 steps 1e-5 (A, A/fs), the Nose-Hoover chain energy
1e-5 eV.  Fixed atoms: forces exactly zero, positions bit-identical to the start.
"""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _mobile(n, seed=0, frac=0.25):
    rng = np.random.default_rng(seed)
    mob = np.ones(n, bool)
    mob[rng.choice(n, int(frac * n), replace=False)] = False
    return mob


def _setup(device_model, nmol=64, seed=0, mob=None):
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    coord, box, types = syn.water_box(nmol, seed=seed)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    vel = syn.maxwell_boltzmann(types, 300.0, seed=seed)
    if mob is not None:
        vel = vel * mob[:, None]
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(coord)
    s.set_velocities(vel)
    if mob is not None:
        s.set_fixed(~mob)
    s.nlist_allocate(6.2, 1.2)
    return s, coord, box, types, m, vel


def test_fixed_atom_forces(device_model, model0):
    print('case fixed forces', flush=True)
    params, var = model0
    mob = _mobile(192, seed=4)
    s, coord, box, types, m, _ = _setup(device_model, 64, 4, mob)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto', mobile=mob)
    e, f, v, _ = s.compute()
    F = F.numpy()
    assert np.all(f[0][~mob] == 0)
    assert np.abs(f[0] - F).max() <= 1e-4 * max(1.0, np.abs(F).max())
    assert abs(e[0] - E.item()) / len(types) <= 1e-5
    assert np.abs(v[0] - W.numpy()).max() <= 1e-4 * np.abs(W.numpy()).max()
    # clearing the mask restores the unconstrained forces
    s.set_fixed(None)
    _, f2, _, _ = s.compute()
    F0 = O.energy_force_virial(params, var, coord, box, types, lattice='auto')[1].numpy()
    assert np.abs(f2[0] - F0).max() <= 1e-4 * max(1.0, np.abs(F0).max())


@pytest.mark.parametrize('remove_com', [False, True])
def test_fixed_nve_against_oracle(device_model, model0, remove_com):
    from deepmd_jax import _lib
    print('case fixed nve remove_com=%s' % remove_com, flush=True)
    params, var = model0
    mob = _mobile(192, seed=5)
    s, coord, box, types, m, vel = _setup(device_model, 64, 5, mob)
    p = _lib.MDParams()
    p.routine, p.dt, p.remove_com = 0, 0.5, int(remove_com)
    nsteps = 20
    tp, tv, err = s.md_run(p, nsteps)
    assert err == 0
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    ff = lambda x: O.energy_force_virial(params, var, x, box, types, lattice=lat, mobile=mob)[1].numpy()
    x, pm = coord.copy(), vel * m[:, None]
    F = ff(x)
    for k in range(nsteps):
        x, pm, F = O.vv_step(x, pm, m, ff, 0.5, F, mobile=mob)
        if remove_com:
            pm = O.remove_com(pm, m, mob)
        assert np.array_equal(tp[k, 0][~mob], coord[~mob])
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5
    if remove_com:
        w = (m * mob)[:, None]
        assert np.abs((tv[-1, 0] * w).sum(0)).max() < 1e-9 * w.sum()
    else:
        assert np.all(tv[:, 0][:, ~mob] == 0)


@pytest.mark.parametrize('fixed,cs,sy', [(False, 1, 1), (True, 1, 1), (False, 2, 3)])
def test_nvt_nose_hoover_against_oracle(device_model, model0, fixed, cs, sy):
    """20 Nose-Hoover chain steps (chain length 3, tau 20 fs, 400 K over a 300 K start; chain_steps
    x sy_steps substeps) against the oracle step, and the chain energy of the invariant."""
    from deepmd_jax import _lib
    from deepmd_jax.constants import TEMP_UNIT_CONVERSION
    print('case nvt nh fixed=%s' % fixed, flush=True)
    params, var = model0
    mob = _mobile(192, seed=6) if fixed else None
    s, coord, box, types, m, vel = _setup(device_model, 64, 6, mob)
    kT, tau = 400 * TEMP_UNIT_CONVERSION, 20.0
    p = _lib.MDParams()
    p.routine, p.dt, p.kT, p.tau, p.chain_length = 1, 0.5, kT, tau, 3
    p.chain_steps, p.sy_steps = cs, sy
    nsteps = 20
    tp, tv, err = s.md_run(p, nsteps)
    assert err == 0
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    ff = lambda x: O.energy_force_virial(params, var, x, box, types, lattice=lat, mobile=mob)[1].numpy()
    x, pm = coord.copy(), vel * m[:, None]
    F = ff(x)
    xi, px = np.zeros(3), np.zeros(3)
    for _ in range(nsteps):
        x, pm, F = O.nvt_nh_step(x, pm, m, ff, 0.5, F, xi, px, kT, tau, mobile=mob, chain_steps=cs, sy_steps=sy)
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5
    dof = 3 * (len(types) if mob is None else int(mob.sum()))
    assert abs(s.thermostat_energy()[0] - O.nh_chain_energy(xi, px, kT, tau, dof)) < 1e-5
    if fixed:
        assert np.array_equal(tp[-1, 0][~mob], coord[~mob])


def test_simulation_fixed_indices_and_remove_com(model_pkl):
    """Simulation surface: fixed_indices with NVT_langevin and remove_com_motion, the
    mobile-dof temperature, setVelocity masking and remove_com_motion(); NPT with fixed atoms
    raises as in the reference."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case simulation fixed', flush=True)
    coord, box, types = syn.water_box(64, seed=7)
    fixed = np.arange(0, 192, 4)
    mob = np.ones(192, bool)
    mob[fixed] = False
    with pytest.raises(NotImplementedError):
        Simulation(model_pkl, box, types, syn.MASSES, 'NPT', 0.5, coord, temperature=300, pressure=1.0,
                   fixed_indices=fixed)
    sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVT_langevin', 0.5, coord, temperature=300.0,
                     fixed_indices=fixed, remove_com_motion=True, report_interval=50, seed=3)
    assert np.all(sim.getVelocity()[fixed] == 0)
    assert np.all(sim.getForce()[fixed] == 0)
    traj = sim.run(100)
    assert np.isfinite(traj['position']).all()
    assert np.array_equal(traj['position'][:, fixed], np.broadcast_to(coord[fixed], traj['position'][:, fixed].shape))
    m = np.array(syn.MASSES)[types] * mob
    com = (traj['velocity'][1:] * m[None, :, None]).sum(1)
    assert np.abs(com).max() < 1e-9 * m.sum()
    # the report's temperature counts the mobile degrees of freedom
    v = sim.getVelocity()
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    mm = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    ke = 0.5 * (mm[:, None] * v ** 2).sum()
    rep = sim._report_values()
    assert abs(rep['Temperature'] - 2 * ke / (3 * mob.sum()) / TEMP_UNIT_CONVERSION) < 1e-6 * rep['Temperature']
    sim.setVelocity(np.ones((192, 3)))
    assert np.all(sim.getVelocity()[fixed] == 0)
    sim.remove_com_motion()
    vv = sim.getVelocity()
    assert np.abs((vv * m[:, None]).sum(0)).max() < 1e-9


def test_debug_neighbour_update_rate(model_pkl):
    """Simulation(debug=True): the NbrUpdateRate column is the per-step moving average
    profile * (1 - 1/100) + update / 100 of deepmd_jax/md.py:733.  With a zero skin every step
    rebuilds, so after k steps it is 1 - 0.99^k; the column is absent below the reference's
    neighbour-list condition."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case debug rate', flush=True)
    coord, box, types = syn.water_box(216, seed=60)
    with contextlib.redirect_stdout(io.StringIO()) as out:
        sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVE', 0.5, coord, debug=True, report_interval=50,
                         neighbor_skin=0.0)
        sim.run(100)
    assert 'NbrUpdateRate' in out.getvalue()
    rates = [row[-1] for row in sim.log]
    assert rates[0] == 0.0
    np.testing.assert_allclose(rates[1:], [1 - 0.99 ** 50, 1 - 0.99 ** 100], atol=1e-5)
    c2, b2, t2 = syn.water_box(64, seed=61)
    with contextlib.redirect_stdout(io.StringIO()) as out2:
        sim2 = Simulation(model_pkl, b2, t2, syn.MASSES, 'NVE', 0.5, c2, debug=True, report_interval=10)
        sim2.run(10)
    assert 'NbrUpdateRate' not in out2.getvalue()

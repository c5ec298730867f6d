"""Models of one, three and four atom types on the device against the FP64 oracle.

The seeded water boxes are relabelled: one type (every atom type 0), three types (every third
O becomes type 2) and four types (every fifth H also becomes type 3); each model is seeded for
its type count with statistics from its own frame.  Checks: neighbour lists bit-exact against
the oracle's FP32 restatement, energy / forces / virial at the FP32 tolerances of the parity
tests, and 10 velocity-Verlet steps within 1e-5 (A, A/fs) of the oracle's.
"""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _frame(ntypes, nmol=216, seed=90):
    from deepmd_jax import synthetic as syn
    coord, box, types = syn.water_box(nmol, seed=seed)
    t = np.array(types)
    if ntypes == 1:
        t[:] = 0
    if ntypes >= 3:
        o = np.nonzero(t == 0)[0]
        t[o[::3]] = 2
    if ntypes == 4:
        h = np.nonzero(t == 1)[0]
        t[h[::5]] = 3
    return np.asarray(coord, np.float32).astype(np.float64), box, t


@pytest.mark.parametrize('ntypes', [1, 3, 4])
def test_type_counts(ntypes):
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case types', ntypes, flush=True)
    coord, box, types = _frame(ntypes)
    params, var = syn.init_model(3, coord=coord, box=box, type_idx=types, ntypes=ntypes)
    dm = _lib.DeviceModel(params, var)
    s = _lib.DeviceSystem(dm, types, 1)
    s.set_box(box)
    m = np.full(len(types), 15.9994 * MASS_UNIT_CONVERSION)
    m[types == 1] = 1.0078 * MASS_UNIT_CONVERSION
    s.set_masses(m)
    s.set_positions(coord)
    rc_skin = params['rcut'] + 0.5
    s.nlist_allocate(rc_skin, 1.2)
    ref, _, knbr = O.typed_nlist(coord, box, types, rc_skin, ntypes=ntypes)
    assert s.knbr == knbr
    off = np.concatenate([[0], np.cumsum(np.bincount(types, minlength=ntypes))])
    for j in range(ntypes):
        got = s.nlist_export(j)[0]
        for i in range(ntypes):
            np.testing.assert_array_equal(got[off[i]:off[i + 1]], ref[i][j])
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'], ntypes=ntypes)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
    E, F, W = E.item(), F.numpy(), W.numpy()
    e, f, v, _ = s.compute('fp32')
    N = len(types)
    assert abs(e[0] - E) / N <= 1e-5 and abs(e[0] - E) / abs(E) <= 1e-5, (e[0], E)
    assert np.abs(f[0] - F).max() <= 1e-4 * max(1.0, np.abs(F).max())
    assert np.abs(v[0] - W).max() <= 1e-4 * np.abs(W).max()
    # the BF16 form within its stated force error
    f16 = s.compute('bf16')[1]
    assert np.sqrt(((f16[0] - F) ** 2).mean()) <= 1e-3
    # velocity Verlet against the oracle
    vel = syn.maxwell_boltzmann(np.minimum(types, 1), 300.0, seed=90)
    s.set_velocities(vel)
    p = _lib.MDParams()
    p.routine, p.dt = 0, 0.5
    tp, tv, err = s.md_run(p, 10)
    assert err == 0
    x, pm = coord.copy(), vel * m[:, None]

    def force(xx):
        nbx, _, _ = O.typed_nlist(xx, box, types, params['rcut'], ntypes=ntypes)
        return O.energy_force_virial(params, var, xx, box, types, nbrs_nm=nbx)[1].numpy()
    F = force(x)
    for _ in range(10):
        pm += 0.25 * F
        x += 0.5 * pm / m[:, None]
        F = force(x)
        pm += 0.25 * F
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5


def test_compressed_three_types():
    """compress_model of a three-type model (9 tabulated nets) against the oracle's tabulated
    restatement at the FP32 tolerances."""
    import copy
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import compress_model
    print('case types compressed', flush=True)
    coord, box, types = _frame(3, seed=91)
    params, var = syn.init_model(4, coord=coord, box=box, type_idx=types, ntypes=3)
    m, cvar = compress_model(DPModel(copy.deepcopy(params)), copy.deepcopy(var), 1024, 0.6)
    cparams = m.params
    s = _lib.DeviceSystem(_lib.DeviceModel(cparams, cvar), types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)))
    s.set_positions(coord)
    s.nlist_allocate(cparams['rcut'], 1.2)
    e, f, v, _ = s.compute('fp32')
    nb, _, _ = O.typed_nlist(coord, box, types, cparams['rcut'], ntypes=3)
    E, F, W = O.energy_force_virial(cparams, cvar, coord, box, types, nbrs_nm=nb)
    E, F, W = E.item(), F.numpy(), W.numpy()
    N = len(types)
    assert abs(e[0] - E) / N <= 1e-5, (e[0], E)
    assert np.abs(f[0] - F).max() <= 1e-4 * max(1.0, np.abs(F).max())
    assert np.abs(v[0] - W).max() <= 1e-4 * np.abs(W).max()


def test_simulation_bf16_precision(model_pkl, model0):
    """Simulation(precision='bf16'): the run integrates with BF16-GEMM forces (their RMSE against
    the FP64 oracle within 1e-3 eV/A at the end state), KE + PE holds to 1e-3 eV/atom over 40 steps
    of 0.25 fs, and the reported PE is the BF16 evaluation's."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case simulation bf16', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(64, seed=95)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVE', 0.25, coord, report_interval=10, precision='bf16',
                         initial_velocity=syn.maxwell_boltzmann(types, 300.0, seed=95))
        sim.run(40)
    log = np.array(sim.log)
    inv = log[:, 4]
    assert np.abs(inv - inv[0]).max() / len(types) < 1e-3, inv
    x = sim.getPosition()
    F = O.energy_force_virial(params, var, x, box, types, lattice='auto')[1].numpy()
    f = sim.getForce()
    assert np.sqrt(((f - F) ** 2).mean()) <= 1e-3
    assert abs(sim.getEnergy() - log[-1, 3]) <= 1e-6 * abs(log[-1, 3]) + 1e-6

"""GPU parity of the tabulated (compressed) embedding path against the CPU oracle's
tabulated restatement (deepmd_jax/utils.py:126-151), through the C ABI.

Tolerances as for the MLP path (FP32 device arithmetic against the FP64 oracle on the same
float32 table): energy |dE|/N <= 1e-5 eV/atom and |dE|/|E| <= 1e-5; forces
max |dF| <= 1e-4 * max(1, max |F|); virial max |dW| <= 1e-4 * max |W|.  With BF16 / FP16
fitting GEMMs: energy/atom and force RMSE <= 1e-3.
"""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _system(dm, coord, box, types, rc_skin):
    from deepmd_jax import _lib
    s = _lib.DeviceSystem(dm, types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)))
    s.set_positions(coord[None])
    s.nlist_allocate(rc_skin, 1.2)
    return s


def _check_fp32(e, f, v, E, F, W, N):
    assert abs(e - E) / N <= 1e-5 and abs(e - E) / abs(E) <= 1e-5, (e, E)
    assert np.abs(f - F).max() <= 1e-4 * max(1.0, np.abs(F).max()), np.abs(f - F).max()
    assert np.abs(v - W).max() <= 1e-4 * np.abs(W).max(), (v, W)


@pytest.mark.parametrize('nmol', [64, 216])
def test_compressed_efv(device_model_c, model0c, nmol):
    from deepmd_jax import synthetic as syn
    print('case compressed efv', nmol, flush=True)
    params, var = model0c
    coord, box, types = syn.water_box(nmol, seed=21)
    coord = np.asarray(coord, np.float32).astype(np.float64)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    s = _system(device_model_c, coord, box, types, params['rcut'])
    e, f, v, _ = s.compute('fp32')
    _check_fp32(e[0], f[0], v[0], E.item(), F.numpy(), W.numpy(), len(types))


def test_compressed_neighbour_list_box(device_model_c, model0c):
    """1,000 H2O (L = 31 A) through the oracle's typed neighbour list at rc + 1 A."""
    from deepmd_jax import synthetic as syn
    print('case compressed nlist box', flush=True)
    params, var = model0c
    coord, box, types = syn.water_box(1000, seed=22)
    coord = np.asarray(coord, np.float32).astype(np.float64)
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 1.0)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
    s = _system(device_model_c, coord, box, types, params['rcut'] + 1.0)
    e, f, v, _ = s.compute('fp32')
    _check_fp32(e[0], f[0], v[0], E.item(), F.numpy(), W.numpy(), len(types))


@pytest.mark.parametrize('prec', ['bf16', 'fp16'])
def test_compressed_half_precision_fitting(device_model_c, model0c, prec):
    from deepmd_jax import synthetic as syn
    print('case compressed', prec, flush=True)
    params, var = model0c
    coord, box, types = syn.water_box(216, seed=23)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E, F, W = E.item(), F.numpy(), W.numpy()
    s = _system(device_model_c, coord, box, types, params['rcut'])
    e, f, v, _ = s.compute(prec)
    N = len(types)
    rmse = np.sqrt(((f[0] - F) ** 2).mean())
    print(prec, 'dE/atom', abs(e[0] - E) / N, 'F rmse', rmse, flush=True)
    assert abs(e[0] - E) / N <= 1e-3
    assert rmse <= 1e-3
    assert np.abs(v[0] - W).max() <= 1e-2 * np.abs(W).max()


def test_compressed_nve_trajectory(device_model_c, model0c):
    """20 velocity-Verlet steps with the tabulated model against the FP64 oracle."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case compressed nve', flush=True)
    params, var = model0c
    coord, box, types = syn.water_box(64, seed=0)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=0)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _lib.DeviceSystem(device_model_c, types, 1)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.nlist_allocate(6.2, 1.2)
    p = _lib.MDParams()
    p.routine, p.precision, p.dt = 0, 0, 0.5
    tp, tv, err = s.md_run(p, 20)
    assert err == 0
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    x, pm = coord.copy(), vel * m[:, None]
    F = O.energy_force_virial(params, var, x, box, types, lattice=lat)[1].numpy()
    for k in range(20):
        pm += 0.25 * F
        x += 0.5 * pm / m[:, None]
        F = O.energy_force_virial(params, var, x, box, types, lattice=lat)[1].numpy()
        pm += 0.25 * F
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5


def test_compressed_model_pkl_evaluate_and_simulation(tmp_path, model0c):
    """A compressed model.pkl goes through evaluate() and Simulation like any other."""
    import contextlib
    import copy
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.train import evaluate
    from deepmd_jax.utils import save_model
    print('case compressed model.pkl', flush=True)
    params, var = model0c
    path = str(tmp_path / 'model_c.pkl')
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(copy.deepcopy(params)), var)
    coord, box, types = syn.water_box(216, seed=24)
    coord = np.asarray(coord, np.float32).astype(np.float64)
    E, F, _ = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    with contextlib.redirect_stdout(io.StringIO()):
        out = evaluate(path, coord[None], np.asarray(box)[None], types)
    assert abs(out['energy'][0] - E.item()) / len(types) <= 1e-5
    assert np.abs(out['force'][0] - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(path, box, types, syn.MASSES, 'NVE', 0.5, coord,
                         initial_velocity=syn.maxwell_boltzmann(types, 300.0, seed=5), report_interval=10 ** 6)
        assert abs(sim.getEnergy() - E.item()) / len(types) <= 1e-5
        traj = sim.run(10)
    assert np.isfinite(traj['position']).all()


def test_compressed_pimd_and_general_cell(device_model_c, model0c):
    """The tabulated kernels over several segments and in a general cell: ring-polymer BAOAB
    without noise (4 beads x 64 H2O, 3 steps) against the oracle's tabulated restatement, and
    the energy / forces of a skewed cell."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.pimd import normal_mode_transform
    print('case compressed pimd / cell', flush=True)
    params, var = model0c
    P, dt = 4, 0.25
    coord, box, types = syn.water_box(64, seed=33)
    rng = np.random.default_rng(33)
    x = coord[None] + rng.uniform(-0.05, 0.05, (P,) + coord.shape)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT = 300 * TEMP_UNIT_CONVERSION
    pm = rng.standard_normal(x.shape) * np.sqrt(m * kT * P)[None, :, None]
    freqs, C = normal_mode_transform(P, kT)
    s = _lib.DeviceSystem(device_model_c, types, P)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(x)
    s.set_velocities(pm / m[None, :, None])
    s.nlist_allocate(6.2, 1.2)
    s.pimd_setup(C, freqs, np.ones(P), np.zeros(P))
    p = _lib.MDParams()
    p.routine, p.dt = 3, dt
    tp, tv, err = s.md_run(p, 3)
    assert err == 0
    lat = O.lattice_candidate(np.diag(box), params['rcut'])

    def forces(xx):
        return np.stack([O.energy_force_virial(params, var, xx[b], box, types, lattice=lat)[1].numpy()
                         for b in range(P)])
    F = forces(x)
    xo, po = x.copy(), pm.copy()
    for _ in range(3):
        xo, po, F = O.pimd_step(xo, po, m, forces, dt, C, freqs, np.ones(P), np.zeros(P), F, box)
    assert np.abs(tp[-1] - xo).max() < 1e-5
    assert np.abs(tv[-1] - po / m[None, :, None]).max() < 1e-5
    # general cell
    c0, b0, t0 = syn.water_box(216, seed=34)
    H = np.array([[b0[0], 0, 0], [0.2 * b0[1], b0[1], 0], [-0.1 * b0[2], 0.15 * b0[2], b0[2]]])
    c = (c0 / b0) @ H
    s2 = _lib.DeviceSystem(device_model_c, t0, 1)
    s2.set_box(H)
    s2.set_masses(np.ones(len(t0)))
    s2.set_positions(c)
    s2.nlist_allocate(params['rcut'], 1.2)
    e, f, v, _ = s2.compute('fp32')
    nb, _, _ = O.typed_nlist(c, H, t0, params['rcut'])
    E, Fo, W = O.energy_force_virial(params, var, c, H, t0, nbrs_nm=nb)
    assert abs(e[0] - E.item()) / len(t0) <= 1e-5
    assert np.abs(f[0] - Fo.numpy()).max() <= 1e-4 * max(1.0, np.abs(Fo.numpy()).max())
    assert np.abs(v[0] - W.numpy()).max() <= 1e-4 * np.abs(W.numpy()).max()

"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Tolerances (FP32 device arithmetic against the FP64 oracle):
  energy: |dE| / N <= 1e-5 eV/atom and |dE| / |E| <= 1e-5
  force:  max |dF| <= 1e-4 * max(1, max |F|) eV/A  (FP32 path); RMS relative <= 1e-5
  virial: max |dW| <= 1e-4 * max |W|
Neighbour lists: bit-exact index sets (type-local, ascending, padded with N).
"""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _system(device_model, coord, box, types, rc_skin, nrep=1):
    from deepmd_jax import _lib
    s = _lib.DeviceSystem(device_model, types, nrep)
    s.set_box(box)
    s.set_masses(np.ones(len(types)))
    s.set_positions(np.broadcast_to(coord, (nrep,) + coord.shape))
    s.nlist_allocate(rc_skin, 1.2)
    return s


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


@pytest.mark.parametrize('nmol,rc_skin', [(64, 6.2), (512, 7.0), (512, 6.3)])
def test_nlist_bit_exact(device_model, nmol, rc_skin):
    from deepmd_jax import synthetic as syn
    print('case nlist', nmol, rc_skin, flush=True)
    coord, box, types = syn.water_box(nmol, seed=5)
    coord = _f32(coord)
    ref, overflow, knbr = O.typed_nlist(coord, box, types, rc_skin)
    s = _system(device_model, coord, box, types, rc_skin)
    assert s.knbr == knbr
    N = len(types)
    off = np.concatenate([[0], np.cumsum(np.bincount(types))])
    for j in range(2):
        got = s.nlist_export(j)[0]
        for i in range(2):
            np.testing.assert_array_equal(got[off[i]:off[i + 1]], ref[i][j])
    assert not s.nlist_update()


@pytest.mark.parametrize('nmol', [64, 216])
def test_energy_force_virial(device_model, model0, nmol):
    from deepmd_jax import synthetic as syn
    print('case efv', nmol, flush=True)
    params, var = model0
    coord, box, types = syn.water_box(nmol, seed=1)
    coord = _f32(coord)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E, F, W = E.item(), F.numpy(), W.numpy()
    s = _system(device_model, coord, box, types, params['rcut'])
    e, f, v, ea = s.compute('fp32', want_atom=True)
    N = len(types)
    assert abs(e[0] - E) / N <= 1e-5 and abs(e[0] - E) / abs(E) <= 1e-5, (e[0], E)
    assert np.abs(f[0] - F).max() <= 1e-4 * max(1.0, np.abs(F).max()), np.abs(f[0] - F).max()
    assert np.sqrt(((f[0] - F) ** 2).mean() / (F ** 2).mean()) <= 1e-5
    assert np.abs(v[0] - W).max() <= 1e-4 * np.abs(W).max(), (v[0], W)
    assert abs(ea[0].sum() - e[0]) <= 1e-6 * abs(E)


def test_isolated_atoms_and_empty_rows(device_model, model0):
    """Atoms with no neighbour within rcut (empty rows of both types) beside a dense cluster: the
    isolated atoms' energies come from the fitting net on a zero descriptor, their forces are zero,
    and the cluster is unaffected."""
    from deepmd_jax import synthetic as syn
    print('case isolated', flush=True)
    params, var = model0
    cl, cbox, ctypes = syn.water_box(27, seed=3)
    L = 40.0
    iso = np.array([[20., 20., 20.], [28., 20., 20.], [20., 28., 20.], [20., 20., 28.],
                    [28., 28., 28.], [20., 28., 28.]])
    coord = _f32(np.concatenate([np.mod(cl, np.asarray(cbox).reshape(-1)[:3]), iso]))
    types = np.concatenate([ctypes, [0, 1, 0, 1, 1, 0]]).astype(np.int32)
    box = np.array([L, L, L])
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E, F, W = E.item(), F.numpy(), W.numpy()
    s = _system(device_model, coord, box, types, params['rcut'])
    e, f, v, ea = s.compute('fp32', want_atom=True)
    N = len(types)
    assert abs(e[0] - E) / N <= 1e-5 and abs(e[0] - E) / abs(E) <= 1e-5, (e[0], E)
    assert np.abs(f[0] - F).max() <= 1e-4 * max(1.0, np.abs(F).max()), np.abs(f[0] - F).max()
    assert np.abs(v[0] - W).max() <= 1e-4 * np.abs(W).max(), (v[0], W)
    # the isolated atoms: no force, and each one's energy is its type's zero-descriptor energy
    assert np.abs(f[0][len(ctypes):]).max() == 0.0
    ea_iso = ea[0][len(ctypes):]
    for t in (0, 1):
        sel = types[len(ctypes):] == t
        assert np.ptp(ea_iso[sel]) == 0.0


def test_replicas_and_skin_independent(device_model, model0):
    """Replicas of one system give identical results; the skin does not change E/F."""
    from deepmd_jax import synthetic as syn
    print('case replicas', flush=True)
    coord, box, types = syn.water_box(216, seed=2)
    s1 = _system(device_model, coord, box, types, 6.0)
    e1, f1, v1, _ = s1.compute()
    s3 = _system(device_model, coord, box, types, 7.0, nrep=3)
    e3, f3, v3, _ = s3.compute()
    for r in range(3):
        assert abs(e3[r] - e1[0]) <= 1e-6 * abs(e1[0])
        assert np.abs(f3[r] - f1[0]).max() <= 1e-5 * np.abs(f1).max()


def test_nve_against_oracle_trajectory(device_model, model0):
    """20 velocity-Verlet steps: device FP64 integrator + FP32 forces vs the oracle in FP64."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case nve', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(64, seed=0)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=0)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.nlist_allocate(6.2, 1.2)
    p = _lib.MDParams()
    p.routine, p.precision, p.dt = 0, 0, 0.5
    nsteps = 20
    tp, tv, err = s.md_run(p, nsteps)
    assert err == 0
    x, pm = coord.copy(), vel * m[:, None]
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    _, F, _ = O.energy_force_virial(params, var, x, box, types, lattice=lat)
    F = F.numpy()
    for k in range(nsteps):
        pm += 0.25 * F
        x += 0.5 * pm / m[:, None]
        _, F, _ = O.energy_force_virial(params, var, x, box, types, lattice=lat)
        F = F.numpy()
        pm += 0.25 * F
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5


def test_overflow_is_reported_and_state_restored(device_model, model0):
    from deepmd_jax import _lib, synthetic as syn
    print('case overflow', flush=True)
    coord, box, types = syn.water_box(216, seed=3)
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)) * 100.0)
    s.set_positions(coord)
    s.nlist_allocate(7.0, 1.2, knbr_override=[10, 10])
    x0 = s.get_positions()
    p = _lib.MDParams()
    p.routine, p.dt = 0, 0.5
    _, _, err = s.md_run(p, 3, record=False)
    assert err & 2
    np.testing.assert_array_equal(s.get_positions(), x0)


@pytest.mark.parametrize('prec', ['bf16', 'fp16'])
def test_half_precision_forms(device_model, model0, prec):
    """BF16 / FP16 embedding and fitting GEMMs (FP32 accumulation; FP32 switch, environment,
    contraction and descriptor): energy/atom and force RMSE against the FP64 oracle within
    1e-3 eV/atom and 1e-3 eV/A, virial within 1e-2 relative."""
    from deepmd_jax import synthetic as syn
    print('case prec', prec, flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=11)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E, F, W = E.item(), F.numpy(), W.numpy()
    s = _system(device_model, coord, box, types, params['rcut'])
    e, f, v, _ = s.compute(prec)
    N = len(types)
    rmse = np.sqrt(((f[0] - F) ** 2).mean())
    print(prec, 'dE/atom', abs(e[0] - E) / N, 'F rmse', rmse, 'W maxrel', np.abs(v[0] - W).max() / np.abs(W).max())
    assert abs(e[0] - E) / N <= 1e-3
    assert rmse <= 1e-3
    assert np.abs(v[0] - W).max() <= 1e-2 * np.abs(W).max()
    # the FP32 path of the same system is far tighter
    e32, f32, _, _ = s.compute('fp32')
    assert np.sqrt(((f32[0] - F) ** 2).mean()) < 0.1 * max(rmse, 1e-9) or np.sqrt(((f32[0] - F) ** 2).mean()) < 1e-6


def test_nvt_langevin_and_nose_hoover_run(device_model, model0):
    """Thermostatted runs stay finite; Nose-Hoover invariant is conserved to 1e-3 eV/atom over 100 steps."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    print('case nvt', flush=True)
    coord, box, types = syn.water_box(216, seed=12)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=2)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    for routine in (1, 2):
        s = _lib.DeviceSystem(device_model, types, 1)
        s.set_box(box)
        s.set_masses(m)
        s.set_positions(coord)
        s.set_velocities(vel)
        s.nlist_allocate(7.0, 1.2)
        p = _lib.MDParams()
        p.routine, p.dt, p.kT, p.gamma, p.tau, p.chain_length, p.seed = routine, 0.5, 300 * TEMP_UNIT_CONVERSION, 0.01, 50.0, 3, 7
        e0, _, _, _ = s.compute()
        inv0 = e0[0] + s.kinetic_energy()[0]
        tp, tv, err = s.md_run(p, 100)
        assert err == 0 and np.isfinite(tp).all()
        e1, _, _, _ = s.compute()
        th = s.thermostat_energy()[0]
        if routine == 1:
            inv1 = e1[0] + s.kinetic_energy()[0] + th
            assert abs(inv1 - inv0) / len(types) < 1e-3, (inv0, inv1)
        else:
            inv1 = e1[0] + s.kinetic_energy()[0] - th
            assert abs(inv1 - inv0) / len(types) < 2e-3, (inv0, inv1)


def test_pimd_step_against_oracle(device_model, model0):
    """Ring-polymer BAOAB (exact free ring + PILE-L) with the stochastic term switched off
    (gamma = 0): 4 beads x 64 H2O, 5 steps, device vs the FP64 oracle restatement."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.pimd import normal_mode_transform
    print('case pimd', flush=True)
    params, var = model0
    P, dt = 4, 0.25
    coord, box, types = syn.water_box(64, seed=3)
    rng = np.random.default_rng(3)
    x = coord[None] + rng.uniform(-0.05, 0.05, (P,) + coord.shape)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT = 300 * TEMP_UNIT_CONVERSION
    pm = rng.standard_normal(x.shape) * np.sqrt(m * kT * P)[None, :, None]
    freqs, C = normal_mode_transform(P, kT)
    s = _lib.DeviceSystem(device_model, types, P)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(x)
    s.set_velocities(pm / m[None, :, None])
    s.nlist_allocate(6.2, 1.2)
    s.pimd_setup(C, freqs, np.ones(P), np.zeros(P))
    p = _lib.MDParams()
    p.routine, p.dt = 3, dt
    tp, tv, err = s.md_run(p, 5)
    assert err == 0
    lat = O.lattice_candidate(np.diag(box), params['rcut'])

    def forces(xx):
        return np.stack([O.energy_force_virial(params, var, xx[b], box, types, lattice=lat)[1].numpy()
                         for b in range(P)])
    F = forces(x)
    xo, po = x.copy(), pm.copy()
    for k in range(5):
        xo, po, F = O.pimd_step(xo, po, m, forces, dt, C, freqs, np.ones(P), np.zeros(P), F, box)
    assert np.abs(tp[-1] - xo).max() < 1e-5
    assert np.abs(tv[-1] - po / m[None, :, None]).max() < 1e-5


def test_default_topology_model():
    """The reference's default widths (embed 32/32/64, fit 128x3, axis 12) run on the
    second kernel instantiation; FP32 parity with the FP64 oracle."""
    from deepmd_jax import _lib, synthetic as syn
    print('case default topology', flush=True)
    params, var = syn.init_model(1, embed_widths=(32, 32, 64), fit_widths=(128, 128, 128), axis=12)
    coord, box, types = syn.water_box(216, seed=14)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E, F, W = E.item(), F.numpy(), W.numpy()
    dm = _lib.DeviceModel(params, var)
    s = _system(dm, coord, box, types, params['rcut'])
    for prec, tol_e, tol_f in (('fp32', 1e-5, 1e-5), ('bf16', 1e-3, 1e-3)):
        e, f, v, _ = s.compute(prec)
        N = len(types)
        assert abs(e[0] - E) / N <= tol_e, (prec, e[0], E)
        assert np.sqrt(((f[0] - F) ** 2).mean()) <= tol_f * max(1.0, np.sqrt((F ** 2).mean())), prec


def test_forces_98k_box_against_fd(device_model, model0):
    """Full-size box (32,768 H2O, 98,304 atoms): device FP32 forces on 64 random atoms
    against FP64 central differences (h = 1e-4 A) of the oracle energy.
    Tolerance: max |dF| <= 1e-4 * max(1, max |F|)."""
    from deepmd_jax import _lib, synthetic as syn
    print('case fd 98k', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(32768, seed=0)
    coord = _f32(coord)
    s = _system(device_model, coord, box, types, params['rcut'] + 1.0)
    _, f, _, _ = s.compute('fp32')
    atoms = np.sort(np.random.default_rng(11).choice(len(types), 64, replace=False))
    fd = O.local_fd_forces(params, var, coord, box, types, atoms)
    err = np.abs(f[0][atoms] - fd).max()
    print('fd max err', err, 'max |F|', np.abs(fd).max(), flush=True)
    assert err <= 1e-4 * max(1.0, np.abs(fd).max())


def test_autograd_boundary(device_model, model0):
    """torch.autograd through the device energy: dE/dx = -F, and the energy change
    along a random displacement matches the gradient (first order)."""
    import torch
    from deepmd_jax import synthetic as syn
    from deepmd_jax.autodiff import DeviceEnergy, energy_and_force
    print('case autograd', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(64, seed=2)
    s = _system(device_model, coord, box, types, params['rcut'] + 0.2)
    e, f = energy_and_force(s, coord)
    _, f2, _, _ = s.compute('fp32')
    # two evaluations differ only by the order of the FP32 force atomics
    assert np.abs(f - f2[0]).max() <= 1e-5 * np.abs(f2[0]).max()
    d = np.random.default_rng(0).standard_normal(coord.shape) * 1e-3
    c = torch.as_tensor(coord + d)
    e2 = float(DeviceEnergy.apply(c, s, 'fp32'))
    de_lin = -(f * d).sum()
    assert abs((e2 - e) - de_lin) <= 0.05 * abs(de_lin) + 1e-4


def test_simulation_small_box_images(device_model, model0):
    """Simulation on a box below 2 rcut (32 H2O, L = 9.86 A): held as a 2x2x2 image
    supercell on the device; 20 NVE steps against the FP64 oracle's lattice-image VV, and
    the reported energy is the primary cell's."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    print('case simulation small box', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(32, seed=4, permute_sites=True)
    assert box[0] < 2 * params['rcut']
    vel = syn.maxwell_boltzmann(types, 300.0, seed=4)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=vel,
                         report_interval=10 ** 6, neighbor_skin=0.5, model=(DPModel(params), var, device_model))
        assert sim._sys.nimg == 8
        E0 = sim.getEnergy()
        traj = sim.run(20)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    Eo, F, _ = O.energy_force_virial(params, var, coord, box, types, lattice=lat)
    assert abs(E0 - Eo.item()) <= 1e-5 * len(types)
    x, pm, F = coord.copy(), vel * m[:, None], F.numpy()
    for k in range(20):
        pm += 0.25 * F
        x += 0.5 * pm / m[:, None]
        _, F, _ = O.energy_force_virial(params, var, x, box, types, lattice=lat)
        F = F.numpy()
        pm += 0.25 * F
    assert np.abs(traj['position'][-1] - x).max() < 1e-5
    assert np.abs(traj['velocity'][-1] - pm / m[:, None]).max() < 1e-5


def test_simulation_trajectory_frames(device_model, model0):
    """Simulation.run on a box that needs no images (216 H2O, L = 18.6 A): the device writes
    each chunk's frames (caller order, velocities = momenta / mass) straight into the
    returned trajectory; 25 NVE steps (three chunks of at most 10) against the FP64
    oracle's velocity Verlet, frame 0 the initial state."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    print('case simulation trajectory frames', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=6, permute_sites=True)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=6)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=vel,
                         report_interval=10 ** 6, neighbor_skin=1.0, model=(DPModel(params), var, device_model))
        assert sim._sys.nimg == 1
        sim._step_chunk_size = 10
        traj = sim.run(25)
    assert traj['position'].shape == (26, len(types), 3)
    np.testing.assert_allclose(traj['position'][0], coord, rtol=0, atol=1e-12)
    np.testing.assert_allclose(traj['velocity'][0], vel, rtol=0, atol=1e-12)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 1.0)
    x, pm = coord.copy(), vel * m[:, None]
    F = O.energy_force_virial(params, var, x, box, types, nbrs_nm=nb)[1].numpy()
    for k in range(25):
        pm += 0.25 * F
        x += 0.5 * pm / m[:, None]
        F = O.energy_force_virial(params, var, x, box, types, nbrs_nm=nb)[1].numpy()
        pm += 0.25 * F
        if k in (9, 19, 24):
            assert np.abs(traj['position'][k + 1] - x).max() < 1e-5, k
            assert np.abs(traj['velocity'][k + 1] - pm / m[:, None]).max() < 1e-5, k


@pytest.mark.parametrize('nmol', [32, 216])
def test_ase_energy_forces_stress(device_model, model0, nmol):
    """The ASE calculator's computation (deepmd_jax/ase_calc.py:79-143): stress = (1/V) dU/d(strain)
    in Voigt order against the FP64 oracle's strain derivative, within 1e-4 of its largest
    component; energy and forces at the FP32 tolerances.  32 H2O (L = 9.9 A) runs as a 2x2x2
    supercell, 216 H2O on the neighbour list."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.evaluator import Evaluator, energy_forces_stress
    print('case ase stress', nmol, flush=True)
    params, var = model0
    coord, box, types = syn.water_box(nmol, seed=15)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    S = -W.numpy() / np.prod(box)
    ref = np.array([S[0, 0], S[1, 1], S[2, 2], S[1, 2], S[0, 2], S[0, 1]])
    e, f, s = energy_forces_stress(Evaluator(params, var, model=device_model), coord, np.diag(box), types)
    N = len(types)
    assert abs(e - E.item()) / N <= 1e-5
    assert np.abs(f - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())
    assert np.abs(s - ref).max() <= 1e-4 * np.abs(ref).max(), (s, ref)


@pytest.mark.parametrize('prec', ['bf16x3', 'bf16-fit32'])
def test_mixed_precision_forms(device_model, model0, prec):
    """BF16 embedding GEMMs with split-BF16 ('bf16x3': hi*hi + hi*lo + lo*hi on BF16 MFMAs) or
    FP32 ('bf16-fit32') fitting GEMMs: within the 16-bit forms' tolerances against the FP64
    oracle, and the fitting's share of the error nearly gone.  With the embedding precision
    fixed (BF16), force(plain BF16 fitting) - force(FP32 fitting) is the BF16 fitting's error;
    the split fitting must leave under a twentieth of it (its products keep ~16 significand
    bits against BF16's 8), and the two mixed forms agree to the same degree."""
    from deepmd_jax import synthetic as syn
    print('case prec', prec, flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=11)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E, F, W = E.item(), F.numpy(), W.numpy()
    s = _system(device_model, coord, box, types, params['rcut'])
    N = len(types)
    e, f, v, _ = s.compute(prec)
    rmse = np.sqrt(((f[0] - F) ** 2).mean())
    print(prec, 'dE/atom', abs(e[0] - E) / N, 'F rmse', rmse, 'W maxrel', np.abs(v[0] - W).max() / np.abs(W).max())
    assert abs(e[0] - E) / N <= 1e-3
    assert rmse <= 1e-3
    assert np.abs(v[0] - W).max() <= 1e-2 * np.abs(W).max()
    eb, fb, _, _ = s.compute('bf16')
    ef, ff, _, _ = s.compute('bf16-fit32')
    fit_err = np.sqrt(((fb[0] - ff[0]) ** 2).mean())
    split_err = np.sqrt(((s.compute('bf16x3')[1][0] - ff[0]) ** 2).mean())
    print('BF16 fitting error', fit_err, 'split fitting error', split_err, 'energy', abs(eb[0] - ef[0]) / N)
    assert fit_err > 0
    assert split_err <= 0.05 * fit_err


def test_simulation_bf16x3_runs(device_model, model0):
    """Simulation with the split-BF16 form: NVE steps stay finite and conserve KE + PE to the
    16-bit forms' level over 100 steps of a 216-molecule box."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    import contextlib
    import io
    params, var = model0
    coord, box, types = syn.water_box(216, seed=3)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=4)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION

    def h(sim):
        return float(sim.getEnergy()) + 0.5 * float((m[:, None] * np.asarray(sim.getVelocity()) ** 2).sum())
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=vel,
                         report_interval=10 ** 9, neighbor_skin=1.0, precision='bf16x3',
                         model=(DPModel(params), var, device_model))
        h0 = h(sim)
        sim.run(100)
        h1 = h(sim)
    assert np.isfinite(sim.getPosition()).all()
    print('H change per atom', (h1 - h0) / len(types))
    assert abs(h1 - h0) / len(types) < 1e-3

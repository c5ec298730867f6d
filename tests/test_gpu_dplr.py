"""DPLR on the device (deepmd_jax/md.py:404-415) against the FP64 oracle.

216 H2O (L = 18.6 A), the seeded energy model plus a seeded Wannier-centroid model, charges
O +6, H +1, centroid -8, beta 0.4, the resolution-5 grid.  Tolerances (FP32 model arithmetic,
FP64 PPPM): centroids within 1e-5 A, long-range energy within 1e-5 relative, total energy
within 1e-5 eV/atom, forces within 1e-4 * max |F|, the virial diagonal within 1e-4 * max |W|
of central differences of the oracle's energy under axis strain (atoms, box and the centroid
model's inputs scaled); 10 velocity-Verlet steps with DPLR forces within 1e-5 (A, A/fs) of the
oracle's; NPT with DPLR conserves its invariant.
"""
import numpy as np
import pytest
import torch

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu

Q_ATOMS, Q_WC, BETA = [6.0, 1.0], [-8.0, 0.0], 0.4


def _setup(device_model, model0, nmol=216, seed=40):
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    wdm = _lib.DeviceModel(wp, wv)
    coord, box, types = syn.water_box(nmol, seed=seed)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(coord)
    s.nlist_allocate(params['rcut'] + 0.5, 1.2)
    M = O.p3m_grid(box, BETA, 5)
    s.set_dplr(wdm, Q_ATOMS, Q_WC, BETA, M)
    return s, (wp, wv), coord, box, types, m, M


def _oracle(model0, wc, coord, box, types, M):
    params, var = model0
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 1.0)
    return O.dplr_energy_force(params, var, wc[0], wc[1], coord, box, types, Q_ATOMS, Q_WC[:1], BETA, M,
                               nbrs_nm=nb, wc_nbrs_nm=nb)


def test_dplr_energy_forces_centroids(device_model, model0):
    print('case dplr eval', flush=True)
    s, wc, coord, box, types, m, M = _setup(device_model, model0)
    e, f, v, _ = s.compute()
    elr, cen = s.dplr_get()
    elr, cen = elr[0], cen[0]
    E, Elr, F, W = _oracle(model0, wc, coord, box, types, M)
    N = len(types)
    assert np.abs(cen - W).max() < 1e-5
    assert abs(elr - Elr) <= 1e-5 * abs(Elr), (elr, Elr)
    assert abs(e[0] - E) / N <= 1e-5, (e[0], E)
    assert np.abs(f[0] - F).max() <= 1e-4 * np.abs(F).max()
    # virial diagonal: -dE/de_a of the oracle's DPLR energy with atoms and box scaled along a
    h = 1e-5
    for a in range(3):
        sc = np.ones(3)
        sc[a] = 1 + h
        ep = _oracle(model0, wc, coord * sc, box * sc, types, M)[0]
        sc[a] = 1 - h
        em = _oracle(model0, wc, coord * sc, box * sc, types, M)[0]
        assert abs(v[0][a, a] + (ep - em) / (2 * h)) <= 1e-4 * np.abs(v[0]).max(), (a, v[0][a, a], (ep - em) / (2 * h))
    # removing the term restores the short-range model
    s.set_dplr(None, None, None, 0, None)
    e2, f2, v2, _ = s.compute()
    assert np.isfinite(v2).all() and abs(e2[0] - (E - Elr)) / N <= 1e-5


def test_dplr_nve_against_oracle(device_model, model0):
    from deepmd_jax import _lib, synthetic as syn
    print('case dplr nve', flush=True)
    s, wc, coord, box, types, m, M = _setup(device_model, model0, seed=41)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=41)
    s.set_velocities(vel)
    p = _lib.MDParams()
    p.routine, p.dt = 0, 0.5
    nsteps = 10
    tp, tv, err = s.md_run(p, nsteps)
    assert err == 0
    x, pm = coord.copy(), vel * m[:, None]
    F = _oracle(model0, wc, x, box, types, M)[2]
    for _ in range(nsteps):
        pm = pm + 0.25 * F
        x = x + 0.5 * pm / m[:, None]
        F = _oracle(model0, wc, x, box, types, M)[2]
        pm = pm + 0.25 * F
    assert np.abs(tp[-1, 0] - x).max() < 1e-5
    assert np.abs(tv[-1, 0] - pm / m[:, None]).max() < 1e-5


def test_dplr_model_pkl_simulation_and_evaluate(model0, tmp_path):
    """A 'dplr' model.pkl (deepmd_jax/train.py:305-313) through Simulation and evaluate():
    NVE conserves KE + PE (with the long-range term) to 2e-5 eV/atom over 40 steps of 0.05 fs
    while the long-range energy moves by more than 1 eV (the seeded centroid model with these
    charges drives fast exchanges: at 0.25 fs the integration error alone is 3e-4 eV/atom, for
    the oracle too), and evaluate() returns the oracle's DPLR energy and forces."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.train import evaluate
    from deepmd_jax.utils import save_model
    print('case dplr model.pkl', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    dparams = dict(params)
    dparams.update({'type': 'dplr', 'dplr_wannier_model_and_variables': (DPModel(wp), {'params': wv['params']}),
                    'dplr_q_atoms': Q_ATOMS, 'dplr_q_wc': Q_WC[:1], 'dplr_beta': BETA, 'dplr_resolution': 5})
    path = str(tmp_path / 'dplr.pkl')
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(dparams), var)
    coord, box, types = syn.water_box(216, seed=42)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(path, box, types, syn.MASSES, 'NVE', 0.05, coord, initial_velocity=
                         syn.maxwell_boltzmann(types, 300.0, seed=42), report_interval=10)
        elr0 = sim._sys.dev.dplr_get()[0][0] if sim.getEnergy() else None
        sim.run(40)
    elr1 = sim._sys.dev.dplr_get()[0][0]
    inv = np.array(sim.log)[:, 4]
    assert abs(elr1 - elr0) > 1.0, (elr0, elr1)
    assert np.abs(inv - inv[0]).max() / len(types) < 2e-5, inv
    with contextlib.redirect_stdout(io.StringIO()):
        out = evaluate(path, coord[None], box[None], types)
    M = O.p3m_grid(box, BETA, 5)
    E, _, F, _ = _oracle(model0, (wp, wv), coord, box, types, M)
    assert abs(out['energy'][0] - E) / len(types) <= 1e-5
    assert np.abs(out['force'][0] - F).max() <= 1e-4 * np.abs(F).max()


def test_dplr_npt_invariant(model0, tmp_path):
    """NPT (MTK, isotropic, 1 kbar) with a DPLR model: the barostat reads the long-range virial;
    the NPT invariant holds to 5e-5 eV/atom over 60 steps of 0.05 fs while the box moves."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.utils import save_model
    print('case dplr npt', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    dparams = dict(params)
    dparams.update({'type': 'dplr', 'dplr_wannier_model_and_variables': (DPModel(wp), {'params': wv['params']}),
                    'dplr_q_atoms': Q_ATOMS, 'dplr_q_wc': Q_WC[:1], 'dplr_beta': BETA, 'dplr_resolution': 5})
    path = str(tmp_path / 'dplr.pkl')
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(dparams), var)
    coord, box, types = syn.water_box(216, seed=43)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(path, box, types, syn.MASSES, 'NPT', 0.05, coord, temperature=300.0, pressure=1000.0,
                         tau_p=50.0, tau_t=20.0, report_interval=10, initial_velocity=syn.maxwell_boltzmann(types, 300.0, seed=43))
        traj = sim.run(60)
    log = np.array(sim.log)
    inv = log[:, -2]
    assert np.isfinite(log).all()
    assert np.abs(traj['box'][-1] - box).max() > 1e-6
    assert np.abs(inv - inv[0]).max() / len(types) < 5e-5, inv


@pytest.mark.parametrize('kind', ['atomic', 'atomic_t2'])
def test_evaluate_atomic_models(tmp_path, kind):
    """evaluate() of 'atomic' and 'atomic_t2' model.pkl files (deepmd_jax/train.py:765-788) against
    the oracle: 2 frames of 64 H2O (small box: image supercell) and 216 H2O, outputs within 1e-5
    of the largest component."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.train import evaluate
    from deepmd_jax.utils import save_model
    print('case evaluate', kind, flush=True)
    wp, wv = syn.init_wannier_model(6, kind=kind)
    path = str(tmp_path / ('%s.pkl' % kind))
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(wp), wv)
    for nmol in (64, 216):
        frames = [syn.water_box(nmol, seed=s) for s in (50, 51)]
        coord = np.stack([f[0] for f in frames])
        box = np.stack([f[1] for f in frames])
        types = frames[0][2]
        with contextlib.redirect_stdout(io.StringIO()):
            out = evaluate(path, coord, box, types, batch_size=2)
        key = wp['atomic_data_prefix']
        got = out[key]
        ncomp = 3 if kind == 'atomic' else 9
        assert got.shape == (2, (types == 0).sum(), ncomp)
        for f in range(2):
            lat = O.lattice_candidate(np.diag(box[f]), wp['rcut'])
            want = O.dp_atomic(wp, wv, torch.as_tensor(coord[f]), box[f], types, lattice=lat).numpy()
            assert np.abs(got[f] - want).max() <= 1e-5 * np.abs(want).max(), (nmol, f)


def test_dplr_segments_with_own_boxes(device_model, model0):
    """Two frames as two segments of one system, the second in a box 2 % larger (same grid):
    each segment's energy, forces, long-range energy and centroids against the oracle."""
    from deepmd_jax import _lib, synthetic as syn
    print('case dplr segments', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    wdm = _lib.DeviceModel(wp, wv)
    c0, box, types = syn.water_box(128, seed=60)
    c1 = syn.water_box(128, seed=61)[0] * 1.02
    boxes = np.stack([box, box * 1.02])
    M = O.p3m_grid(box, BETA, 5)
    s = _lib.DeviceSystem(device_model, types, 2)
    s.set_boxes(boxes)
    s.set_masses(np.ones(len(types)))
    s.set_positions(np.stack([c0, c1]))
    s.nlist_allocate(params['rcut'] + 0.5, 1.2)
    s.set_dplr(wdm, Q_ATOMS, Q_WC, BETA, M)
    e, f, v, _ = s.compute()
    elr, cen = s.dplr_get()
    N = len(types)
    for k, (c, b) in enumerate(((c0, box), (c1, box * 1.02))):
        E, Elr, F, Wc = _oracle(model0, (wp, wv), c, b, types, M)
        assert abs(elr[k] - Elr) <= 1e-5 * abs(Elr), (k, elr[k], Elr)
        assert np.abs(cen[k] - Wc).max() < 1e-5
        assert abs(e[k] - E) / N <= 1e-5, (k, e[k], E)
        assert np.abs(f[k] - F).max() <= 1e-4 * np.abs(F).max()


def test_dplr_pimd_step_against_oracle(device_model, model0):
    """Ring-polymer BAOAB without noise (4 beads x 64 H2O) with the DPLR term on every bead,
    3 steps, against the oracle's ring-polymer step with per-bead DPLR forces."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.pimd import normal_mode_transform
    print('case dplr pimd', flush=True)
    params, var = model0
    P, dt = 4, 0.1
    wp, wv = syn.init_wannier_model(4)
    wdm = _lib.DeviceModel(wp, wv)
    coord, box, types = syn.water_box(64, seed=62)
    rng = np.random.default_rng(62)
    x = coord[None] + rng.uniform(-0.03, 0.03, (P,) + coord.shape)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT = 300 * TEMP_UNIT_CONVERSION
    pm = rng.standard_normal(x.shape) * np.sqrt(m * kT * P)[None, :, None]
    freqs, C = normal_mode_transform(P, kT)
    M = O.p3m_grid(box, BETA, 5)
    s = _lib.DeviceSystem(device_model, types, P)
    s.set_box(box)
    s.set_masses(m)
    s.set_positions(x)
    s.set_velocities(pm / m[None, :, None])
    s.nlist_allocate(params['rcut'] + 0.2, 1.2)
    s.set_dplr(wdm, Q_ATOMS, Q_WC, BETA, M)
    s.pimd_setup(C, freqs, np.ones(P), np.zeros(P))
    p = _lib.MDParams()
    p.routine, p.dt = 3, dt
    tp, tv, err = s.md_run(p, 3)
    assert err == 0

    def forces(xx):
        return np.stack([_oracle(model0, (wp, wv), xx[b], box, types, M)[2] for b in range(P)])
    F = forces(x)
    xo, po = x.copy(), pm.copy()
    for _ in range(3):
        xo, po, F = O.pimd_step(xo, po, m, forces, dt, C, freqs, np.ones(P), np.zeros(P), F, box)
    assert np.abs(tp[-1] - xo).max() < 1e-5
    assert np.abs(tv[-1] - po / m[None, :, None]).max() < 1e-5


def test_dplr_evaluate_batches_and_pimd_simulation(model0, tmp_path):
    """evaluate() of a 'dplr' model.pkl with batch_size 3 (frames as segments) equals the oracle
    frame by frame; Simulation with 4 beads and DPLR runs and keeps its path-integral invariant."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.train import evaluate
    from deepmd_jax.utils import save_model
    print('case dplr batches', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    dparams = dict(params)
    dparams.update({'type': 'dplr', 'dplr_wannier_model_and_variables': (DPModel(wp), {'params': wv['params']}),
                    'dplr_q_atoms': Q_ATOMS, 'dplr_q_wc': Q_WC[:1], 'dplr_beta': BETA, 'dplr_resolution': 5})
    path = str(tmp_path / 'dplr.pkl')
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(dparams), var)
    frames = [syn.water_box(64, seed=s) for s in (63, 64, 65)]
    coord = np.stack([f[0] for f in frames])
    box = frames[0][1]
    types = frames[0][2]
    with contextlib.redirect_stdout(io.StringIO()):
        out = evaluate(path, coord, np.stack([box] * 3), types, batch_size=3)
    M = O.p3m_grid(box, BETA, 5)
    for k in range(3):
        E, _, F, _ = _oracle(model0, (wp, wv), coord[k], box, types, M)
        assert abs(out['energy'][k] - E) / len(types) <= 1e-5, k
        assert np.abs(out['force'][k] - F).max() <= 1e-4 * np.abs(F).max()
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(path, box, types, syn.MASSES, 'NVT', 0.05, coord[0], temperature=300.0, n_bead=4,
                         pimd_print_invariant=True, report_interval=10, tau_t=100.0)
        sim.run(40)
    log = np.array(sim.log)
    inv = log[:, 4]
    assert np.isfinite(log).all()
    assert np.abs(inv - inv[0]).max() / len(types) < 1e-4, inv


def test_dplr_small_box_on_images(model0):
    """32 H2O (L = 9.86 A < 2 rcut): the short-range and centroid models see the 2x2x2 supercell
    of images, the PPPM term the cell (the reference's lattice path, deepmd_jax/utils.py:66-78,
    with PPPM on the cell).  Energy, forces and the virial diagonal against the oracle's lattice
    form; NVE on the images conserves KE + PE."""
    import contextlib
    import io
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.evaluator import Evaluator
    print('case dplr images', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    coord, box, types = syn.water_box(32, seed=66)
    assert box.min() < 2 * params['rcut']
    M = O.p3m_grid(box, BETA, 5)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])

    def oracle(c, b):
        la = O.lattice_candidate(np.diag(b), params['rcut'])
        return O.dplr_energy_force(params, var, wp, wv, c, b, types, Q_ATOMS, Q_WC[:1], BETA, M, lattice=la)
    ev = Evaluator(params, var)
    ev.dplr = (_lib.DeviceModel(wp, wv), np.array(Q_ATOMS), np.array(Q_WC), BETA, 5.0)
    e, f, v = ev.evaluate_batch(coord[None], box, types)
    E, Elr, F, _ = O.dplr_energy_force(params, var, wp, wv, coord, box, types, Q_ATOMS, Q_WC[:1], BETA, M,
                                       lattice=lat)
    N = len(types)
    assert abs(e[0] - E) / N <= 1e-5, (e[0], E)
    assert np.abs(f[0] - F).max() <= 1e-4 * np.abs(F).max()
    h = 1e-5
    for a in range(3):
        sc = np.ones(3)
        sc[a] = 1 + h
        ep = oracle(coord * sc, box * sc)[0]
        sc[a] = 1 - h
        em = oracle(coord * sc, box * sc)[0]
        assert abs(v[0][a, a] + (ep - em) / (2 * h)) <= 1e-4 * np.abs(v[0]).max(), (a, v[0][a, a], (ep - em) / (2 * h))
    # NVE with the images moving together
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.utils import save_model
    import tempfile
    import os
    dparams = dict(params)
    dparams.update({'type': 'dplr', 'dplr_wannier_model_and_variables': (DPModel(wp), {'params': wv['params']}),
                    'dplr_q_atoms': Q_ATOMS, 'dplr_q_wc': Q_WC[:1], 'dplr_beta': BETA, 'dplr_resolution': 5})
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, 'dplr.pkl')
        with contextlib.redirect_stdout(io.StringIO()):
            save_model(path, DPModel(dparams), var)
            sim = Simulation(path, box, types, syn.MASSES, 'NVE', 0.05, coord, report_interval=10,
                             initial_velocity=syn.maxwell_boltzmann(types, 300.0, seed=66))
            sim.run(40)
    inv = np.array(sim.log)[:, 4]
    assert np.isfinite(inv).all()
    assert np.abs(inv - inv[0]).max() / N < 2e-5, inv


def test_dplr_path_integral_npt(model0, tmp_path):
    """Path-integral NPT (4 beads, Langevin box) with a DPLR model: every bead's PPPM follows the
    box the device moves; the path-integral invariant holds to 2e-3 eV/atom over 40 steps of 0.1 fs
    while the box moves."""
    import contextlib
    import io
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.utils import save_model
    print('case dplr pimd npt', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    dparams = dict(params)
    dparams.update({'type': 'dplr', 'dplr_wannier_model_and_variables': (DPModel(wp), {'params': wv['params']}),
                    'dplr_q_atoms': Q_ATOMS, 'dplr_q_wc': Q_WC[:1], 'dplr_beta': BETA, 'dplr_resolution': 5})
    path = str(tmp_path / 'dplr.pkl')
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(dparams), var)
    coord, box, types = syn.water_box(128, seed=96)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(path, box, types, syn.MASSES, 'NPT', 0.1, coord, temperature=300.0, pressure=1.0e4,
                         n_bead=4, tau_p=50.0, pimd_print_invariant=True, report_interval=10, seed=7)
        traj = sim.run(40)
    keys = ['Step', 'Temperature', 'KE', 'PE', 'Pressure', 'box_x', 'box_y', 'box_z', 'Invariant']
    log = np.array(sim.log)
    inv = log[:, keys.index('Invariant')]
    assert np.isfinite(log).all()
    assert np.abs(traj['box'][-1] - box).max() > 1e-6
    assert np.abs(inv - inv[0]).max() / len(types) < 2e-3, inv

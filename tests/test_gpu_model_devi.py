"""Simulation(model_deviation_paths=...) on the device (deepmd_jax/md.py:299-302,559-567).

Two seeded models of different weights: the Model_Devi column at the last report equals
max over atoms and components of std([F_0, F_1], axis=0) = max |F_0 - F_1| / 2 from the FP64
oracle at the reported positions, within 1e-4 * max(1, max |F|).
"""
import contextlib
import io

import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def test_model_deviation_column(model_pkl, model0, tmp_path):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.utils import save_model
    print('case model deviation', flush=True)
    params1, var1 = syn.init_model(1)
    pkl1 = str(tmp_path / 'model1.pkl')
    save_model(pkl1, DPModel(params1), var1)
    coord, box, types = syn.water_box(64, seed=30)
    with contextlib.redirect_stdout(io.StringIO()) as out:
        sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVT', 0.5, coord, temperature=300.0, report_interval=10,
                         model_deviation_paths=[model_pkl, pkl1], seed=2)
        sim.run(20)
    assert 'Model_Devi' in out.getvalue()
    dev = sim.log[-1][-2]
    x = sim.getPosition()
    F0 = O.energy_force_virial(model0[0], model0[1], x, box, types, lattice='auto')[1].numpy()
    F1 = O.energy_force_virial(params1, var1, x, box, types, lattice='auto')[1].numpy()
    want = np.abs(F0 - F1).max() / 2
    assert abs(dev - want) <= 1e-4 * max(1.0, np.abs(F0).max(), np.abs(F1).max()), (dev, want)
    with pytest.raises(NotImplementedError):
        Simulation(model_pkl, box, types, syn.MASSES, 'NVT', 0.5, coord, temperature=300.0, n_bead=2,
                   model_deviation_paths=[pkl1])


def test_model_deviation_with_dplr_models(model_pkl, model0, tmp_path):
    """A DPLR model and the plain model of the same short-range weights as the deviation models:
    the column is max |F_dplr - F_sr| / 2 of the FP64 oracle at the reported positions, which
    only holds when the DPLR deviation model carries its long-range term."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.utils import save_model
    print('case model devi dplr', flush=True)
    params, var = model0
    wp, wv = syn.init_wannier_model(4)
    qa, qw, beta = [6.0, 1.0], [-8.0], 0.4
    dparams = dict(params)
    dparams.update({'type': 'dplr', 'dplr_wannier_model_and_variables': (DPModel(wp), {'params': wv['params']}),
                    'dplr_q_atoms': qa, 'dplr_q_wc': qw, 'dplr_beta': beta, 'dplr_resolution': 5})
    path = str(tmp_path / 'dplr.pkl')
    with contextlib.redirect_stdout(io.StringIO()):
        save_model(path, DPModel(dparams), var)
    coord, box, types = syn.water_box(64, seed=80)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVE', 0.1, coord, report_interval=5,
                         model_deviation_paths=[path, model_pkl],
                         initial_velocity=syn.maxwell_boltzmann(types, 300.0, seed=80))
        sim.run(10)
    dev = sim.log[-1][-2]
    x = sim.getPosition()
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    M = O.p3m_grid(box, beta, 5)
    _, _, Fd, _ = O.dplr_energy_force(params, var, wp, wv, x, box, types, qa, qw, beta, M, lattice=lat)
    Fs = O.energy_force_virial(params, var, x, box, types, lattice=lat)[1].numpy()
    want = np.abs(Fd - Fs).max() / 2
    assert want > 0.1
    assert abs(dev - want) <= 1e-4 * max(1.0, np.abs(Fd).max()), (dev, want)

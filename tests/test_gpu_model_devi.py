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

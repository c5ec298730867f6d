import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'deepmd-jax_amd'))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels)')


@pytest.fixture(scope='session')
def water64():
    from deepmd_jax import synthetic as syn
    return syn.water_box(64, seed=0)


@pytest.fixture(scope='session')
def model0():
    from deepmd_jax import synthetic as syn
    return syn.init_model(0)


@pytest.fixture(scope='session')
def model_pkl(tmp_path_factory, model0):
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import save_model
    path = str(tmp_path_factory.mktemp('model') / 'model.pkl')
    save_model(path, DPModel(model0[0]), model0[1])
    return path


@pytest.fixture(scope='session')
def device_model(model0):
    from deepmd_jax import _lib
    return _lib.DeviceModel(model0[0], model0[1])


@pytest.fixture(scope='session')
def model0c(model0):
    """The seeded model after compress_model (1,024 intervals, rmin 0.6 A: the defaults of
    deepmd_jax/train.py:49-50)."""
    import copy
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import compress_model
    params, var = copy.deepcopy(model0)
    m, var = compress_model(DPModel(params), var, 1024, 0.6)
    return m.params, var


@pytest.fixture(scope='session')
def device_model_c(model0c):
    from deepmd_jax import _lib
    return _lib.DeviceModel(model0c[0], model0c[1])

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

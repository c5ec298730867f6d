"""model.pkl round trip and the restricted unpickler (deepmd_jax/utils.py:225-237)."""
import io
import pickle

import numpy as np
import pytest


def test_round_trip(tmp_path, model0):
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import load_model, save_model
    params, var = model0
    p = str(tmp_path / 'm.pkl')
    save_model(p, DPModel(params), var)
    m, v = load_model(p)
    assert m.params['axis'] == 16 and m.params['embed_widths'] == [25, 50, 100]
    for k in range(4):
        np.testing.assert_array_equal(v['params']['embedding_net_%d' % k]['Dense_1']['kernel'],
                                      var['params']['embedding_net_%d' % k]['Dense_1']['kernel'])
    np.testing.assert_array_equal(v['params']['fitting_net_1']['dt2'], var['params']['fitting_net_1']['dt2'])


class _Evil:
    def __reduce__(self):
        import os
        return (os.system, ('echo pwned',))


def test_restricted_unpickler_refuses_arbitrary_globals(tmp_path):
    from deepmd_jax.utils import load_model
    p = tmp_path / 'bad.pkl'
    p.write_bytes(pickle.dumps({'model': _Evil(), 'variables': {}}))
    with pytest.raises(pickle.UnpicklingError):
        load_model(str(p))


def test_flax_style_module_state_is_accepted(tmp_path, model0):
    """A pickled module whose state is a dataclass dict with extra flax fields."""
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import load_model
    params, var = model0
    m = DPModel(params)
    blob = pickle.dumps({'model': m, 'variables': var})
    # emulate extra state keys a flax Module carries
    m2 = DPModel.__new__(DPModel)
    m2.__setstate__({'params': params, 'parent': None, 'name': None})
    assert m2.params['rcut'] == 6.0
    p = tmp_path / 'ok.pkl'
    p.write_bytes(blob)
    mm, vv = load_model(str(p))
    assert mm.params['ntypes'] == 2

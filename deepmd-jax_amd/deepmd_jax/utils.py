"""model.pkl I/O (deepmd_jax/utils.py:225-237) with a restricted unpickler.

A deepmd-jax model.pkl is a pickle of {'model': DPModel(params), 'variables':
{'params': {...}}}.  Loading resolves only an allowlist of globals: numpy array
reconstruction, this package's DPModel, and the jax / flax names such a file
refers to, which are mapped to plain-numpy stand-ins.  Anything else is refused.
"""
import pickle

import numpy as np

from .dpmodel import DPModel


def _jax_reconstruct_array(fun, args, arr_state, aval_state):
    arr = fun(*args)
    if arr_state is not None:
        arr.__setstate__(arr_state)
    return np.asarray(arr)


class _Stub:
    """Stand-in for flax/jax objects embedded in a pickled module's state."""

    def __init__(self, *a, **k):
        pass

    def __setstate__(self, state):
        self.state = state


_ALLOWED = {
    ('numpy', 'ndarray'), ('numpy', 'dtype'), ('numpy.core.multiarray', '_reconstruct'),
    ('numpy._core.multiarray', '_reconstruct'), ('numpy.core.multiarray', 'scalar'),
    ('numpy._core.multiarray', 'scalar'), ('builtins', 'slice'), ('builtins', 'range'),
    ('collections', 'OrderedDict'), ('copyreg', '_reconstructor'), ('builtins', 'object'),
}


class _Unpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _ALLOWED:
            return super().find_class(module, name)
        if module in ('deepmd_jax.dpmodel',) and name == 'DPModel':
            return DPModel
        if module.startswith('jax') and name == '_reconstruct_array':
            return _jax_reconstruct_array
        if module.startswith('flax') or module.startswith('jax') or module.startswith('deepmd_jax'):
            return _Stub
        raise pickle.UnpicklingError('global %s.%s is not allowed in a model file' % (module, name))


def _to_numpy(tree):
    if isinstance(tree, dict):
        return {k: _to_numpy(v) for k, v in tree.items()}
    if isinstance(tree, (list, tuple)):
        return type(tree)(_to_numpy(v) for v in tree)
    if hasattr(tree, '__array__') and not isinstance(tree, np.ndarray):
        return np.asarray(tree)
    return tree


def load_model(path, replicate=True):
    """Return (DPModel, variables) with variables as numpy arrays."""
    with open(path, 'rb') as f:
        m = _Unpickler(f).load()
    print("# Model loaded from '%s'." % path)
    model = m['model']
    if not isinstance(model, DPModel):
        raise ValueError('model.pkl does not hold a DPModel')
    return model, _to_numpy(m['variables'])


def save_model(path, model, variables):
    with open(path, 'wb') as f:
        pickle.dump({'model': model, 'variables': _to_numpy(variables)}, f)
    print("# Model saved to '%s'." % path)

"""DPModel: the model object stored in model.pkl (deepmd_jax/dpmodel.py:8-9).

Here it is a parameter holder; evaluation runs on the GPU engine
(`deepmd_jax._lib.DeviceModel`).  `energy_and_force` keeps the reference signature
(deepmd_jax/dpmodel.py:106-108) for orthorhombic frames.
"""
import numpy as np


class DPModel:
    def __init__(self, params=None, **kwargs):
        self.params = dict(params) if params is not None else {}

    def __setstate__(self, state):
        # accept both this class's own state and the dataclass state of a pickled
        # flax module, which carries the same `params` field
        if isinstance(state, tuple):
            state = next((s for s in state if isinstance(s, dict) and 'params' in s), {})
        self.params = dict(state.get('params', {}))

    def __getstate__(self):
        return {'params': self.params}

    def __repr__(self):
        return 'DPModel(type=%r, rcut=%r, embed_widths=%r, fit_widths=%r, axis=%r)' % (
            self.params.get('type'), self.params.get('rcut'), self.params.get('embed_widths'),
            self.params.get('fit_widths'), self.params.get('axis'))

    def energy_and_force(self, variables, coord_N3, box_33, static_args, nbrs_nm=None, precision='fp32'):
        """(energy, force) of one frame; static_args['type_idx'] gives the atom types."""
        from .evaluator import Evaluator
        ev = Evaluator(self.params, variables)
        box = np.asarray(box_33, np.float64)
        box3 = np.diag(box) if box.shape == (3, 3) else box.reshape(3)
        e, f, _ = ev.evaluate_frame(np.asarray(coord_N3, np.float64), box3,
                                    np.asarray(static_args['type_idx']), precision=precision)
        return e, f

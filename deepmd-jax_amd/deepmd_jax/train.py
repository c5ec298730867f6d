"""`evaluate()` with the signature and return value of deepmd_jax/train.py:703-788.

Training (deepmd_jax/train.py:15-520) is not part of this engine.
"""
import numpy as np

from .evaluator import Evaluator, canonical_cell
from .utils import load_model


def _boxes(box, nframes):
    """(F, 3, 3) cells of the frames' boxes: scalars, (3,) lengths, (9,) or (3,3) cells."""
    box = np.asarray(box, np.float64)
    if box.ndim == 1 or (box.ndim == 2 and box.shape[1] in (1, 3, 9)) or (box.ndim == 3 and box.shape[1:] == (3, 3)):
        return np.stack([canonical_cell(b) for b in box])
    raise ValueError


def evaluate(model_path, coord, box, type_idx, batch_size=1, precision='fp32'):
    """Evaluate a trained model on frames (no ground truth needed).

    coord (F, N, 3); box (F,), (F,1), (F,3), (F,9) or (F,3,3); type_idx (N,).
    Returns {'energy': (F,), 'force': (F, N, 3)} in the caller's atom order.
    precision: 'fp32' (default), 'bf16' or 'fp16' embedding/fitting GEMMs, or the mixed forms
        'bf16x3' (BF16 embedding GEMMs, fitting GEMMs as three BF16 products of split operands)
        and 'bf16-fit32' (BF16 embedding, FP32 fitting GEMMs).
    """
    try:
        coord = np.asarray(coord)
        type_idx = np.asarray(type_idx)
        assert coord.ndim == 3 and coord.shape[2] == 3
        assert type_idx.ndim == 1 and np.asarray(box).ndim in [1, 2, 3]
        assert coord.shape[1] == type_idx.shape[0]
        assert coord.shape[0] == np.asarray(box).shape[0]
        boxes = _boxes(box, coord.shape[0])
    except (AssertionError, ValueError):
        raise ValueError('Input shapes are incorrect: \n' +
                         'coord: (n_frames, n_atoms, 3) \n' +
                         'box: (n_frames) + (,) or (1,) or (3,) or (9), or (3,3) \n' +
                         'type_idx (n_atoms).')
    model, variables = load_model(model_path, replicate=False)
    params = model.params
    kind = params.get('type', 'energy')
    if kind not in ('energy', 'dplr', 'atomic', 'atomic_t2'):
        raise NotImplementedError('evaluate() supports energy, DPLR, atomic and atomic_t2 models')
    ct = params.get('chemical_types')
    if ct is not None:
        unknown = set(type_idx.tolist()) - set(ct)
        if unknown:
            raise ValueError('Atomic numbers %s in type_idx are not in model.params["chemical_types"]=%s'
                             % (sorted(unknown), ct))
        z2i = {z: i for i, z in enumerate(ct)}
        type_idx = np.array([z2i[z] for z in type_idx.tolist()])
    if type_idx.min() < 0 or type_idx.max() >= params['ntypes']:
        raise ValueError('type index out of trained scope of model (ntypes=%d)' % params['ntypes'])
    ev = Evaluator(params, variables)
    F = coord.shape[0]
    if kind in ('atomic', 'atomic_t2'):
        # atomic outputs per frame (deepmd_jax/train.py:765-788): (F, M, 3) or (F, M, 9)
        key = params.get('atomic_data_prefix') or ('atomic_dipole' if kind == 'atomic' else 'atomic_polarizability')
        outs = []
        for i, j in _batches(ev, boxes, batch_size):
            outs.append(ev.evaluate_atomic_frames(coord[i:j], boxes[i:j], type_idx))
        return {key: np.concatenate(outs, 0)}
    energy = np.zeros(F)
    force = np.zeros(coord.shape)
    for i, j in _batches(ev, boxes, batch_size):
        e, f, _ = ev.evaluate_frames(coord[i:j], boxes[i:j], type_idx, precision)
        energy[i:j] = e
        force[i:j] = f
    return {'energy': energy, 'force': force}


def _batches(ev, boxes, batch_size):
    """Runs of at most batch_size consecutive frames with one batch key (the same images per
    axis and DPLR grid); each frame keeps its own box."""
    keys = [ev.batch_key(b) for b in boxes]
    i, F = 0, len(boxes)
    while i < F:
        j = i + 1
        while j < F and j - i < batch_size and keys[j] == keys[i]:
            j += 1
        yield i, j
        i = j

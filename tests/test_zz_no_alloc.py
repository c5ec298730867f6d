"""dpmd_compute_noalloc through the C ABI against dpmd_compute and the FP64 oracle: two
calls in a row on one GPU (a 512-molecule box, then a smaller 64-molecule box, so the
second call sees the first call's leftovers in the shared scratch array), FP32 and BF16,
and the size limit rejected with DPMD_E_ARG."""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _system(device_model, nmol, seed):
    from deepmd_jax import _lib, synthetic as syn
    coord, box, types = syn.water_box(nmol, seed=seed)
    coord = np.asarray(coord, np.float32).astype(np.float64)
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)))
    s.set_positions(coord)
    s.nlist_allocate(6.0 + (1.0 if nmol > 64 else 0.2), 1.2)
    return s, coord, box, types


def test_noalloc_matches_compute_two_calls(device_model, model0):
    params, var = model0
    for nmol, seed in ((512, 3), (64, 4)):
        print('case noalloc', nmol, flush=True)
        s, coord, box, types = _system(device_model, nmol, seed)
        for prec in ('fp32', 'bf16'):
            e0, f0, v0, _ = s.compute(prec)
            e1, f1, v1 = s.compute_noalloc(prec)
            fs = max(1.0, np.abs(f0).max())
            assert abs(e1[0] - e0[0]) <= 1e-12 * abs(e0[0]), (prec, e0, e1)
            assert np.abs(f1 - f0).max() <= 1e-5 * fs, prec
            assert np.abs(v1 - v0).max() <= 1e-9 * np.abs(v0).max(), prec
        E, F, W = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
        e1, f1, v1 = s.compute_noalloc('fp32')
        N = len(types)
        assert abs(e1[0] - E.item()) / N < 1e-5
        assert np.abs(f1[0] - F.numpy()).max() < 1e-4 * max(1.0, np.abs(F.numpy()).max())
        assert np.abs(v1[0] - W.numpy()).max() < 1e-4 * np.abs(W.numpy()).max()


def test_noalloc_rejects_beyond_scratch(device_model):
    from deepmd_jax import _lib, synthetic as syn
    print('case noalloc limit', flush=True)
    coord, box, types = syn.water_box(4096, seed=1)
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)))
    s.set_positions(coord)
    s.nlist_allocate(7.0, 1.2)
    with pytest.raises(ValueError, match='scratch'):
        s.compute_noalloc('fp32')
    e0, _, _, _ = s.compute('fp32')    # the allocating entry point still serves it
    assert np.isfinite(e0).all()

"""The fitting net's hidden layers in one kernel (fit_hidden_kernel: both hidden forward
layers, the energy head and their reverse pass, activations kept in registers) against the
per-GEMM path (one GEMM launch per layer, selected with DPMD_FIT_PER_GEMM=1 when a system is
created) and against the FP64 oracle: energies, forces and virials of the 240-wide model and
of a 128-wide one, in each precision."""
import os

import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _eval(device_model, coord, box, types, rc, prec, per_gemm):
    from deepmd_jax import _lib
    old = os.environ.get('DPMD_FIT_PER_GEMM')
    os.environ['DPMD_FIT_PER_GEMM'] = '1' if per_gemm else '0'
    try:
        s = _lib.DeviceSystem(device_model, types, 1)
    finally:
        if old is None:
            del os.environ['DPMD_FIT_PER_GEMM']
        else:
            os.environ['DPMD_FIT_PER_GEMM'] = old
    s.set_box(box)
    s.set_masses(np.ones(len(types)))
    s.set_positions(coord)
    s.nlist_allocate(rc, 1.2)
    return s.compute(prec)[:3]


@pytest.mark.parametrize('widths', [(240, 240, 240), (128, 128, 128)])
@pytest.mark.parametrize('prec', ['fp32', 'bf16', 'fp16'])
def test_fused_hidden_layers_match_per_gemm(widths, prec):
    from deepmd_jax import _lib, synthetic as syn
    print('case fit fused', widths, prec, flush=True)
    params, var = syn.init_model(4, fit_widths=widths)
    dm = _lib.DeviceModel(params, var)
    coord, box, types = syn.water_box(1024, seed=4)
    coord = np.asarray(coord, np.float32).astype(np.float64)
    ef, ff, vf = _eval(dm, coord, box, types, params['rcut'], prec, per_gemm=False)
    eg, fg, vg = _eval(dm, coord, box, types, params['rcut'], prec, per_gemm=True)
    N = len(types)
    fs = max(1.0, np.abs(fg).max())
    # K summed in another order (FP32 accumulation in both); in the 16-bit forms an activation
    # that lands on a rounding boundary can round to the neighbouring 16-bit value (2^-9 relative)
    tol = 2e-5 if prec == 'fp32' else 2e-3
    assert abs(ef[0] - eg[0]) / N <= tol * 0.1
    assert np.abs(ff - fg).max() <= tol * fs
    assert np.abs(vf - vg).max() <= tol * max(1.0, np.abs(vg).max())
    if prec == 'fp32':
        nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'])
        E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
        assert abs(ef[0] - E.item()) / N < 1e-5
        assert np.abs(ff[0] - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())
        assert np.abs(vf[0] - W.numpy()).max() <= 1e-4 * np.abs(W.numpy()).max()

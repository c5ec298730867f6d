"""The C++/OpenMP CPU restatement (oracle/dp_cpu.cpp) against the torch oracle (CPU).

The two are independent derivations of the same path: the torch oracle differentiates with
autograd, the C++ code carries a hand-written backward (fitting net, G, T, then per pair the
recomputed embedding net, env matrix and switch).  Agreement in FP64 to 1e-10 (relative to the
largest component) pins the oracle's forces and virial and the C++ derivatives at once; a
dropped term or a wrong index in either breaks it.  The FP32 build, the CPU baseline timed
beside the GPU, stays within the FP32 tolerances of the GPU parity tests.
"""
import numpy as np
import pytest

from oracle import dp_oracle as O


@pytest.fixture(scope='module')
def cpu():
    from oracle import dp_cpu
    dp_cpu.build()
    return dp_cpu


@pytest.mark.parametrize('topology', ['default', 'wide'])
def test_fp64_restatement_matches_torch_oracle(cpu, topology):
    from deepmd_jax import synthetic as syn
    if topology == 'default':
        params, var = syn.init_model(0)
    else:
        params, var = syn.init_model(1, embed_widths=(32, 32, 64), fit_widths=(64, 64, 96), axis=8)
    coord, box, types = syn.water_box(64, seed=3)
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 0.5)
    E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
    pk = cpu.Packed(params, var)
    rows, ks = cpu.neighbour_rows(nb, types)
    e, f, w, ea = cpu.energy_force(pk, coord, box, types, rows, ks, 'f64')
    assert abs(e - E.item()) <= 1e-10 * abs(E.item())
    assert np.abs(f - F.numpy()).max() <= 1e-10 * np.abs(F.numpy()).max()
    assert np.abs(w - W.numpy()).max() <= 1e-10 * np.abs(W.numpy()).max()
    assert abs(ea.sum() - e) <= 1e-12 * abs(e)
    # translation invariance of the hand-written forces: they sum to zero
    assert np.abs(f.sum(0)).max() <= 1e-10 * np.abs(f).max()


def test_fp32_restatement_within_fp32_tolerances(cpu):
    from deepmd_jax import synthetic as syn
    params, var = syn.init_model(0)
    coord, box, types = syn.water_box(216, seed=4)
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'])
    E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb)
    pk = cpu.Packed(params, var)
    rows, ks = cpu.neighbour_rows(nb, types)
    e, f, w, _ = cpu.energy_force(pk, coord, box, types, rows, ks, 'f32')
    N = len(types)
    assert abs(e - E.item()) / N <= 1e-5
    assert np.abs(f - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max())
    assert np.abs(w - W.numpy()).max() <= 1e-4 * np.abs(W.numpy()).max()

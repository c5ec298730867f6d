"""Compressed (tabulated) embedding nets: compress_model (deepmd_jax/utils.py:250-297) and
the tabulated evaluation (deepmd_jax/utils.py:126-151), on the CPU.

Pins:
  * the product's closed-form quintic Hermite coefficients against scipy's
    BPoly.from_derivatives -> PPoly.from_bernstein_basis (the routine the reference calls),
    fed by the oracle's autograd derivatives;
  * the product's analytic net derivatives against torch autograd;
  * the interpolation conditions (value, slope, curvature at every node) and the
    compression error between nodes;
  * the compressed oracle's energy against the uncompressed oracle (the error of the
    table), its forces against central finite differences (no gradient through the
    interval index), and model.pkl round trips of compress_var.
"""
import copy

import numpy as np
import pytest
import torch

from oracle import dp_oracle as O


def _range(params, i, rmin=0.6):
    from deepmd_jax.utils import sr
    return (np.array([0.0, sr(rmin, params['rcut'])]) - params['sr_mean'][i]) / params['sr_std'][i]


@pytest.mark.parametrize('k', [0, 3])
def test_hermite_coefficients_match_scipy(model0, k):
    from deepmd_jax.utils import embedding_net_jet, quintic_hermite
    params, var = model0
    net = var['params']['embedding_net_%d' % k]
    ref, nodes = O.compress_net(net, _range(params, k // 2), 128)
    f0, f1, f2 = embedding_net_jet(net, nodes)
    mine = quintic_hermite(nodes, f0, f1, f2)
    h = np.diff(nodes).max()
    scale = np.abs(f0).max()
    for c in range(6):
        # coefficient of x0^(5-c) weighted by its largest contribution on the interval
        assert np.abs(mine[:, c] - ref[:, c]).max() * h ** (5 - c) <= 1e-12 * scale, c


def test_net_derivatives_match_autograd(model0):
    from deepmd_jax.utils import embedding_net_jet
    params, var = model0
    net = var['params']['embedding_net_1']
    x = np.linspace(-3.0, 4.0, 37)
    f0, f1, f2 = embedding_net_jet(net, x)
    red = torch.ones(1, 1, 1, dtype=torch.float64)

    def f(t):
        return O.embedding_net(net, t.reshape(1, 1, 1), red)[0, 0]

    xt = torch.as_tensor(x)
    g0 = torch.stack([f(t) for t in xt]).numpy()
    g1 = torch.func.vmap(torch.func.jacfwd(f))(xt).numpy()
    g2 = torch.func.vmap(torch.func.jacfwd(torch.func.jacfwd(f)))(xt).numpy()
    np.testing.assert_allclose(f0, g0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(f1, g1, rtol=0, atol=1e-12 * np.abs(g1).max())
    np.testing.assert_allclose(f2, g2, rtol=0, atol=1e-11 * np.abs(g2).max())


def test_interpolation_conditions_and_error(model0):
    """The FP64 quintic reproduces value/slope/curvature at both ends of every interval,
    and the stored (float32) table stays within 1e-6 of the net in between."""
    from deepmd_jax.utils import embedding_net_jet, quintic_hermite, tabulated_jet
    params, var = model0
    net = var['params']['embedding_net_2']
    lo, hi = _range(params, 1)
    nodes = np.linspace(lo, hi, 65)
    f0, f1, f2 = embedding_net_jet(net, nodes)
    c = quintic_hermite(nodes, f0, f1, f2)
    h = np.diff(nodes)[:, None]
    # right end of each interval from the power basis
    v = sum(c[:, i] * h ** (5 - i) for i in range(6))
    d1 = sum((5 - i) * c[:, i] * h ** (4 - i) for i in range(5))
    d2 = sum((5 - i) * (4 - i) * c[:, i] * h ** (3 - i) for i in range(4))
    np.testing.assert_allclose(c[:, 5], f0[:-1], atol=0)
    np.testing.assert_allclose(c[:, 4], f1[:-1], atol=0)
    np.testing.assert_allclose(2 * c[:, 3], f2[:-1], rtol=1e-15, atol=0)
    np.testing.assert_allclose(v, f0[1:], rtol=0, atol=1e-12)
    np.testing.assert_allclose(d1, f1[1:], rtol=0, atol=1e-9 * np.abs(f1).max())
    np.testing.assert_allclose(d2, f2[1:], rtol=0, atol=1e-6 * np.abs(f2).max())
    # stored table (1,024 intervals, float32) against the net between the nodes
    cv = {'srmin': np.float32(lo), 'srmax': np.float32(hi)}
    nodes = np.linspace(lo, hi, 1025)
    cv['poly_coeff'] = quintic_hermite(nodes, *embedding_net_jet(net, nodes)).astype(np.float32)
    xs = np.linspace(lo, hi * (1 - 1e-6), 5001)
    p0, p1, _ = tabulated_jet(cv, xs)
    t0, t1, _ = embedding_net_jet(net, xs)
    assert np.abs(p0 - t0).max() < 1e-6
    assert np.abs(p1 - t1).max() < 1e-4 * np.abs(t1).max()


def test_oracle_interval_rule(model0c):
    """At s~ = srmin the tabulated net is the constant term of interval 0, i.e. the net's
    value at srmin; the torch restatement agrees with the NumPy jet everywhere."""
    from deepmd_jax.utils import embedding_net_jet, tabulated_jet
    params, var = model0c
    cv = var['compress_var']['embedding_net_0']
    s = torch.tensor([[[float(np.float32(cv['srmin']))]]], dtype=torch.float64)
    out = O.embedding_net_compressed(cv, s, torch.ones(1, 1, 1, dtype=torch.float64))[0, 0].numpy()
    np.testing.assert_allclose(out, np.asarray(cv['poly_coeff'][0, 5], np.float64), rtol=0, atol=0)
    f0, _, _ = embedding_net_jet(var['params']['embedding_net_0'], [float(np.float32(cv['srmin']))])
    assert np.abs(out - f0[0]).max() < 1e-6
    xs = np.linspace(float(cv['srmin']) - 0.001, float(cv['srmax']) + 0.05, 301)
    tj, _, _ = tabulated_jet(cv, xs)
    st = torch.as_tensor(xs).reshape(1, -1, 1)
    to = O.embedding_net_compressed(cv, st, torch.eye(301, dtype=torch.float64)[None])[0].numpy()
    np.testing.assert_allclose(to, tj, rtol=0, atol=1e-12)


def test_compressed_energy_close_to_mlp(model0, model0c, water64):
    coord, box, types = water64
    E, F, W = O.energy_force_virial(*model0, coord, box, types, lattice='auto')
    Ec, Fc, Wc = O.energy_force_virial(*model0c, coord, box, types, lattice='auto')
    N = len(types)
    assert abs(Ec.item() - E.item()) / N < 1e-6
    assert np.abs(Fc.numpy() - F.numpy()).max() < 1e-5
    assert np.abs(Wc.numpy() - W.numpy()).max() < 1e-5 * np.abs(W.numpy()).max() + 1e-6


def test_compressed_oracle_forces_vs_fd(model0):
    """Autograd forces of the tabulated oracle against central differences.  The table is
    kept in FP64 here: float32 coefficients leave ~1e-7 steps at interval ends, which a
    difference quotient over h = 1e-4 A would turn into errors of ~1e-5."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.utils import embedding_net_jet, quintic_hermite
    params, var = copy.deepcopy(model0)
    params['is_compressed'] = True
    var['compress_var'] = {}
    for k in range(4):
        lo, hi = _range(params, k // 2)
        nodes = np.linspace(lo, hi, 1025)
        var['compress_var']['embedding_net_%d' % k] = {
            'srmin': np.float64(lo), 'srmax': np.float64(hi),
            'poly_coeff': quintic_hermite(nodes, *embedding_net_jet(var['params']['embedding_net_%d' % k], nodes))}
    coord, box, types = syn.water_box(8, seed=4)
    lat = O.lattice_candidate(np.diag(box), params['rcut'])
    _, F, _ = O.energy_force_virial(params, var, coord, box, types, lattice=lat)
    h = 1e-4
    for i in (0, 5, 17):
        for d in range(3):
            es = []
            for sg in (1, -1):
                c2 = np.array(coord, np.float64)
                c2[i, d] += sg * h
                es.append(O.dp_energy(params, var, torch.as_tensor(c2), box, types, lattice=lat).item())
            fd = -(es[0] - es[1]) / (2 * h)
            assert abs(fd - F[i, d].item()) < 1e-7 * max(1.0, abs(fd)), (i, d, fd, F[i, d].item())


def test_compressed_model_pkl_roundtrip(tmp_path, model0c):
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import load_model, save_model
    params, var = model0c
    path = str(tmp_path / 'model_c.pkl')
    save_model(path, DPModel(copy.deepcopy(params)), var)
    m, v = load_model(path)
    assert m.params['is_compressed'] is True
    for k in range(4):
        a = var['compress_var']['embedding_net_%d' % k]
        b = v['compress_var']['embedding_net_%d' % k]
        assert b['poly_coeff'].dtype == np.float32 and b['poly_coeff'].shape == (1024, 6, 100)
        np.testing.assert_array_equal(a['poly_coeff'], b['poly_coeff'])
        assert np.float32(a['srmin']) == np.float32(b['srmin'])
        assert np.float32(a['srmax']) == np.float32(b['srmax'])

"""Device PPPM (deepmd_jax/utils.py:177-222) against the FP64 oracle: energy within 1e-10
relative, forces (the oracle's autograd gradient) within 1e-9 * max |F|, the virial diagonal
(-dE/de_a with charges and box scaled along a) within 1e-7 of central differences; all FP64.
Cases: a neutral random set in a non-cubic box with odd and even grid sizes, a charge set
with atoms outside the primary cell, and a water-like box at the resolution DPLR uses."""
import numpy as np
import pytest
import torch

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('L,res,n,shift', [((10., 11., 12.), 5, 40, 0), ((13., 13., 9.5), 3, 200, 1),
                                           ((18.6, 18.6, 18.6), 5, 864, 0)])
def test_pppm_against_oracle(L, res, n, shift):
    from deepmd_jax import _lib
    print('case pppm', L, res, n, flush=True)
    rng = np.random.default_rng(n)
    L = np.asarray(L)
    x = rng.uniform(0, 1, (n, 3)) * L + shift * L * rng.integers(-2, 3, (n, 3))
    q = rng.standard_normal(n)
    q -= q.mean()
    M = O.p3m_grid(L, 0.4, res)
    xt = torch.as_tensor(x).clone().requires_grad_(True)
    E = O.p3m_energy(xt, q, L, 0.4, M)
    F = -torch.autograd.grad(E, [xt])[0].numpy()
    e, f, w = _lib.pppm_compute(L, 0.4, M, x, q)
    assert abs(e - E.item()) <= 1e-10 * abs(E.item()), (e, E.item())
    assert np.abs(f - F).max() <= 1e-9 * np.abs(F).max()
    h = 1e-5
    for a in range(3):
        s = np.ones(3)
        s[a] = 1 + h
        ep = O.p3m_energy(torch.as_tensor(x * s), q, L * s, 0.4, M).item()
        s[a] = 1 - h
        em = O.p3m_energy(torch.as_tensor(x * s), q, L * s, 0.4, M).item()
        assert abs(w[a, a] + (ep - em) / (2 * h)) <= 1e-7 * max(1.0, abs(E.item())), (a, w[a, a], (ep - em) / (2 * h))
    np.testing.assert_allclose(w, w.T, atol=1e-12 * abs(E.item()))


def test_get_p3mlr_fn_matches_oracle():
    """deepmd_jax.utils.get_p3mlr_fn (deepmd_jax/utils.py:178-222) on the GPU: energy of random
    neutral charges against the oracle's PPPM, forces and virial from energy_force_virial."""
    from deepmd_jax.utils import get_p3mlr_fn
    print('case get_p3mlr_fn', flush=True)
    rng = np.random.default_rng(7)
    box = np.array([11.0, 12.5, 10.2])
    x = rng.uniform(0, 1, (60, 3)) * box
    q = rng.uniform(-1, 1, 60)
    q -= q.mean()
    fn = get_p3mlr_fn(box, 0.4)
    M = O.p3m_grid(box, 0.4, 5)
    assert tuple(fn.grid) == tuple(int(m) for m in M)
    e_ref = O.p3m_energy(torch.as_tensor(x), q, box, 0.4, M).item()
    assert abs(fn(x, q, box) - e_ref) <= 1e-10 * max(1.0, abs(e_ref))
    e, f, w = fn.energy_force_virial(x, q, box)
    xt = torch.as_tensor(x).clone().requires_grad_(True)
    g, = torch.autograd.grad(O.p3m_energy(xt, q, box, 0.4, M), [xt])
    np.testing.assert_allclose(f, -g.numpy(), rtol=0, atol=1e-9 * np.abs(g.numpy()).max())

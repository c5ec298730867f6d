"""Pins of the oracle's stochastic steps: the Philox4x32-10 stream against the published
known-answer vectors (Random123, Salmon et al. SC'11), the Box-Muller normals by their
moments and distribution function, and the BAOAB / PILE-L O steps by their gamma -> 0 limit
and their stationary temperature (a wrong c1 or c2 moves it)."""
import numpy as np
import pytest

from oracle import dp_oracle as O


@pytest.mark.parametrize('ctr,key,out', [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, out):
    got = O.philox4x32_10(np.array([ctr], np.uint64), np.array([key], np.uint64))[0]
    assert [int(v) for v in got] == out


def test_normals_moments_and_distribution():
    from scipy import stats
    z = O.normal4(12345, 7, np.arange(100000), O.LANGEVIN_STREAM)
    flat = z.reshape(-1)
    n = flat.size
    assert abs(flat.mean()) < 5 / np.sqrt(n)
    assert abs(flat.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs(stats.skew(flat)) < 5 * np.sqrt(6 / n) and abs(stats.kurtosis(flat)) < 5 * np.sqrt(24 / n)
    assert stats.kstest(flat, 'norm').pvalue > 1e-4
    # the four words of a draw are uncorrelated, and so are successive counters and streams
    c = np.corrcoef(z.T)
    assert np.abs(c - np.eye(4)).max() < 5 / np.sqrt(len(z))
    z2 = O.normal4(12345, 8, np.arange(100000), O.LANGEVIN_STREAM)
    z3 = O.normal4(12345, 7, np.arange(100000), O.PIMD_STREAM)
    for other in (z2, z3):
        assert abs(np.corrcoef(z[:, 0], other[:, 0])[0, 1]) < 5 / np.sqrt(len(z))


def test_streams_are_keyed():
    """Segment seeds give distinct streams; the same key and counter repeat exactly."""
    a = O.langevin_noise(7, 0, 3, 50)
    assert np.array_equal(a, O.langevin_noise(7, 0, 3, 50))
    assert not np.allclose(a, O.langevin_noise(7, 1, 3, 50))
    assert not np.allclose(a, O.langevin_noise(7, 0, 4, 50))
    assert O.segment_key(7, 0) == 7 and O.segment_key(0, 1) == 0x9E3779B97F4A7C15


def test_baoab_gamma_zero_is_velocity_verlet():
    rng = np.random.default_rng(0)
    m = rng.uniform(1, 20, 10)
    k = 0.7
    ff = lambda x: -k * x
    x, p = rng.standard_normal((10, 3)), rng.standard_normal((10, 3))
    F = ff(x)
    xa, pa, Fa, dke = O.nvt_langevin_step(x, p, m, ff, 0.3, F, 0.0, 0.05, rng.standard_normal((10, 3)))
    xb, pb, Fb = O.vv_step(x, p, m, ff, 0.3, F)
    np.testing.assert_allclose(xa, xb, atol=1e-14)
    np.testing.assert_allclose(pa, pb, atol=1e-14)
    assert dke == 0.0


def test_baoab_stationary_temperature_and_work():
    """Free particles under the O step relax to <p^2/m> = kT per component, and the
    thermostat work equals the kinetic-energy change (no forces)."""
    n, kT, m = 20000, 0.04, np.full(20000, 3.0)
    x, p = np.zeros((n, 3)), np.zeros((n, 3))
    F = np.zeros((n, 3))
    work = 0.0
    for step in range(150):
        x, p, F, dke = O.nvt_langevin_step(x, p, m, lambda y: np.zeros_like(y), 1.0, F, 0.05, kT,
                                           O.langevin_noise(3, 0, step, n))
        work += dke
    ke = 0.5 * (p ** 2 / m[:, None]).sum()
    assert abs(work - ke) < 1e-9 * ke
    assert abs((p ** 2 / m[:, None]).mean() / kT - 1) < 0.02


def test_pile_l_stationary_bead_temperature():
    """Free ring polymers under PILE-L: every bead's momentum component relaxes to
    <p^2/m> = P kT (the ring-polymer temperature of deepmd_jax/pimd.py:99-100)."""
    from deepmd_jax.pimd import normal_mode_transform, pile_l_factors
    P, n, kT, dt = 4, 4000, 0.025, 0.5
    freqs, C = normal_mode_transform(P, kT)
    _, c1, c2 = pile_l_factors(freqs, dt, kT, P, 10.0, 1.0)
    m = np.full(n, 2.0)
    x = np.zeros((P, n, 3)) + 5.0
    p = np.zeros((P, n, 3))
    F = np.zeros((P, n, 3))
    for step in range(300):
        x, p, F = O.pimd_step(x, p, m, lambda y: np.zeros_like(y), dt, C, freqs, c1, c2, F, np.full(3, 100.0),
                              noise=O.pimd_noise(9, step, n, P))
    t = (p ** 2 / m[None, :, None]).mean() / (P * kT)
    assert abs(t - 1) < 0.03

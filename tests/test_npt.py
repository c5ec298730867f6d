"""NPT restatement (deepmd_jax/routine.py:161-456) on the CPU.

The oracle's NPT step is pinned by conservation of the NPT invariant under an analytic
soft-sphere potential (forces and the per-axis strain derivative in closed form), for
isotropic, coupled-plus-independent, uniaxial and fully anisotropic couple_axes, while
the box volume moves by more than 5 %.  A dropped or mis-signed term (P V, box kinetic
energy, the tr(v)/N_f coupling, a chain term) breaks the conservation.

Finding pinned here: with G > 1 groups the barostat chain has G degrees of freedom, so the
conserved quantity carries G kT xi_0 for the chain head; the reference's invariant
(deepmd_jax/routine.py:447-449) writes kT xi_0 and therefore drifts by (G-1) kT dxi_0.
The reported Invariant follows the reference's formula; the conserved form is tested.
"""
import numpy as np
import pytest

from oracle import dp_oracle as O


def soft_spheres(x, box, sig=1.0, k=20.0):
    """U = sum_{i<j, r<sig} k (sig - r)^2 with minimum image; returns (U, F, dU/deps_axis)."""
    d = x[:, None] - x[None]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(len(x))
    ov = (r < sig) & ~np.eye(len(x), dtype=bool)
    dUdr = np.where(ov, -2 * k * (sig - r), 0.0)
    U = 0.5 * np.where(ov, k * (sig - r) ** 2, 0.0).sum()
    F = -(dUdr[..., None] * d / r[..., None]).sum(1)
    dudeps = 0.5 * (dUdr[..., None] * d ** 2 / r[..., None]).sum((0, 1))
    return U, F, dudeps


def _start(seed=0, n=4, L=4.2, kT=0.5):
    rng = np.random.default_rng(seed)
    g = (np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing='ij'), -1).reshape(-1, 3) + 0.5) * (L / n)
    x = g + rng.uniform(-0.1, 0.1, g.shape)
    p = rng.standard_normal(x.shape) * np.sqrt(kT)
    return x, p, np.ones(len(x)), np.full(3, L), kT


def test_soft_sphere_strain_derivative_matches_fd():
    x, _, _, box, _ = _start(1)
    _, _, d = soft_spheres(x, box)
    for a in range(3):
        e = 1e-6
        s = np.ones(3)
        s[a] += e
        up = soft_spheres(x * s, box * s)[0]
        s[a] -= 2 * e
        dn = soft_spheres(x * s, box * s)[0]
        assert abs((up - dn) / (2 * e) - d[a]) < 1e-5 * max(1.0, abs(d[a]))


@pytest.mark.parametrize('couple', [((0, 1, 2),), ((0, 1), (2,)), ((0,),), ((0,), (1,), (2,))])
def test_npt_invariant_conserved(couple):
    x, p, m, box, kT = _start()
    o = O.NPTOracle(x, p, m, box, lambda xx, b: soft_spheres(xx, b)[1:], 0.005, kT, 1.0, 0.5, 2.0,
                    couple_axes=couple)
    H, V, ref_form = [], [], []
    for s in range(2000):
        o.step()
        if s % 20 == 0:
            Hr = o.invariant(soft_spheres(o.x, o.box)[0])
            ref_form.append(Hr)
            H.append(Hr + (o.G - 1) * kT * o.bxi[0])
            V.append(np.prod(o.box))
    H, V = np.array(H), np.array(V)
    assert V.max() / V.min() > 1.05
    assert abs(H[-1] - H[0]) < 0.05 and H.std() < 0.02, (H[-1] - H[0], H.std())
    # the reference's form (kT xi_0 for the chain head) is conserved only for one group
    if o.G > 1:
        assert abs(ref_form[-1] - ref_form[0]) > 5 * abs(H[-1] - H[0])


def test_npt_fixed_axis_stays_fixed():
    x, p, m, box, kT = _start(2)
    o = O.NPTOracle(x, p, m, box, lambda xx, b: soft_spheres(xx, b)[1:], 0.005, kT, 1.0, 0.5, 2.0,
                    couple_axes=((0, 1),))
    for _ in range(200):
        o.step()
    assert o.box[2] == box[2] and abs(o.box[0] - box[0]) > 1e-3 and o.box[0] == o.box[1]


@pytest.mark.parametrize('bad', [((0, 3),), ((0, 1), (1,)), ((),), (), 5])
def test_parse_couple_axes_rejects(bad):
    from deepmd_jax.md import parse_couple_axes
    with pytest.raises(ValueError):
        parse_couple_axes(bad)


def test_parse_couple_axes_groups():
    from deepmd_jax.md import parse_couple_axes
    groups, gid = parse_couple_axes(((0, 1), (2,)))
    assert groups == ((0, 1), (2,)) and list(gid) == [0, 0, 1]
    _, gid = parse_couple_axes(((1,),))
    assert list(gid) == [-1, 0, -1]


def _pimd_start(nb=4, seed=3, kT=0.5, hbar=0.25):
    x, p, m, box, kT = _start(seed, kT=kT)
    rng = np.random.default_rng(seed + 10)
    xb = x[None] + rng.normal(0, 0.05, (nb,) + x.shape)
    pb = rng.standard_normal((nb,) + x.shape) * np.sqrt(kT * nb)
    freqs, C = O.ring_modes(nb, kT, hbar)
    return xb, pb, m, box, kT, freqs, C, nb * kT / hbar


@pytest.mark.parametrize('couple', [((0, 1, 2),), ((0, 1), (2,))])
def test_pimd_npt_invariant_conserved(couple):
    """deepmd_jax/pimd.py:297-518 with the O steps switched off (c1 = 1, no noise): the reported
    invariant (H_rp / P + KE_box + P V) is conserved while the volume moves by > 5 %; a dropped
    internal-virial term, centroid-only coupling or sign breaks it."""
    xb, pb, m, box, kT, freqs, C, wP = _pimd_start()
    nb = len(xb)

    def ff(xx, b):
        _, F, d = soft_spheres(xx, b)
        return F, d
    o = O.PIMDNPTOracle(xb, pb, m, box, ff, 0.004, kT, 1.0, C, freqs, np.ones(nb), np.zeros(nb), 0.0, 2.0,
                        couple_axes=couple)
    H, V = [], []
    for s in range(1500):
        o.step()
        if s % 25 == 0:
            pe = [soft_spheres(o.x[b], o.box)[0] for b in range(nb)]
            H.append(o.invariant(pe, wP))
            V.append(np.prod(o.box))
    H, V = np.array(H), np.array(V)
    assert V.max() / V.min() > 1.05, V.max() / V.min()
    assert abs(H[-1] - H[0]) < 0.05 and H.std() < 0.02, (H[-1] - H[0], H.std())


def test_pimd_npt_one_bead_is_classical_mtk_without_chains():
    """With one bead the centroid is the particle: the step reduces to the MTK barostat step of
    NPTOracle with its thermostat and barostat chains absent (c1 = 1, gamma_box = 0)."""
    x, p, m, box, kT = _start(4)

    def ff(xx, b):
        _, F, d = soft_spheres(xx, b)
        return F, d
    o1 = O.PIMDNPTOracle(x[None], p[None], m, box, ff, 0.004, kT, 1.0, np.ones((1, 1)), np.zeros(1), np.ones(1),
                         np.zeros(1), 0.0, 2.0)
    o2 = O.NPTOracle(x, p, m, box, ff, 0.004, kT, 1.0, 1e12, 2.0)
    o2._chains = lambda: (np.full(3, 1e300), np.full(3, 1e300))
    for _ in range(50):
        o1.step()
        o2.step()
    np.testing.assert_allclose(o1.box, o2.box, rtol=1e-12)
    np.testing.assert_allclose(o1.x[0], o2.x, atol=1e-10)
    np.testing.assert_allclose(o1.p[0], o2.p, atol=1e-10)

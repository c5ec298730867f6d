"""Oracle pins for fixed atoms, Nose-Hoover NVT and centre-of-mass removal (CPU).

Fixed atoms (deepmd_jax/md.py:221-234,428-430; deepmd_jax/routine.py:110-136) are pinned
by properties the stop-gradient fixes: the energy and the strain derivative do not change,
the forces on mobile atoms are those of the unconstrained model and the forces on fixed
atoms are zero.  The Nose-Hoover step is pinned by conservation of its invariant with and
without fixed atoms; centre-of-mass removal by the zero mobile momentum it leaves.
"""
import numpy as np

from oracle import dp_oracle as O


def _mobile(n, seed=0, frac=0.25):
    rng = np.random.default_rng(seed)
    mob = np.ones(n, bool)
    mob[rng.choice(n, int(frac * n), replace=False)] = False
    return mob


def test_fixed_forces_are_the_unconstrained_ones_on_mobile_atoms(model0, water64):
    params, var = model0
    coord, box, types = water64
    mob = _mobile(len(types))
    E0, F0, W0 = O.energy_force_virial(params, var, coord, box, types, lattice='auto')
    E1, F1, W1 = O.energy_force_virial(params, var, coord, box, types, lattice='auto', mobile=mob)
    assert abs(E1.item() - E0.item()) < 1e-12
    np.testing.assert_allclose(F1.numpy()[mob], F0.numpy()[mob], rtol=0, atol=1e-12)
    assert np.all(F1.numpy()[~mob] == 0)
    # the strain acts on every coordinate: the virial is unchanged
    np.testing.assert_allclose(W1.numpy(), W0.numpy(), rtol=0, atol=1e-12)


def _force_fn(params, var, box, types, mob=None):
    lat = O.lattice_candidate(np.diag(box), params['rcut'])

    def f(x):
        return O.energy_force_virial(params, var, x, box, types, lattice=lat, mobile=mob)[1].numpy()

    def e(x):
        return O.energy_force_virial(params, var, x, box, types, lattice=lat)[0].item()
    return f, e


def test_fixed_nve_conserves_energy_and_holds_fixed_atoms(model0):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    params, var = model0
    coord, box, types = syn.water_box(27, seed=5)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    mob = _mobile(len(types), seed=1)
    f, e = _force_fn(params, var, box, types, mob)
    p = syn.maxwell_boltzmann(types, 300.0, seed=1) * m[:, None] * mob[:, None]
    x, F = coord.copy(), f(coord)
    h0 = e(x) + 0.5 * (p ** 2 / m[:, None]).sum()
    for _ in range(30):
        x, p, F = O.vv_step(x, p, m, f, 0.25, F, mobile=mob)
    h1 = e(x) + 0.5 * (p ** 2 / m[:, None]).sum()
    assert np.array_equal(x[~mob], coord[~mob])
    assert np.all(p[~mob] == 0)
    assert abs(h1 - h0) / len(types) < 2e-5, (h0, h1)


def test_nvt_nose_hoover_invariant_with_and_without_fixed(model0):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    params, var = model0
    coord, box, types = syn.water_box(27, seed=6)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT, tau = 400 * TEMP_UNIT_CONVERSION, 20.0
    for mob in (None, _mobile(len(types), seed=2)):
        f, e = _force_fn(params, var, box, types, mob)
        p = syn.maxwell_boltzmann(types, 300.0, seed=2) * m[:, None]
        if mob is not None:
            p = p * mob[:, None]
        dof = 3 * (len(types) if mob is None else int(mob.sum()))
        xi, px = np.zeros(3), np.zeros(3)
        x, F = coord.copy(), f(coord)
        h = lambda: e(x) + 0.5 * (p ** 2 / m[:, None]).sum() + O.nh_chain_energy(xi, px, kT, tau, dof)
        h0 = h()
        for _ in range(40):
            x, p, F = O.nvt_nh_step(x, p, m, f, 0.25, F, xi, px, kT, tau, mobile=mob)
        h1 = h()
        # the thermostat did work (the chain moved) and the invariant held
        assert abs(px[0]) > 0
        assert abs(h1 - h0) / len(types) < 2e-5, (mob is None, h0, h1)
        if mob is not None:
            assert np.array_equal(x[~mob], coord[~mob])


def test_remove_com_zeroes_mobile_momentum():
    rng = np.random.default_rng(3)
    m = rng.uniform(1, 16, 50)
    p = rng.standard_normal((50, 3))
    mob = _mobile(50, seed=3)
    q = O.remove_com(p, m, mob)
    np.testing.assert_allclose((q * mob[:, None]).sum(0), 0, atol=1e-12)
    # every atom loses the same velocity, the mobile-mass-weighted mean (fixed ones too)
    dv = (p - q) / m[:, None]
    np.testing.assert_allclose(dv, np.broadcast_to(dv[0], dv.shape), atol=1e-12)
    vcom = (p * mob[:, None]).sum(0) / m[mob].sum()
    np.testing.assert_allclose(dv[0], vcom, atol=1e-12)
    np.testing.assert_allclose(O.remove_com(p, m).sum(0), 0, atol=1e-12)


def test_nose_hoover_substeps_conserve_and_refine(model0):
    """chain_steps x sy_steps substeps (jax_md's Suzuki-Yoshida composition): the invariant holds
    and the chain integration of a stiff chain (tau 2 fs) is closer to a fine-step reference with
    3 Suzuki-Yoshida substeps than with one."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    params, var = model0
    coord, box, types = syn.water_box(8, seed=8)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT, tau, dt = 600 * TEMP_UNIT_CONVERSION, 2.0, 0.25
    f, e = _force_fn(params, var, box, types)
    dof = 3 * len(types)

    def run(cs, sy, steps=20):
        p = syn.maxwell_boltzmann(types, 300.0, seed=8) * m[:, None]
        xi, px = np.zeros(3), np.zeros(3)
        x, F = coord.copy(), f(coord)
        h0 = e(x) + 0.5 * (p ** 2 / m[:, None]).sum() + O.nh_chain_energy(xi, px, kT, tau, dof)
        for _ in range(steps):
            x, p, F = O.nvt_nh_step(x, p, m, f, dt, F, xi, px, kT, tau, chain_steps=cs, sy_steps=sy)
        h1 = e(x) + 0.5 * (p ** 2 / m[:, None]).sum() + O.nh_chain_energy(xi, px, kT, tau, dof)
        return h1 - h0, px.copy()
    d1, px1 = run(1, 1)
    d3, px3 = run(2, 3)
    dref, pxref = run(8, 7)
    assert abs(d3) / len(types) < 1e-4
    assert np.abs(px3 - pxref).max() < np.abs(px1 - pxref).max()

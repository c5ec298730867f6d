"""Neighbour lists on the device at the shapes the configs run them, bit-exact against the
oracle's FP32 restatement of deepmd_jax/md_utils.py:364-373,393-459, plus the rebuild
decision of deepmd_jax/md.py:720-732 and the error bits of md.py:616-642.

- the config-3 bead box (4,096 H2O, L = 49.67 A, rc+skin 7 A: 7 cells per axis) and the
  config-1 box (32,768 H2O, L = 99.34 A: 14 cells per axis): every row of both lists;
- a list whose reference box differs from the current box (scale != 1, the NPT case);
- the list rebuilt by the drift test inside md_run, against the positions it was built
  from (unwrapped: atoms have left the box);
- one atom displaced to just below and just above the drift threshold, at scale 1 and
  scale != 1, NVE and NPT (the 1.02 safety factor);
- NaN / Inf: error bit 16 from md_run (plain and path-integral) and the Simulation stop;
- NPT: error bit 1 when the box shrinks past 0.8 skin / (rc + skin) of the list's box.
"""
import contextlib
import io

import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _system(device_model, coord, box, types, rc_skin, masses=None):
    from deepmd_jax import _lib
    s = _lib.DeviceSystem(device_model, types, 1)
    s.set_box(box)
    s.set_masses(np.ones(len(types)) if masses is None else masses)
    s.set_positions(coord)
    s.nlist_allocate(rc_skin, 1.2)
    return s


def _rows_equal(s, ref, types):
    off = np.concatenate([[0], np.cumsum(np.bincount(types))])
    for j in range(2):
        got = s.nlist_export(j)[0]
        for i in range(2):
            np.testing.assert_array_equal(got[off[i]:off[i + 1]], ref[i][j], err_msg='block %d,%d' % (i, j))


def _bead_box():
    """One bead of the config-3 system: 4,096 H2O on a 16^3 grid, seed 3, +-0.05 A jitter."""
    from deepmd_jax import synthetic as syn
    coord, box, types = syn.water_box(4096, seed=3)
    rng = np.random.default_rng(3)
    coord = np.mod(coord + rng.uniform(-0.05, 0.05, coord.shape), box)
    return coord, box, types


@pytest.mark.parametrize('which', ['bead4096', 'water98k'])
def test_nlist_rows_at_config_shapes(device_model, which):
    from deepmd_jax import synthetic as syn
    print('case nlist rows', which, flush=True)
    coord, box, types = _bead_box() if which == 'bead4096' else syn.water_box(32768, seed=0)
    s = _system(device_model, coord, box, types, 7.0)
    ref, overflow, knbr = O.typed_nlist(coord, box, types, 7.0)
    assert s.knbr == knbr and not overflow
    if which == 'water98k':
        assert knbr == [68, 135]
    _rows_equal(s, ref, types)


@pytest.mark.parametrize('factor', [1.004, 0.996])
def test_nlist_rows_scaled_reference_box(device_model, factor):
    """The box moves after allocation (as under NPT): the rebuilt list is binned on the
    reference box with canonical coordinates scaled by L_ref / box."""
    print('case nlist scaled', factor, flush=True)
    coord, box, types = _bead_box()
    s = _system(device_model, coord, box, types, 7.0)
    nbox = box * factor
    x = coord * factor
    s.set_box(nbox)
    s.set_positions(x)
    assert not s.nlist_update()
    ref, _, _ = O.typed_nlist(x, nbox, types, 7.0, knbr=s.knbr, reference_box=box)
    _rows_equal(s, ref, types)


def test_nlist_rows_after_drift_rebuild(device_model):
    """NVE steps one at a time on the bead box (skin 0.5 A, dt 1 fs), started from positions
    shifted by whole box vectors, until the drift test has rebuilt the list twice; the rows
    then equal the oracle's list of the positions at the start of that step (FP64,
    unwrapped, canonicalised through the FP64 wrap)."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case nlist drift rebuild', flush=True)
    coord, box, types = _bead_box()
    coord = coord + box * np.array([1.0, -2.0, 3.0])   # every coordinate outside [0, L)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    s = _system(device_model, coord, box, types, 6.5, masses=m)
    s.nlist_allocate(6.5, 1.6)   # headroom: the hot seeded potential clusters atoms
    s.set_velocities(syn.maxwell_boltzmann(types, 600.0, seed=3))
    p = _lib.MDParams()
    p.routine, p.dt = 0, 1.0
    seen = 0
    for step in range(400):
        x = s.get_positions()[0]
        _, _, err = s.md_run(p, 1, record=False)
        assert err == 0
        if s.nlist_rebuilds() > seen:
            seen = s.nlist_rebuilds()
            ref, _, _ = O.typed_nlist(x, box, types, 6.5, knbr=s.knbr)
            _rows_equal(s, ref, types)
            if seen >= 2:
                break
    assert seen >= 2
    assert np.any((x < 0) | (x >= box)), 'no atom left the box: the unwrapped path was not exercised'


def _f32_drift(x, x0, box_now, ref_box):
    """deepmd_jax/md.py:701-707 in FP32: |minimg_Lref((f32(x) mod box) * scale - canon(x0))|."""
    f = np.float32
    scale = (ref_box.astype(f) / box_now.astype(f)).astype(f)
    cur = (np.mod(np.mod(x, box_now).astype(f), box_now.astype(f)) * scale).astype(f)
    c0 = O.canonicalize_f32(x0[None], ref_box)[0]
    d = (cur - c0).astype(f)
    L = ref_box.astype(f)
    h = (L / f(2)).astype(f)
    d = (np.mod((d - h).astype(f), L) - h).astype(f)
    return f(np.sqrt(f((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])))


def _f32_threshold(rc_skin, rc, scale_max, ref_box, npt):
    f = np.float32
    safe = f(f(scale_max) * f(1.02)) if npt else f(scale_max)
    allowed = f(f(f(rc_skin) - f(f(rc) * safe)) / f(2))
    return f(allowed - f(f(ref_box.min()) * f(1e-6)))


@pytest.mark.parametrize('routine,factor', [('NVE', 1.0), ('NVE', 1.003), ('NPT', 0.999), ('NPT', 1.003)])
def test_rebuild_decision_at_the_threshold(device_model, routine, factor):
    """One atom is moved so its FP32 drift sits 1e-5 A below, then 1e-5 A above, the FP32
    threshold ((rc+skin) - rc max(L_ref/box) [1.02 under NPT]) / 2 - 1e-6 L_min; one step with
    a negligible dt rebuilds the list only in the second case.  The margin is below the
    1e-6 L_min term (2.5e-5 A), so that term is pinned too."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case rebuild decision', routine, factor, flush=True)
    rc, rc_skin = 6.0, 7.0
    coord, box, types = syn.water_box(512, seed=8)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    nbox = box * factor
    a = int(np.nonzero((coord[:, 0] > 6) & (coord[:, 0] < 18) & (types == 0))[0][0])
    thr = _f32_threshold(rc_skin, rc, np.float32(box[0]) / np.float32(nbox[0]), box.astype(np.float32),
                         routine == 'NPT')
    x = coord.copy()
    others = max(_f32_drift(x[i], coord[i], nbox, box) for i in range(len(x)))
    assert others < thr - 0.05

    def delta_for(target):
        lo, hi = 0.0, 1.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            y = x[a].copy()
            y[0] += mid
            lo, hi = (mid, hi) if _f32_drift(y, coord[a], nbox, box) < target else (lo, mid)
        return 0.5 * (lo + hi)

    for side, target in (('below', thr - np.float32(1e-5)), ('above', thr + np.float32(1e-5))):
        s = _system(device_model, coord, box, types, rc_skin, masses=m)
        s.set_velocities(np.zeros_like(coord))
        if factor != 1.0:
            s.set_box(nbox)
        y = x.copy()
        y[a, 0] += delta_for(target)
        d = _f32_drift(y[a], coord[a], nbox, box)
        assert (d < thr) if side == 'below' else (d > thr)
        s.set_positions(y)
        p = _lib.MDParams()
        p.routine, p.dt = (0, 1e-6) if routine == 'NVE' else (4, 1e-6)
        if routine == 'NPT':
            p.kT, p.tau, p.chain_length = 300 * 8.617333e-5, 100.0, 3
            p.pressure, p.tau_p, p.chain_length_p = 6.241509e-7, 1e6, 3
            p.couple[0] = p.couple[1] = p.couple[2] = 0
        _, _, err = s.md_run(p, 1, record=False)
        assert err == 0
        assert s.nlist_rebuilds() == (1 if side == 'above' else 0), (routine, factor, side, d, thr)


def test_nan_sets_bit16_and_stops_simulation(device_model, model0):
    """A NaN coordinate gives error bit 16 from md_run (plain and path-integral NVT) and
    Simulation.run prints the warning and stops after the chunk (deepmd_jax/md.py:909-911)."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    from deepmd_jax.pimd import normal_mode_transform, pile_l_factors
    print('case nan', flush=True)
    params, var = model0
    coord, box, types = syn.water_box(216, seed=4)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    bad = coord.copy()
    bad[7, 1] = np.nan
    s = _system(device_model, coord, box, types, 7.0, masses=m)
    s.set_positions(bad)
    p = _lib.MDParams()
    p.routine, p.dt = 0, 0.5
    _, _, err = s.md_run(p, 2, record=False)
    assert err & 16
    # path-integral NVT: a non-finite momentum on one bead
    P = 2
    kT = 300 * TEMP_UNIT_CONVERSION
    sp = _lib.DeviceSystem(device_model, types, P)
    sp.set_box(box)
    sp.set_masses(m)
    sp.set_positions(np.broadcast_to(coord, (P,) + coord.shape))
    v = np.zeros((P,) + coord.shape)
    v[1, 3, 0] = np.inf
    sp.set_velocities(v)
    sp.nlist_allocate(7.0, 1.2)
    freqs, C = normal_mode_transform(P, kT)
    _, c1, c2 = pile_l_factors(freqs, 0.5, kT, P, 100.0, 1.0)
    sp.pimd_setup(C, freqs, c1, c2)
    pp = _lib.MDParams()
    pp.routine, pp.dt, pp.seed = 3, 0.5, 1
    _, _, err = sp.md_run(pp, 2, record=False)
    assert err & 16
    # Simulation: the chunk runs, the warning prints, the run ends
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=np.zeros_like(coord),
                         report_interval=10 ** 6,
                         neighbor_skin=1.0, model=(DPModel(params), var, device_model))
        sim._step_chunk_size = 5
        sim.setPosition(bad)
        sim.run(20)
    assert 'Nan or Inf encountered' in out.getvalue()
    assert sim.step == 5


def test_npt_box_shrink_sets_bit1(device_model):
    """NPT with the list allocated on box L: at 5 % linear shrink (below 0.8 skin/(rc+skin)
    = 6.15 % for skin 0.5 A) the step runs; at 7 % md_run returns bit 1 and restores the
    entry state (md.py:630-636)."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    print('case npt shrink', flush=True)
    coord, box, types = syn.water_box(512, seed=9)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    for shrink, expect in ((0.05, 0), (0.07, 1)):
        s = _system(device_model, coord, box, types, 6.5, masses=m)
        s.set_velocities(np.zeros_like(coord))
        nbox = box * (1 - shrink)
        s.set_box(nbox)
        s.set_positions(coord * (1 - shrink))
        p = _lib.MDParams()
        p.routine, p.dt, p.kT, p.tau, p.chain_length = 4, 0.5, 300 * 8.617333e-5, 100.0, 3
        p.pressure, p.tau_p, p.chain_length_p = 6.241509e-7, 1000.0, 3
        p.couple[0] = p.couple[1] = p.couple[2] = 0
        _, _, err = s.md_run(p, 2, record=False)
        assert err == expect, (shrink, err)
        if expect:
            np.testing.assert_allclose(s.get_positions()[0], coord * (1 - shrink), rtol=0, atol=1e-12)

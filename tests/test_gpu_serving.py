"""Ragged (uneven-size) batches and the serving loop on the GPU, through the C ABI.

Tolerances:
  same request alone vs in a batch: energy |dE| <= 1e-6 |E| + 1e-6 eV (the per-segment
  energy reduction does not depend on the batch), forces max |dF| <= 1e-5 max |F|
  (FP32 force atomics accumulate in a launch-dependent order), virial <= 1e-5 max |W|;
  MD end state after 40-90 Langevin steps: positions <= 1e-4 A, velocities <= 1e-4 A/fs.
  small boxes held as supercells vs the FP64 oracle's lattice-image sum: the plain
  FP32 tolerances of test_gpu_parity.py.
Neighbour lists of a segment: bit-exact against the same box alone.
"""
import numpy as np
import pytest

from oracle import dp_oracle as O

pytestmark = pytest.mark.gpu


def _boxes(ns, seed0=20):
    from deepmd_jax import synthetic as syn
    return [syn.water_box(n, seed=seed0 + i, permute_sites=True) for i, n in enumerate(ns)]


def _f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def test_ragged_nlist_and_efv_match_alone(device_model):
    from deepmd_jax import _lib
    print('case ragged nlist/efv', flush=True)
    boxes = _boxes([70, 150, 100, 343])
    rc = 6.3
    segs = [t for _, _, t in boxes]
    rs = _lib.DeviceSystem(device_model, segments=segs)
    rs.set_boxes(np.stack([b for _, b, _ in boxes]))
    rs.set_positions(np.concatenate([_f32(c) for c, _, _ in boxes]))
    knbr = rs.nlist_allocate(rc, 1.2)
    e, f, v, _ = rs.compute('fp32')
    fs = rs.split(f)
    for k, (c, b, t) in enumerate(boxes):
        s = _lib.DeviceSystem(device_model, t, 1)
        s.set_box(b)
        s.set_positions(_f32(c)[None])
        s.nlist_allocate(rc, 1.2, knbr)
        off = rs.seg_offsets
        for j in range(2):
            np.testing.assert_array_equal(rs.nlist_export(j)[off[k]:off[k + 1]], s.nlist_export(j)[0])
        e1, f1, v1, _ = s.compute('fp32')
        assert abs(e[k] - e1[0]) <= 1e-6 * abs(e1[0]) + 1e-6
        fm = np.abs(f1).max()
        assert np.abs(fs[k] - f1[0]).max() <= 1e-5 * fm
        assert np.abs(v[k] - v1[0]).max() <= 1e-5 * np.abs(v1[0]).max()


def test_small_box_images_vs_oracle(device_model, model0):
    """Boxes below 2 rcut held as 2x2x2 supercell segments next to a plain box."""
    from deepmd_jax import _lib
    from deepmd_jax.evaluator import replicate
    from deepmd_jax.serving import image_factors
    print('case small-box images', flush=True)
    params, var = model0
    boxes = _boxes([32, 45, 64], seed0=40)
    segs, coords, sboxes, nimg, ref = [], [], [], [], []
    for c, b, t in boxes:
        c = _f32(c)
        k = image_factors(b, params['rcut'])
        sc, sb, st = replicate(c, b, t, k)
        segs.append(st)
        coords.append(sc)
        sboxes.append(sb)
        nimg.append(int(np.prod(k)))
        E, F, W = O.energy_force_virial(params, var, c, b, t, lattice='auto')
        ref.append((E.item(), F.numpy(), W.numpy(), len(t)))
    assert nimg[0] == 8 and nimg[2] == 1
    rs = _lib.DeviceSystem(device_model, segments=segs)
    rs.set_boxes(np.stack(sboxes))
    rs.set_images(nimg)
    rs.set_positions(np.concatenate(coords))
    rs.nlist_allocate(params['rcut'], 1.2)
    e, f, v, _ = rs.compute('fp32')
    for k, (E, F, W, N) in enumerate(ref):
        fk = rs.split(f)[k][:N]
        assert abs(e[k] - E) / N <= 1e-5 and abs(e[k] - E) <= 1e-5 * abs(E)
        assert np.abs(fk - F).max() <= 1e-4 * max(1.0, np.abs(F).max())
        assert np.abs(v[k].reshape(3, 3) - W).max() <= 1e-4 * np.abs(W).max()


def _requests(seed=7):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.serving import Request
    spec = [('eval', 40, 0), ('md', 40, 60), ('eval', 90, 0), ('md', 120, 90), ('eval', 200, 0), ('md', 75, 40)]
    reqs = []
    for i, (kind, n, steps) in enumerate(spec):
        c, b, t = syn.water_box(n, seed=seed + i, permute_sites=True)
        v = syn.maxwell_boltzmann(t, 300.0, seed=seed + i) if kind == 'md' else None
        reqs.append(Request(rid=i, kind=kind, coord=c, box=b, type_idx=t, nsteps=steps, velocity=v,
                            seed=100 + i, arrival=0.0))
    return reqs


def _compare(a, b):
    assert a.kind == b.kind and a.error == 0 and b.error == 0
    if a.kind == 'eval':
        assert abs(a.energy - b.energy) <= 1e-6 * abs(b.energy) + 1e-6
        assert np.abs(a.force - b.force).max() <= 1e-5 * np.abs(b.force).max()
        assert np.abs(a.virial - b.virial).max() <= 1e-5 * np.abs(b.virial).max()
    else:
        assert np.abs(a.position - b.position).max() <= 1e-4
        assert np.abs(a.velocity - b.velocity).max() <= 1e-4
        assert abs(a.thermostat_work - b.thermostat_work) <= 1e-4 * max(1.0, abs(b.thermostat_work))


def test_serving_alone_vs_mix(device_model):
    """Every request gives the same result alone and batched with the others, including
    MD segments that join and leave a running batch (continuous batching)."""
    from deepmd_jax.serving import Server
    import time
    print('case serving alone vs mix', flush=True)
    reqs = _requests()
    srv = Server(device_model, md_chunk=10)
    try:
        for r in reqs[:-1]:
            srv.submit(r)
        # the last MD request joins a batch that is already running
        t_end = time.monotonic() + 60
        while srv.stats['md_chunks'] < 2 and time.monotonic() < t_end:
            time.sleep(0.001)
        srv.submit(reqs[-1])
        assert srv.wait(len(reqs), timeout=300)
    finally:
        srv.close()
    mixed = dict(srv.results)
    print('stats', srv.stats, flush=True)
    assert srv.stats['max_md_segments'] >= 2 and srv.stats['md_relayouts'] >= 2
    for r in _requests():
        one = Server(device_model, md_chunk=10)
        try:
            one.submit(r)
            assert one.wait(1, timeout=120)
        finally:
            one.close()
        _compare(mixed[r.rid], one.results[r.rid])


def test_serving_md_images_stay_together(device_model):
    """A small MD box held as 8 images: every image follows the primary trajectory."""
    from deepmd_jax.serving import Server, image_factors
    print('case serving images', flush=True)
    r = _requests()[1]
    k = image_factors(r.box, 7.0)
    assert int(np.prod(k)) == 8
    srv = Server(device_model, md_chunk=10, keep_images=True)
    try:
        srv.submit(r)
        assert srv.wait(1, timeout=120)
    finally:
        srv.close()
    res = srv.results[r.rid]
    assert res.error == 0 and np.isfinite(res.position).all()
    # the finished segment's last device state (all 8 images, pulled at completion): every
    # image is the primary shifted by its box multiple
    from deepmd_jax.evaluator import replicate
    N = r.natoms
    sh = replicate(np.zeros((N, 3)), r.box, r.type_idx, k)[0]
    full = res.image_position
    assert full.shape == (8 * N, 3)
    np.testing.assert_allclose(full[:N], res.position, rtol=0, atol=0)
    d = full - np.tile(full[:N], (8, 1)) - sh
    assert np.abs(d).max() <= 1e-4
    assert np.abs(full[:N] - r.coord).max() > 1e-3   # the segment moved


def test_serving_general_cells(device_model, model0):
    """Evaluation requests with (3,3) cells, one wide and one narrow (held as a supercell of
    images), served together with an orthorhombic one: energy, forces and virial against the
    oracle's lattice form at the plain FP32 tolerances."""
    from deepmd_jax import synthetic as syn
    from deepmd_jax.serving import Request, Server
    print('case serving cells', flush=True)
    params, var = model0
    reqs = []
    for i, (n, tilt) in enumerate(((128, (0.2, -0.1, 0.15)), (24, (0.1, 0.05, -0.2)), (64, None))):
        c, b, t = syn.water_box(n, seed=70 + i)
        if tilt is not None:
            H = np.array([[b[0], 0, 0], [tilt[0] * b[1], b[1], 0], [tilt[1] * b[2], tilt[2] * b[2], b[2]]])
            c, b = (c / b) @ H, H
        reqs.append(Request(rid=i, kind='eval', coord=c, box=b, type_idx=t, arrival=0.0))
    srv = Server(device_model)
    try:
        for r in reqs:
            srv.submit(r)
        assert srv.wait(len(reqs), timeout=120)
    finally:
        srv.close()
    for r in reqs:
        res = srv.results[r.rid]
        assert res.error == 0
        E, F, W = O.energy_force_virial(params, var, r.coord, r.box, r.type_idx, lattice='auto')
        N = len(r.type_idx)
        assert abs(res.energy - E.item()) / N <= 1e-5, r.rid
        assert np.abs(res.force - F.numpy()).max() <= 1e-4 * max(1.0, np.abs(F.numpy()).max()), r.rid
        assert np.abs(res.virial - W.numpy()).max() <= 1e-4 * np.abs(W.numpy()).max(), r.rid


def test_serving_request_masses_and_capacity_reuse(device_model):
    """An MD request carrying its own masses (heavy water, D = 2.0141 Da) gives the same end
    state as a server whose default masses are those, and a different one from plain water;
    a stream of evaluation batches reuses the first batch's row capacities (measuring again
    only when a row overflows) and still matches the requests evaluated alone."""
    from deepmd_jax import _lib
    from deepmd_jax.serving import Request, Server
    from deepmd_jax import synthetic as syn
    print('case serving masses / capacity reuse', flush=True)
    c, b, t = syn.water_box(216, seed=31, permute_sites=True)
    v = syn.maxwell_boltzmann(t, 300.0, seed=31)
    heavy = (15.9994, 2.0141)
    ends = []
    for req_m, srv_m in ((heavy, syn.MASSES), (None, heavy), (None, syn.MASSES)):
        srv = Server(device_model, md_chunk=10, masses=srv_m)
        try:
            srv.submit(Request(rid=0, kind='md', coord=c, box=b, type_idx=t, nsteps=30, velocity=v, seed=5,
                               masses=req_m))
            assert srv.wait(1, timeout=120)
        finally:
            srv.close()
        ends.append(srv.results[0].position)
    np.testing.assert_allclose(ends[0], ends[1], rtol=0, atol=1e-5)   # force-atomic order only
    assert np.abs(ends[0] - ends[2]).max() > 1e-4
    # evaluation batches: capacities measured once, reused while they hold
    boxes = _boxes([64, 125, 216, 100, 343, 150], seed0=40)
    srv = Server(device_model)
    try:
        for i, (cc, bb, tt) in enumerate(boxes):
            srv.submit(Request(rid=i, kind='eval', coord=cc, box=bb, type_idx=tt))
            assert srv.wait(i + 1, timeout=120)
    finally:
        srv.close()
    st = srv.stats
    assert st['eval_batches'] == len(boxes) and st['eval_capacity_measurements'] < st['eval_batches']
    for i, (cc, bb, tt) in enumerate(boxes[2:4]):
        s = _lib.DeviceSystem(device_model, tt, 1)
        s.set_box(bb)
        s.set_positions(cc[None])
        s.nlist_allocate(6.0, 1.2)
        e1, f1, _, _ = s.compute('fp32')
        r = srv.results[i + 2]
        assert abs(r.energy - e1[0]) <= 1e-6 * abs(e1[0]) + 1e-6
        assert np.abs(r.force - f1[0]).max() <= 1e-5 * np.abs(f1).max()

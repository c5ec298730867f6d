"""Serving loop: continuous batching of E/F/virial evaluations and short NVT Langevin
segments of uneven size on one GPU.

The reference evaluates one shape at a time (`evaluate`, deepmd_jax/train.py:703-788, and
`Simulation.run`, deepmd_jax/md.py:874-964, both re-traced per shape).  Here every
request is a *segment* of a ragged device system (`_lib.DeviceSystem(segments=...)`):
one launch of each kernel covers all segments in flight, whatever their sizes.

Two workers share the GPU, each with its own HIP stream, so their kernels overlap:

* the evaluation worker takes every queued evaluation that fits the memory budget,
  lays them out as one ragged batch, builds the neighbour list, evaluates and returns;
* the MD worker runs the NVT Langevin (BAOAB, deepmd_jax/routine.py:46-107) segments in
  chunks of `md_chunk` steps.  Between chunks finished segments leave and queued ones
  join (continuous batching); segments that stay carry positions, velocities,
  thermostat work and their random-stream position into the new layout, so a segment's
  trajectory does not depend on what it was batched with.

Admission is by device memory: a request is admitted when the bytes it needs
(`request_bytes`) fit in the remaining budget; otherwise it waits for running work to
finish.  Boxes shorter than 2 (rcut + skin) are held as k^3-image supercells (the
equivalent of the reference's lattice-image branch, deepmd_jax/utils.py:66-78).
"""
import collections
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
from .evaluator import canonical_box, cell_matrix, cell_widths, replicate
from .synthetic import MASSES


@dataclass
class Request:
    rid: int
    kind: str                       # 'eval' or 'md'
    coord: np.ndarray               # (N, 3) Angstrom
    box: np.ndarray                 # (3,) Angstrom, or a (3,3) cell (rows: lattice vectors)
    type_idx: np.ndarray            # (N,)
    nsteps: int = 0                 # MD steps ('md')
    velocity: np.ndarray = None     # (N, 3) Angstrom/fs ('md')
    seed: int = 0                   # Langevin stream of this request
    arrival: float = 0.0            # seconds after the stream start
    masses: np.ndarray = None       # per-type masses (Dalton) of an 'md' request; None: the server's

    @property
    def natoms(self):
        return len(self.type_idx)


@dataclass
class Result:
    rid: int
    kind: str
    natoms: int
    steps: int
    energy: float = None
    force: np.ndarray = None
    virial: np.ndarray = None
    position: np.ndarray = None
    velocity: np.ndarray = None
    thermostat_work: float = None
    error: int = 0
    t_arrival: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    image_position: np.ndarray = None   # every image of a supercell segment ('md', keep_images=True)

    @property
    def latency(self):
        return self.t_done - self.t_arrival


def image_factors(box3, cutoff):
    """Images per lattice direction so that every supercell width is >= 2 * cutoff."""
    return np.maximum(1, np.ceil(2.0 * cutoff / cell_widths(box3) * (1 + 1e-9))).astype(int)


def request_bytes(natoms_with_images, model_params):
    """Device bytes one segment of this many atoms occupies (work buffers, state, lists)."""
    W = int(model_params['embed_widths'][-1])
    A = int(model_params['axis'])
    F = int(np.ceil(int(model_params['fit_widths'][0]) / 32) * 32)
    per_atom = (4 * 10 * W            # T
                + 2 * 4 * A * W       # G, dG
                + 7 * 4 * F + 4       # fitting activations, per-row energy
                + 4 * 8 * 3 + 8 + 16 + 12 + 8 + 4  # pos, mom (+saves), mass, fp32 mirror, force, scratch, pairs
                + 16 + 12             # canonical coordinates, cell bookkeeping
                + 4 * 4 * 80)         # neighbour rows (about 80 per type at 1.2x of water at rc + 1 A)
    return int(natoms_with_images * per_atom * 1.25) + 64 * 1024


class _Seg:
    """An MD request in flight: its supercell state between chunks."""

    def __init__(self, req, k, masses, t_admit):
        self.req = req
        self.k = k
        self.nimg = int(np.prod(k))
        c, sbox, stype = replicate(np.asarray(req.coord, np.float64), req.box, req.type_idx, k)
        self.coord, self.box, self.types = c, sbox, np.asarray(stype, np.int32)
        self.vel = np.tile(np.asarray(req.velocity, np.float64), (self.nimg, 1))
        m = masses if req.masses is None else req.masses
        self.mass = np.asarray(m, np.float64)[self.types] * MASS_UNIT_CONVERSION
        self.thermo = np.zeros(64)
        self.step = 0
        self.remaining = int(req.nsteps)
        self.t_admit = t_admit


class Server:
    """Request loop over one DeviceModel.

    submit() is thread-safe; results land in `results` (rid -> Result).  close() drains
    the queues and stops the workers.
    """

    def __init__(self, model, *, precision='fp32', hbm_budget=16 << 30, max_batch_atoms=600_000, md_chunk=20,
                 dt=0.5, temperature=300.0, tau_t=100.0, skin=1.0, masses=MASSES, seed=0, clock=None,
                 keep_images=False):
        self.model = model
        self.params = model.params
        self.rcut = float(model.rcut)
        self.precision = precision
        self.budget = int(hbm_budget)
        self.max_batch_atoms = int(max_batch_atoms)
        self.md_chunk = int(md_chunk)
        self.skin = float(skin)
        self.masses = masses
        self.keep_images = bool(keep_images)
        self._eval_knbr = None   # capacities of the last evaluation batch, reused while they hold
        self.md_params = _lib.MDParams()
        self.md_params.routine = _lib.ROUTINE['NVT_langevin']
        self.md_params.precision = _lib.PREC[precision]
        self.md_params.dt = float(dt)
        self.md_params.kT = float(temperature) * TEMP_UNIT_CONVERSION
        self.md_params.gamma = 1.0 / float(tau_t)
        self.md_params.tau = float(tau_t)
        self.md_params.chain_length = 1
        self.md_params.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._clock = clock if clock is not None else time.perf_counter
        self.t0 = self._clock()
        self._cv = threading.Condition()
        self._eval_q = collections.deque()
        self._md_q = collections.deque()
        self._in_flight = 0
        self._closing = False
        self.results = {}
        self.errors = []
        self.stats = {'eval_batches': 0, 'md_chunks': 0, 'md_relayouts': 0, 'max_in_flight_bytes': 0,
                      'max_md_segments': 0, 'max_eval_batch': 0, 'eval_requests': 0,
                      'eval_capacity_measurements': 0}
        self._eval_sys = None
        self._md_sys = None
        self._threads = [threading.Thread(target=self._guard, args=(self._eval_worker,), daemon=True),
                         threading.Thread(target=self._guard, args=(self._md_worker,), daemon=True)]
        for t in self._threads:
            t.start()

    # ------------------------------------------------------------------ public
    def now(self):
        return self._clock() - self.t0

    def cost(self, req):
        cut = self.rcut + (self.skin if req.kind == 'md' else 0.0)
        k = image_factors(canonical_box(req.box), cut)
        return request_bytes(req.natoms * int(np.prod(k)), self.params)

    def submit(self, req):
        if req.kind not in ('eval', 'md'):
            raise ValueError("request kind must be 'eval' or 'md'")
        if req.kind == 'md' and (req.velocity is None or req.nsteps < 1):
            raise ValueError('MD requests need velocities and nsteps >= 1')
        if self.cost(req) > self.budget:
            raise ValueError('request %d needs %d bytes, above the budget of %d' % (req.rid, self.cost(req),
                                                                                   self.budget))
        req.box = canonical_box(req.box)
        with self._cv:
            if self._closing:
                raise RuntimeError('server is closed')
            (self._eval_q if req.kind == 'eval' else self._md_q).append((req, self.now()))
            self._cv.notify_all()

    def close(self, timeout=None):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for t in self._threads:
            t.join(timeout)
        if self.errors:
            raise self.errors[0]

    def wait(self, n_results, timeout=None):
        """Block until n_results requests have completed."""
        end = None if timeout is None else time.monotonic() + timeout
        with self._cv:
            while len(self.results) < n_results and not self.errors:
                rem = None if end is None else end - time.monotonic()
                if rem is not None and rem <= 0:
                    return False
                self._cv.wait(rem if rem is not None else 0.5)
        if self.errors:
            raise self.errors[0]
        return True

    # --------------------------------------------------------------- internals
    def _guard(self, fn):
        try:
            fn()
        except BaseException as e:  # surfaced by wait()/close()
            with self._cv:
                self.errors.append(e)
                self._cv.notify_all()

    def _admit(self, q, limit_atoms, take_all):
        """Pop queued requests (FIFO) that fit the memory budget; called under the lock."""
        out = []
        atoms = 0
        while q:
            req, t_arr = q[0]
            b = self.cost(req)
            if self._in_flight + b > self.budget or (out and atoms + req.natoms > limit_atoms):
                break
            q.popleft()
            self._in_flight += b
            self.stats['max_in_flight_bytes'] = max(self.stats['max_in_flight_bytes'], self._in_flight)
            out.append((req, t_arr, b))
            atoms += req.natoms
            if not take_all:
                break
        return out

    def _release(self, nbytes):
        self._in_flight -= nbytes
        self._cv.notify_all()

    def _eval_worker(self):
        while True:
            with self._cv:
                while True:
                    batch = self._admit(self._eval_q, self.max_batch_atoms, True)
                    if batch or (self._closing and not self._eval_q):
                        break
                    self._cv.wait(0.05)
            if not batch:
                return
            t_admit = self.now()
            res = self._run_eval(batch, t_admit)
            with self._cv:
                for (req, t_arr, b), r in zip(batch, res):
                    r.t_done = self.now()
                    self.results[req.rid] = r
                    self._release(b)
                self.stats['eval_batches'] += 1
                self.stats['eval_requests'] += len(batch)
                self.stats['max_eval_batch'] = max(self.stats['max_eval_batch'], len(batch))

    def _run_eval(self, batch, t_admit):
        segs, coords, boxes, nimg, ks = [], [], [], [], []
        for req, _, _ in batch:
            k = image_factors(req.box, self.rcut)
            c, sbox, st = replicate(np.asarray(req.coord, np.float64), req.box, req.type_idx, k)
            segs.append(np.asarray(st, np.int32))
            coords.append(c)
            boxes.append(sbox)
            nimg.append(int(np.prod(k)))
            ks.append(k)
        if self._eval_sys is None:
            self._eval_sys = _lib.DeviceSystem(self.model, segments=segs)
        else:
            self._eval_sys.reshape(segs)
        s = self._eval_sys
        s.set_cells(np.stack([cell_matrix(b) for b in boxes]))
        s.set_images(nimg)
        s.set_positions(np.concatenate(coords))
        # reuse the last batch's row capacities (no measuring pass, no host round trip for it);
        # measure again only when a row overflowed them
        if self._eval_knbr is not None:
            s.nlist_allocate(self.rcut, 1.2, knbr_override=self._eval_knbr)
        if self._eval_knbr is None or s.nlist_overflow():
            self._eval_knbr = list(s.nlist_allocate(self.rcut, 1.2))
            self.stats['eval_capacity_measurements'] += 1
        e, f, v, _ = s.compute(self.precision)
        out = []
        for i, ((req, t_arr, _), fs) in enumerate(zip(batch, s.split(f))):
            out.append(Result(rid=req.rid, kind='eval', natoms=req.natoms, steps=1, energy=float(e[i]),
                              force=fs[:req.natoms].copy(), virial=v[i].copy(), t_arrival=t_arr, t_admit=t_admit))
        return out

    def _md_worker(self):
        active = []
        host_stale = False   # device state of `active` is newer than the host copies
        while True:
            with self._cv:
                while True:
                    new = self._admit(self._md_q, self.max_batch_atoms, True)
                    if new or active or (self._closing and not self._md_q):
                        break
                    self._cv.wait(0.05)
            if not new and not active:
                return
            if new:
                if host_stale:
                    self._md_pull(active)
                    host_stale = False
                t_admit = self.now()
                for req, t_arr, b in new:
                    sg = _Seg(req, image_factors(req.box, self.rcut + self.skin), self.masses, t_admit)
                    sg.t_arrival, sg.nbytes = t_arr, b
                    active.append(sg)
                self._md_layout(active)
            s = self._md_sys
            nrun = min(self.md_chunk, min(sg.remaining for sg in active))
            _, _, err = s.md_run(self.md_params, nrun, record=False)
            self.stats['md_chunks'] += 1
            if err & 2:
                # neighbour capacity exceeded: the chunk was rolled back; grow and re-run
                s.nlist_allocate(self.rcut + self.skin, 1.2)
                continue
            host_stale = True
            for sg in active:
                sg.remaining -= nrun
            failed = bool(err & 16)
            done = [sg for sg in active if sg.remaining == 0 or failed]
            if not done:
                continue
            self._md_pull(active)
            host_stale = False
            finished = [(sg, Result(rid=sg.req.rid, kind='md', natoms=sg.req.natoms, steps=sg.req.nsteps,
                                    position=sg.coord[:sg.req.natoms].copy(), velocity=sg.vel[:sg.req.natoms].copy(),
                                    thermostat_work=float(sg.thermo[0]), error=int(err), t_arrival=sg.t_arrival,
                                    t_admit=sg.t_admit,
                                    image_position=sg.coord.copy() if self.keep_images else None)) for sg in done]
            active = [sg for sg in active if sg.remaining > 0 and not failed]
            if active:
                self._md_layout(active)
            with self._cv:
                for sg, r in finished:
                    r.t_done = self.now()
                    self.results[sg.req.rid] = r
                    self._release(sg.nbytes)

    def _md_layout(self, active):
        segs = [sg.types for sg in active]
        if self._md_sys is None:
            self._md_sys = _lib.DeviceSystem(self.model, segments=segs)
        else:
            self._md_sys.reshape(segs)
        s = self._md_sys
        s.set_cells(np.stack([cell_matrix(sg.box) for sg in active]))
        s.set_images([sg.nimg for sg in active])
        s.set_masses(np.concatenate([sg.mass for sg in active]))
        s.set_positions(np.concatenate([sg.coord for sg in active]))
        s.set_velocities(np.concatenate([sg.vel for sg in active]))
        s.thermostat_state(np.stack([sg.thermo for sg in active]))
        s.set_rng([sg.req.seed for sg in active], [sg.step for sg in active])
        s.nlist_allocate(self.rcut + self.skin, 1.2)
        self.stats['md_relayouts'] += 1
        self.stats['max_md_segments'] = max(self.stats['max_md_segments'], len(active))

    def _md_pull(self, active):
        s = self._md_sys
        pos = s.split(s.get_positions())
        vel = s.split(s.get_velocities())
        th = s.thermostat_state()
        _, steps = s.get_rng()
        for i, sg in enumerate(active):
            sg.coord, sg.vel, sg.thermo, sg.step = pos[i].copy(), vel[i].copy(), th[i].copy(), int(steps[i])


def request_stream(n_requests=256, seed=4, n_min=32, n_max=8192, md_fraction=0.3, md_steps=200, rate=10.0,
                   temperature=300.0):
    """The synthetic request stream: water boxes at 1 g/cm^3 (molecules on a ceil(n^(1/3))^3
    grid, first n sites of a seeded permutation), n log-uniform in [n_min, n_max],
    `md_fraction` NVT Langevin segments of md_steps steps, Poisson arrivals at `rate`/s."""
    from .synthetic import water_box, maxwell_boltzmann
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_requests)).astype(int)
    kinds = np.where(rng.uniform(size=n_requests) < md_fraction, 'md', 'eval')
    gaps = rng.exponential(1.0 / rate, n_requests)
    t = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i in range(n_requests):
        c, box, ty = water_box(int(n[i]), seed=1000 + i, permute_sites=True)
        v = maxwell_boltzmann(ty, temperature, seed=1000 + i) if kinds[i] == 'md' else None
        reqs.append(Request(rid=i, kind=str(kinds[i]), coord=c, box=box, type_idx=ty,
                            nsteps=md_steps if kinds[i] == 'md' else 0, velocity=v, seed=1000 + i,
                            arrival=float(t[i])))
    return reqs


def replay(server, requests, time_scale=1.0):
    """Submit requests at their arrival times (scaled) and wait for all results."""
    t_start = server.now()
    for r in sorted(requests, key=lambda q: q.arrival):
        delay = t_start + r.arrival * time_scale - server.now()
        if delay > 0:
            time.sleep(delay)
        server.submit(r)
    server.wait(len(requests))
    return [server.results[r.rid] for r in requests]


def summarize(results, t_span=None):
    lat = np.array([r.latency for r in results])
    work = sum(r.natoms * r.steps for r in results)
    t0 = min(r.t_arrival for r in results)
    t1 = max(r.t_done for r in results)
    span = t_span if t_span is not None else (t1 - t0)
    out = {'requests': len(results), 'p50_latency_s': float(np.percentile(lat, 50)),
           'p99_latency_s': float(np.percentile(lat, 99)), 'mean_latency_s': float(lat.mean()),
           'makespan_s': float(t1 - t0), 'atoms_steps_per_s': float(work / max(span, 1e-9))}
    for kind in ('eval', 'md'):
        lk = np.array([r.latency for r in results if r.kind == kind])
        if len(lk):
            out['p50_%s_s' % kind] = float(np.percentile(lk, 50))
            out['p99_%s_s' % kind] = float(np.percentile(lk, 99))
    return out

"""MD steps/s on the 98,304-atom water box (32,768 H2O, L = 99.34 A) with the seeded
se_a model (embed 25/50/100, fit 240x3, axis 16, rcut 6 A), NVE velocity Verlet at
dt = 0.5 fs, 1 A neighbour skin, one MI355X.

`value` is device-resident throughput: positions, momenta, neighbour lists and model
stay in HBM; each timed step is the drift test + conditional neighbour rebuild, the
integrator, and one full energy/force/virial evaluation.  The Simulation.run rate
(host trajectory copies every chunk) is reported beside it as `simulation_run_steps_per_s`.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'deepmd-jax_amd'))

PEAK_FP32_MFMA = 157.3e12
PEAK_BF16_MFMA = 2.5e15
PEAK_HBM = 8.0e12
KERNELS = {0: 'nlist', 1: 'embed_fwd', 2: 'fit', 3: 'embed_bwd', 4: 'integrator'}
# HBM bytes per launch of the FP32 embed_bwd kernel on the 98,304-atom box, from two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE in KiB per dispatch) scaled by what the
# same counters read for a 256 MiB elementwise copy (x2.0 and x1.0): 1,075.7 MiB read +
# 471.5 MiB written (profiles/embed_fit_nlist_water98k_trace.md).  Counters cannot be read
# without the profiler, so the figure is carried here and re-measured with the kernel.
EMBED_BWD_TRAFFIC_FP32 = (1075.7 + 471.5) * 2 ** 20


def flops_per_step(N, pairs, A=16, W=100, emb=(25, 50, 100), fit=240):
    """Algorithmic FLOPs of one energy+force evaluation (no recompute, no weight grads):
    2 (forward + input-gradient) x 2 FLOP/MAC x N x [pairs x (embedding + reducer MACs)
    + descriptor MACs + fitting MACs]."""
    e_mac = emb[0] + emb[0] * emb[1] + emb[1] * emb[2]
    red_mac = 10 * W
    g_mac = 10 * A * W
    f_mac = A * W * fit + 2 * fit * fit + fit
    per_kernel = {
        'embed_fwd': 2.0 * N * (pairs * (e_mac + red_mac) + g_mac),
        'embed_bwd': 2.0 * N * (pairs * (e_mac + red_mac) + g_mac),
        'fit': 2.0 * 2.0 * N * f_mac,
    }
    return 2 * 2.0 * N * (pairs * (e_mac + red_mac) + g_mac + f_mac), per_kernel


def cpu_baseline(params, var, seconds_budget=30.0):
    """The C++/OpenMP FP32 CPU restatement of the path (oracle/dp_cpu.cpp: descriptor, fitting
    net, hand-written backward, one thread per centre block) timed on this host's cores on a
    bounded sample: 4,096 H2O (12,288 atoms), E+F+virial with a neighbour list; steps/s for
    the 98,304-atom box extrapolated linearly in N (the work is per atom)."""
    sys.path.insert(0, ROOT)
    from oracle import dp_oracle as O, dp_cpu
    from deepmd_jax import synthetic as syn
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault('OMP_NUM_THREADS', str(ncores))
    if not os.path.exists(dp_cpu.LIB):
        dp_cpu.build()
    coord, box, types = syn.water_box(4096, seed=0)
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'])
    pk = dp_cpu.Packed(params, var)
    rows, ks = dp_cpu.neighbour_rows(nb, types)
    dp_cpu.energy_force(pk, coord, box, types, rows, ks, 'f32')  # warm-up (thread pool, pages)
    times = []
    t_start = time.time()
    while len(times) < 5 and (time.time() - t_start) < seconds_budget:
        t0 = time.time()
        dp_cpu.energy_force(pk, coord, box, types, rows, ks, 'f32')
        times.append(time.time() - t0)
    t = float(np.median(times))
    scale = 98304 / 12288
    return {'value': 1.0 / (t * scale), 'unit': 'MD steps/s', 'cores': dp_cpu.threads(), 'kind': 'port',
            'sample': '12,288-atom water box (4,096 H2O, seed 0), %d E+F+virial evaluations of the C++/OpenMP FP32 '
                      'restatement (oracle/dp_cpu.cpp), median %.3f s, x8 atoms extrapolated to 98,304; '
                      'host cores %d' % (len(times), t, ncores)}


def accuracy(params, var, model, frames=1):
    """Energy/atom, force RMSE and virial error of fp32 / bf16 / fp16 against the FP64 CPU
    oracle on `frames` seeded 12,288-atom water frames (4,096 H2O, seeds 0, 1, ...; the
    FP64 autograd oracle on the full 98,304-atom box needs tens of GB per frame).  Errors
    are the worst frame's value (max) and the mean over frames."""
    import torch
    sys.path.insert(0, ROOT)
    from oracle import dp_oracle as O
    from deepmd_jax import _lib, synthetic as syn
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    keys = ('energy_per_atom_abs_err', 'force_rmse_eV_per_A', 'force_max_abs_err', 'virial_max_rel_err')
    per = {prec: {k: [] for k in keys} for prec in ('fp32', 'bf16', 'fp16')}
    frms = []
    s = None
    for fr in range(frames):
        coord, box, types = syn.water_box(4096, seed=fr)
        nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'])
        E, F, W = O.energy_force_virial(params, var, coord, box, types, nbrs_nm=nb, dtype=torch.float64)
        E, F, W = E.item(), F.numpy(), W.numpy()
        if s is None:
            s = _lib.DeviceSystem(model, types, 1)
            s.set_box(box)
        s.set_positions(coord)
        s.nlist_allocate(params['rcut'], 1.2)
        frms.append(float(np.sqrt((F ** 2).mean())))
        print('# accuracy frame %d done' % fr, file=sys.stderr, flush=True)
        for prec in per:
            e, f, v, _ = s.compute(prec)
            vals = (abs(e[0] - E) / len(types), float(np.sqrt(((f[0] - F) ** 2).mean())),
                    float(np.abs(f[0] - F).max()), float(np.abs(v[0] - W).max() / np.abs(W).max()))
            for k, x in zip(keys, vals):
                per[prec][k].append(x)
    res = {'frames': '%d x 4,096 H2O (12,288 atoms), seeds 0..%d' % (frames, frames - 1)}
    for prec in per:
        res[prec] = {k: float(np.max(v)) for k, v in per[prec].items()}
        res[prec]['mean'] = {k: float(np.mean(v)) for k, v in per[prec].items()}
    res['force_rms_eV_per_A'] = float(np.mean(frms))
    return res


def nve_drift(model, params, coord, box, types, vel, steps, sample=100, precisions=('fp32', 'bf16'), equil_steps=0):
    """Total-energy drift of NVE velocity Verlet (dt 0.5 fs) from the same initial state
    per precision: KE + PE sampled every `sample` steps (PE evaluated in the same
    precision), least-squares slope in eV/atom/ps over the whole run and over the first
    2,000 steps.  equil_steps > 0: the initial state is first equilibrated in FP32 by NVT
    Langevin (BAOAB, 300 K, tau_t 100 fs) for that many steps, so the NVE runs start from
    a state at 300 K under the seeded potential instead of the water grid (whose excess
    potential energy under the random-init potential otherwise turns into heat)."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    N = len(types)
    out = {'steps': steps, 'sample_every': sample, 'dt_fs': 0.5, 'equilibration_steps': equil_steps}
    if equil_steps:
        s = _lib.DeviceSystem(model, types, 1)
        s.set_box(box)
        s.set_masses(m)
        s.set_positions(coord)
        s.set_velocities(vel)
        s.nlist_allocate(params['rcut'] + 1.0, 1.2)
        p = _lib.MDParams()
        p.routine, p.dt, p.kT, p.gamma, p.seed = _lib.ROUTINE['NVT_langevin'], 0.5, 300 * TEMP_UNIT_CONVERSION, 0.01, 11
        done, Teq = 0, []
        while done < equil_steps:
            _, _, err = s.md_run(p, 100, record=False)
            if err & 2:
                s.nlist_allocate(params['rcut'] + 1.0, 1.3)
                continue
            if err & 16:
                raise RuntimeError('NaN/Inf during the NVT equilibration')
            done += 100
            if done % 1000 == 0:
                Teq.append(2 * float(s.kinetic_energy()[0]) / (3 * N) / TEMP_UNIT_CONVERSION)
                print('# drift equilibration step %d T %.1f K' % (done, Teq[-1]), file=sys.stderr, flush=True)
        coord, vel = s.get_positions()[0], s.get_velocities()[0]
        out['equilibration_temperature_K'] = Teq
        del s
    for prec in precisions:
        s = _lib.DeviceSystem(model, types, 1)
        s.set_box(box)
        s.set_masses(m)
        s.set_positions(coord)
        s.set_velocities(vel)
        s.nlist_allocate(params['rcut'] + 1.0, 1.2)
        p = _lib.MDParams()
        p.routine, p.dt, p.precision = 0, 0.5, _lib.PREC[prec]
        t, H, T = [], [], []
        step, realloc, t0 = 0, 0, time.perf_counter()
        while True:
            e, _, _, _ = s.compute(prec)
            ke = float(s.kinetic_energy()[0])
            t.append(step * 0.5e-3)
            H.append(float(e[0]) + ke)
            T.append(2 * ke / (3 * N) / 8.617333262e-5)
            if step >= steps:
                break
            _, _, err = s.md_run(p, sample, record=False)
            if err & 2:
                realloc += 1
                s.nlist_allocate(params['rcut'] + 1.0, 1.2 + 0.05 * realloc)
                continue
            if err & 16:
                break
            step += sample
            if step % 1000 == 0:
                print('# drift %s step %d H %.4f' % (prec, step, H[-1]), file=sys.stderr, flush=True)
        t, H = np.array(t), np.array(H)
        k = t <= 1.0 + 1e-9
        out[prec] = {'drift_eV_per_atom_per_ps': float(np.polyfit(t, H, 1)[0] / N),
                     'drift_first_1ps_eV_per_atom_per_ps': float(np.polyfit(t[k], H[k], 1)[0] / N),
                     'energy_range_eV_per_atom': float((H.max() - H.min()) / N),
                     'temperature_K_start_end': [T[0], T[-1]], 'steps_done': step,
                     'nonfinite': bool(step < steps), 'nlist_reallocations': realloc,
                     'wall_s': time.perf_counter() - t0}
        print('# drift %s: %s' % (prec, out[prec]), file=sys.stderr, flush=True)
    return out


def pimd(model, params, var, steps=100, warmup=10, precisions=('fp32', 'bf16'), acc_beads=4):
    """Ring-polymer NVT (PILE-L, 300 K, dt 0.25 fs): 32 beads x 4,096 H2O (12,288 atoms per
    bead, L = 49.7 A), per-bead +-0.05 A uniform jitter from seed 3, all beads in one
    system (one launch per kernel for all 393,216 atoms).  Device-resident bead*steps/s per
    precision, and the centroid-force error against the FP64 oracle over the first
    `acc_beads` beads of the initial configuration."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    from deepmd_jax.pimd import normal_mode_transform, pile_l_factors
    P, dt = 32, 0.25
    coord, box, types = syn.water_box(4096, seed=3)
    rng = np.random.default_rng(3)
    x = coord[None] + rng.uniform(-0.05, 0.05, (P,) + coord.shape)
    m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
    kT = 300 * TEMP_UNIT_CONVERSION
    v = rng.standard_normal(x.shape) * np.sqrt(kT * P / m)[None, :, None]
    freqs, C = normal_mode_transform(P, kT)
    _, c1, c2 = pile_l_factors(freqs, dt, kT, P, 100.0)
    s = _lib.DeviceSystem(model, types, P)
    s.set_box(box)
    s.set_masses(m)
    s.pimd_setup(C, freqs, c1, c2)
    p = _lib.MDParams()
    p.routine, p.dt, p.kT, p.seed, p.gamma, p.tau = _lib.ROUTINE['PIMD'], dt, kT, 3, 0.01, 100.0
    N = len(types)
    out = {'workload': '32 beads x 4,096 H2O (12,288 atoms per bead, L=%.2f A), PILE-L 300 K, dt 0.25 fs, '
                       'skin 1 A' % box[0], 'atoms_total': P * N}

    def reset():
        s.set_box(box)
        s.set_positions(x)
        s.set_velocities(v)
        s.nlist_allocate(params['rcut'] + 1.0, 1.2)

    def run(n):
        done = 0
        while done < n:
            k = min(50, n - done)
            _, _, err = s.md_run(p, k, record=False)
            if err & 2:
                s.nlist_allocate(params['rcut'] + 1.0, 1.25)
                continue
            if err & 16:
                raise RuntimeError('non-finite state in the PIMD run')
            done += k
    for prec in precisions:
        p.precision = _lib.PREC[prec]
        reset()
        run(warmup)
        s.synchronize()
        t0 = time.perf_counter()
        run(steps)
        s.synchronize()
        el = time.perf_counter() - t0
        out[prec] = {'steps_per_s': steps / el, 'bead_steps_per_s': P * steps / el,
                     'atoms_steps_per_s': P * N * steps / el, 'ms_per_step': 1e3 * el / steps, 'steps': steps}
    # path-integral NPT on the same beads (1 bar, isotropic, tau_p 1000 fs, Langevin barostat)
    from deepmd_jax.constants import PRESS_UNIT_CONVERSION
    p.routine, p.precision, p.gamma = _lib.ROUTINE['PIMD_NPT'], _lib.PREC['fp32'], 1.0 / 1000.0
    p.pressure, p.tau_p = 1.0 * PRESS_UNIT_CONVERSION, 1000.0
    for a in range(3):
        p.couple[a] = 0
    reset()
    run(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    run(steps)
    s.synchronize()
    el = time.perf_counter() - t0
    out['npt_fp32'] = {'steps_per_s': steps / el, 'bead_steps_per_s': P * steps / el,
                       'ms_per_step': 1e3 * el / steps, 'steps': steps,
                       'box_rel_change': float(s.get_boxes()[0][0] / box[0] - 1.0)}
    p.routine, p.gamma = _lib.ROUTINE['PIMD'], 0.01
    if acc_beads:
        import torch
        sys.path.insert(0, ROOT)
        from oracle import dp_oracle as O
        torch.set_num_threads(len(os.sched_getaffinity(0)))
        reset()
        Fo = []
        for b in range(acc_beads):
            nb, _, _ = O.typed_nlist(x[b], box, types, params['rcut'])
            Fo.append(O.energy_force_virial(params, var, x[b], box, types, nbrs_nm=nb,
                                            dtype=torch.float64)[1].numpy())
        Fc = np.mean(Fo, 0)
        out['centroid_force'] = {'beads': acc_beads, 'oracle': 'FP64, per bead',
                                 'centroid_force_rms_eV_per_A': float(np.sqrt((Fc ** 2).mean()))}
        for prec in precisions:
            _, f, _, _ = s.compute(prec)
            fc = f[:acc_beads].mean(0)
            out['centroid_force'][prec] = {'rmse_eV_per_A': float(np.sqrt(((fc - Fc) ** 2).mean())),
                                           'max_abs_err': float(np.abs(fc - Fc).max())}
    return out


def compressed(params, var, coord, box, types, vel, steps, warmup, acc=True):
    """The same NVE workload with the model after compress_model (1,024 intervals, rmin
    0.6 A, the training defaults): tabulated embedding nets (FP32) with FP32 and BF16
    fitting GEMMs.  Device-resident steps/s, per-kernel ms per step, and (acc) the errors
    against the FP64 oracle of the same tabulated model on a 12,288-atom frame."""
    import contextlib
    import copy
    import io
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import compress_model
    pc, vc = copy.deepcopy(params), copy.deepcopy(var)
    with contextlib.redirect_stdout(io.StringIO()):
        m, vc = compress_model(DPModel(pc), vc, 1024, 0.6)
    pc = m.params
    model = _lib.DeviceModel(pc, vc)
    s = _lib.DeviceSystem(model, types, 1)
    s.set_box(box)
    s.set_masses(np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)
    p = _lib.MDParams()
    p.routine, p.dt = 0, 0.5
    out = {'model': 'seeded model after compress_model(Ngrids=1024, rmin=0.6)'}

    def run(n):
        done = 0
        while done < n:
            k = min(100, n - done)
            _, _, err = s.md_run(p, k, record=False)
            if err & 16:
                raise RuntimeError('NaN/Inf in the compressed-model run')
            if err & 2:
                s.nlist_allocate(pc['rcut'] + 1.0, 1.3)
                continue
            done += k
        s.synchronize()
    for prec in ('fp32', 'bf16'):
        p.precision = _lib.PREC[prec]
        s.set_positions(coord)
        s.set_velocities(vel)
        s.nlist_allocate(pc['rcut'] + 1.0, 1.2)
        run(warmup)
        t0 = time.perf_counter()
        run(steps)
        el = time.perf_counter() - t0
        s.timing(on=True)
        run(20)
        kt = {KERNELS[w]: s.timing(which=w)[0] / 20 for w in KERNELS}
        s.timing(on=False)
        out[prec] = {'steps_per_s': steps / el, 'ms_per_step': 1e3 * el / steps,
                     'atoms_steps_per_s': steps * len(types) / el, 'kernel_ms_per_step': kt}
        print('# compressed %s: %.1f steps/s' % (prec, steps / el), file=sys.stderr, flush=True)
    if acc:
        import torch
        sys.path.insert(0, ROOT)
        from oracle import dp_oracle as O
        torch.set_num_threads(len(os.sched_getaffinity(0)))
        c2, b2, t2 = syn.water_box(4096, seed=0)
        nb, _, _ = O.typed_nlist(c2, b2, t2, pc['rcut'])
        E, F, W = O.energy_force_virial(pc, vc, c2, b2, t2, nbrs_nm=nb, dtype=torch.float64)
        E, F, W = E.item(), F.numpy(), W.numpy()
        s2 = _lib.DeviceSystem(model, t2, 1)
        s2.set_box(b2)
        s2.set_positions(c2)
        s2.nlist_allocate(pc['rcut'], 1.2)
        out['errors_vs_fp64_oracle'] = {'frame': '4,096 H2O (12,288 atoms), seed 0, same tabulated model'}
        for prec in ('fp32', 'bf16'):
            e, f, v, _ = s2.compute(prec)
            out['errors_vs_fp64_oracle'][prec] = {
                'energy_per_atom_abs_err': abs(e[0] - E) / len(t2),
                'force_rmse_eV_per_A': float(np.sqrt(((f[0] - F) ** 2).mean())),
                'virial_max_rel_err': float(np.abs(v[0] - W).max() / np.abs(W).max())}
    return out


def npt(model, params, coord, box, types, vel, steps, warmup):
    """NPT (MTK Nose-Hoover barostat + Nose-Hoover chain thermostat, isotropic coupling) at
    300 K and 1 bar on the same 98,304-atom box, tau_t 100 fs, tau_p 1000 fs, skin 1 A:
    device-resident steps/s and the relative box change over the timed steps."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION, PRESS_UNIT_CONVERSION, TEMP_UNIT_CONVERSION
    s = _lib.DeviceSystem(model, types, 1)
    s.set_box(box)
    s.set_masses(np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.nlist_allocate(params['rcut'] + 1.0, 1.2)
    p = _lib.MDParams()
    p.routine, p.dt, p.kT, p.tau, p.chain_length = _lib.ROUTINE['NPT'], 0.5, 300 * TEMP_UNIT_CONVERSION, 100.0, 3
    p.pressure, p.tau_p, p.chain_length_p = 1.0 * PRESS_UNIT_CONVERSION, 1000.0, 3
    for a in range(3):
        p.couple[a] = 0
    out = {'workload': 'NPT 300 K, 1 bar, isotropic, 98,304-atom water, tau_t 100 fs, tau_p 1000 fs, skin 1 A'}

    def run(n):
        done = 0
        while done < n:
            k = min(100, n - done)
            _, _, err = s.md_run(p, k, record=False)
            if err & 16:
                raise RuntimeError('NaN/Inf in the NPT run')
            if err & 2:
                s.nlist_allocate(params['rcut'] + 1.0, 1.3)
                continue
            done += k
        s.synchronize()
    run(warmup)
    b0 = s.get_boxes()[0]
    t0 = time.perf_counter()
    run(steps)
    el = time.perf_counter() - t0
    out['fp32'] = {'steps_per_s': steps / el, 'ms_per_step': 1e3 * el / steps,
                   'box_rel_change': float(s.get_boxes()[0][0] / b0[0] - 1.0)}
    print('# npt fp32: %.1f steps/s' % (steps / el), file=sys.stderr, flush=True)
    return out


def published_model(coord, box, types, vel, steps, warmup):
    """The model of the reference's published figure (README: 2.2e6 MD steps/day, 25.5 steps/s, for
    ~100k-atom water on one RTX 4090): its training defaults (embed 32/32/64, fit 128x3, axis 12,
    rcut 6, compressed with 1,024 intervals from rmin 0.6; deepmd_jax/train.py:24-31,49-50), seeded
    here, FP32, on the same 98,304-atom box; device-resident NVE steps/s.  Same model class and
    size as the published run, other hardware; not the headline metric's config."""
    import copy
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.utils import compress_model
    import contextlib
    import io
    params, var = syn.init_model(1, embed_widths=(32, 32, 64), fit_widths=(128, 128, 128), axis=12)
    with contextlib.redirect_stdout(io.StringIO()):
        m, cvar = compress_model(DPModel(copy.deepcopy(params)), copy.deepcopy(var), 1024, 0.6)
    dm = _lib.DeviceModel(m.params, cvar)
    s = _lib.DeviceSystem(dm, types, 1)
    s.set_box(box)
    s.set_masses(np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.nlist_allocate(m.params['rcut'] + 1.0, 1.2)
    p = _lib.MDParams()
    p.routine, p.dt = 0, 0.5
    N = len(types)
    chunk = max(10, min(100, 100000 // N))

    def run(n):
        done = 0
        while done < n:
            k = min(chunk, n - done)
            _, _, err = s.md_run(p, k, record=False)
            if err & 2:
                s.nlist_allocate(m.params['rcut'] + 1.0, 1.25)
                continue
            done += k
        s.synchronize()
    run(warmup)
    t0 = time.perf_counter()
    run(steps)
    el = time.perf_counter() - t0
    sps = steps / el
    print('# published-model leg: %.1f steps/s' % sps, file=sys.stderr, flush=True)
    return {'model': 'embed 32/32/64, fit 128x3, axis 12, rcut 6, compressed (1024 intervals, rmin 0.6), FP32',
            'atoms': N, 'steps_per_s': sps, 'steps_per_day': sps * 86400,
            'published_steps_per_s': 2.2e6 / 86400, 'published_hardware': '1x RTX 4090 (README.md:11)',
            'ratio_to_published': sps / (2.2e6 / 86400)}


def dplr(model, params, coord, box, types, vel, steps, warmup):
    """DPLR NVE on the same 98,304-atom box: the energy model plus a seeded Wannier-centroid
    model (same embedding widths, fitting 100x3, centroids on O), charges O +6, H +1, centroid
    -8, PPPM beta 0.4 at resolution 5 (grid 198^3); device-resident steps/s."""
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    wp, wv = syn.init_wannier_model(4)
    wdm = _lib.DeviceModel(wp, wv)
    s = _lib.DeviceSystem(model, types, 1)
    s.set_box(box)
    s.set_masses(np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)
    s.set_positions(coord)
    s.set_velocities(vel)
    s.nlist_allocate(params['rcut'] + 1.0, 1.2)
    grid = (np.asarray(box) * 0.4 * 5).astype(int)
    s.set_dplr(wdm, [6.0, 1.0], [-8.0, 0.0], 0.4, grid)
    p = _lib.MDParams()
    p.routine, p.dt = _lib.ROUTINE['NVE'], 0.5
    out = {'workload': 'DPLR NVE, 98,304-atom water, centroid model on O, q (6, 1, -8), beta 0.4, grid %s'
           % 'x'.join(str(int(g)) for g in grid)}

    def run(n):
        done = 0
        while done < n:
            k = min(100, n - done)
            _, _, err = s.md_run(p, k, record=False)
            if err & 16:
                raise RuntimeError('NaN/Inf in the DPLR run')
            if err & 2:
                s.nlist_allocate(params['rcut'] + 1.0, 1.3)
                continue
            done += k
        s.synchronize()
    run(warmup)
    t0 = time.perf_counter()
    run(steps)
    el = time.perf_counter() - t0
    out['fp32'] = {'steps_per_s': steps / el, 'ms_per_step': 1e3 * el / steps,
                   'long_range_energy_eV': float(s.dplr_get()[0][0])}
    print('# dplr fp32: %.1f steps/s' % (steps / el), file=sys.stderr, flush=True)
    return out


def serving(model, rates, precisions, n_requests=256, md_chunk=20, budget_gib=32):
    """The serving stream: 256 water-box requests (n log-uniform in [32, 8192] H2O, first n
    sites of a seeded grid permutation), 70 % single E/F/virial evaluations, 30 % 200-step
    NVT Langevin segments at 300 K, dt 0.5 fs; Poisson arrivals at each rate (requests/s).
    Reports p50/p99 latency (arrival -> result) and atoms*steps/s (evaluation = 1 step)."""
    from deepmd_jax.serving import Server, request_stream, replay, summarize
    reqs = request_stream(n_requests, seed=4, rate=1.0)
    out = {'stream': '%d requests, n log-uniform [32, 8192] H2O (96-24,576 atoms), 70%% E/F/virial, '
                     '30%% 200-step NVT Langevin 300 K dt 0.5 fs; Poisson arrivals' % n_requests,
           'hbm_budget_GiB': budget_gib, 'md_chunk_steps': md_chunk,
           'atoms_total': int(sum(r.natoms for r in reqs)),
           'atoms_steps_total': int(sum(r.natoms * max(1, r.nsteps) for r in reqs))}
    for prec in precisions:
        warm = Server(model, precision=prec, md_chunk=md_chunk, hbm_budget=budget_gib << 30)
        replay(warm, reqs[:8], time_scale=0.0)
        warm.close()
        res = {}
        for rate in rates:
            srv = Server(model, precision=prec, md_chunk=md_chunk, hbm_budget=budget_gib << 30)
            t0 = time.perf_counter()
            results = replay(srv, reqs, time_scale=1.0 / rate)
            wall = time.perf_counter() - t0
            srv.close()
            sm = summarize(results)
            sm['wall_s'] = wall
            sm['stats'] = srv.stats
            res['%g' % rate] = sm
            print('# serving %s rate %g/s: p50 %.3f s p99 %.3f s, %.3g atoms*steps/s' %
                  (prec, rate, sm['p50_latency_s'], sm['p99_latency_s'], sm['atoms_steps_per_s']),
                  file=sys.stderr, flush=True)
        out[prec] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--nmol', type=int, default=32768)
    ap.add_argument('--precision', default='fp32')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-sim', action='store_true')
    ap.add_argument('--no-half', action='store_true')
    ap.add_argument('--no-acc', action='store_true')
    ap.add_argument('--no-serving', action='store_true')
    ap.add_argument('--no-pimd', action='store_true')
    ap.add_argument('--no-compressed', action='store_true')
    ap.add_argument('--no-npt', action='store_true')
    ap.add_argument('--no-dplr', action='store_true')
    ap.add_argument('--no-published', action='store_true', help="skip the reference's published-model leg")
    ap.add_argument('--nve-run-steps', type=int, default=2000,
                    help='length of the full NVE run reported beside the timed steps (0: skip)')
    ap.add_argument('--frames', type=int, default=1, help='seeded frames for the precision errors')
    ap.add_argument('--drift-steps', type=int, default=0, help='NVE drift run length per precision (0: skip)')
    ap.add_argument('--drift-equil-steps', type=int, default=0,
                    help='NVT Langevin equilibration steps before the NVE drift runs (0: start from the grid)')
    ap.add_argument('--pimd-steps', type=int, default=100)
    ap.add_argument('--pimd-acc-beads', type=int, default=2)
    ap.add_argument('--serving-rates', default='20', help='comma-separated Poisson arrival rates (requests/s)')
    ap.add_argument('--serving-requests', type=int, default=256)
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='run a plain 256 MiB elementwise copy (read 256 MiB, write 256 MiB) five times, for '
                         'scaling FETCH_SIZE / WRITE_SIZE counters')
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(64 * 1024 * 1024, device='cuda')
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_read': x.numel() * 4, 'copy_bytes_written': y.numel() * 4, 'launches': 5}))
        return
    if args.gpus != 1:
        print('bench.py runs on one GPU; multi-GPU runs are not part of this build (--gpus %d requested)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION

    params, var = syn.init_model(0)
    coord, box, types = syn.water_box(args.nmol, seed=0)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=0)
    N = len(types)
    t_setup = time.perf_counter()
    model = _lib.DeviceModel(params, var)
    sys_ = _lib.DeviceSystem(model, types, 1)
    sys_.set_box(box)
    sys_.set_masses(np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)

    def reset():
        """Every timed pass starts from the same initial positions and velocities."""
        sys_.set_positions(coord)
        sys_.set_velocities(vel)
        sys_.nlist_allocate(params['rcut'] + 1.0, 1.2)

    reset()
    p = _lib.MDParams()
    p.routine = 0
    p.precision = _lib.PREC[args.precision]
    p.dt = 0.5
    # first chunk, reported separately (the reference's compile + first chunk): model upload,
    # system set-up, neighbour-list allocation, first force evaluation and the first 10 steps
    sys_.md_run(p, 10, record=False)
    sys_.synchronize()
    first_chunk_s = time.perf_counter() - t_setup
    reset()
    # warmup (includes first force evaluation)
    done = 0
    while done < args.warmup:
        k = min(50, args.warmup - done)
        _, _, err = sys_.md_run(p, k, record=False)
        if err & 2:
            sys_.nlist_allocate(params['rcut'] + 1.0, 1.25)
            continue
        done += k
    sys_.synchronize()

    # chunks of the reference's run protocol (deepmd_jax/md.py:874-916): max(10, min(100, 100000 // N))
    # steps per device call; an overflowing chunk is re-run after the reallocation
    chunk = max(10, min(100, 100000 // N))

    def run_steps(n):
        done, realloc = 0, 0
        while done < n:
            k = min(chunk, n - done)
            _, _, err = sys_.md_run(p, k, record=False)
            if err & 16:
                raise RuntimeError('NaN/Inf during the benchmark')
            if err & 2:
                realloc += 1
                sys_.nlist_allocate(params['rcut'] + 1.0, 1.2 + 0.05 * realloc)
                continue
            done += k
        sys_.synchronize()
        return realloc

    t0 = time.perf_counter()
    realloc = run_steps(args.steps)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # per-kernel HIP-event timing on the library's stream, in a separate pass
    ntimed = max(10, min(50, args.steps))
    sys_.timing(on=True)
    run_steps(ntimed)
    kt = {KERNELS[w]: sys_.timing(which=w) for w in KERNELS}
    sys_.timing(on=False)
    pairs = sys_.pair_count()
    steps_per_s = args.steps / elapsed
    ftot, fk = flops_per_step(N, pairs)
    # dominant kernel by summed event time
    dom = max(fk, key=lambda k: kt[k][0])
    ms_dom, cnt_dom = kt[dom]
    avg_s = ms_dom / 1e3 / max(cnt_dom, 1)
    peak = PEAK_FP32_MFMA if args.precision == 'fp32' else PEAK_BF16_MFMA
    roof = {'bound': 'mfma', 'kernel': dom, 'achieved': fk[dom] / avg_s / 1e12, 'peak': peak / 1e12,
            'unit': 'TFLOP/s', 'frac': fk[dom] / avg_s / peak,
            'traffic': EMBED_BWD_TRAFFIC_FP32 if (dom == 'embed_bwd' and args.precision == 'fp32'
                                                  and args.nmol == 32768) else None,
            'launch_ms': avg_s * 1e3}
    out = {
        'metric': 'MD steps/s, 98k-atom water DP, fp32 vs BF16 at stated force err; p99 latency',
        'value': steps_per_s, 'unit': 'MD steps/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * elapsed / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32' if args.precision == 'fp32' else args.precision,
        'data': 'synthetic (seeded water box + seeded random-init model)',
        'config': {'workload': 'NVE velocity Verlet, %d-atom water (%d H2O, L=%.3f A), se_a 25/50/100 fit 240x3 '
                               'axis 16 rcut 6, dt 0.5 fs, skin 1 A' % (N, args.nmol, box[0]),
                   'atoms': N, 'precision': args.precision},
        'atoms_steps_per_s': steps_per_s * N,
        'ns_per_day': steps_per_s * 0.5 * 86400 * 1e-6,
        # the reference's run-profile unit (deepmd_jax/md_utils.py:172-179): 1e-6 steps natoms / s
        'step_per_us_per_atom': 1e-6 * steps_per_s * N,
        'first_chunk_s': first_chunk_s,
        'model_tflops': ftot * steps_per_s / 1e12,
        'pairs_within_rcut_per_atom': pairs,
        'kernel_ms_per_step': {k: v[0] / ntimed for k, v in kt.items()},
        'nlist_reallocations': realloc,
        'roofline': roof,
    }
    # the full configured run: 2,000 NVE steps from the initial state, rebuilds included
    if args.nve_run_steps and args.nmol == 32768:
        out['nve_run'] = {'steps': args.nve_run_steps}
        for prec in (('fp32', 'bf16') if args.precision == 'fp32' and not args.no_half else (args.precision,)):
            p.precision = _lib.PREC[prec]
            reset()
            sys_.synchronize()
            t0 = time.perf_counter()
            ra = run_steps(args.nve_run_steps)
            el = time.perf_counter() - t0
            out['nve_run'][prec] = {'steps_per_s': args.nve_run_steps / el, 'wall_s': el,
                                    'nlist_rebuilds': sys_.nlist_rebuilds(), 'nlist_reallocations': ra}
            print('# nve run %s: %d steps at %.1f steps/s, %d rebuilds' % (
                prec, args.nve_run_steps, args.nve_run_steps / el, sys_.nlist_rebuilds()), file=sys.stderr, flush=True)
        p.precision = _lib.PREC[args.precision]
    # precision forms: steps/s, same workload
    if args.precision == 'fp32' and not args.no_half:
        for prec in ('bf16', 'fp16'):
            p.precision = _lib.PREC[prec]
            reset()
            run_steps(args.warmup)
            t0 = time.perf_counter()
            run_steps(args.steps)
            out['%s_steps_per_s' % prec] = args.steps / (time.perf_counter() - t0)
            out['%s_speedup_vs_fp32' % prec] = out['%s_steps_per_s' % prec] / steps_per_s
        p.precision = _lib.PREC['fp32']
    if not args.no_sim:
        from deepmd_jax.md import Simulation
        from deepmd_jax.dpmodel import DPModel
        import io
        import contextlib
        ns = max(20, args.steps // 4)
        with contextlib.redirect_stdout(io.StringIO()):
            sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=vel,
                             report_interval=10 ** 9, neighbor_skin=1.0, precision=args.precision,
                             model=(DPModel(params), var, model))
            sim.run(10)
            t0 = time.perf_counter()
            sim.run(ns)
            t1 = time.perf_counter()
        out['simulation_run_steps_per_s'] = ns / (t1 - t0)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(params, var)
    else:
        out['cpu_baseline'] = None
    if not args.no_acc:
        out['precision_errors_vs_fp64_oracle'] = accuracy(params, var, model, args.frames)
    if args.drift_steps:
        out['nve_drift'] = nve_drift(model, params, coord, box, types, vel, args.drift_steps,
                                     equil_steps=args.drift_equil_steps)
    if not args.no_compressed:
        out['compressed'] = compressed(params, var, coord, box, types, vel, args.steps, args.warmup,
                                       acc=not args.no_acc)
    if not args.no_npt:
        out['npt'] = npt(model, params, coord, box, types, vel, args.steps, args.warmup)
    if not args.no_dplr:
        out['dplr'] = dplr(model, params, coord, box, types, vel, args.steps, args.warmup)
    if not args.no_published and args.nmol == 32768:
        out['published_model'] = published_model(coord, box, types, vel, args.steps, args.warmup)
    if not args.no_pimd:
        out['pimd'] = pimd(model, params, var, steps=args.pimd_steps, acc_beads=0 if args.no_acc else args.pimd_acc_beads)
    if not args.no_serving:
        rates = [float(x) for x in args.serving_rates.split(',') if x]
        out['serving'] = serving(model, rates, ('fp32', 'bf16'), args.serving_requests)
    print(json.dumps(out))


if __name__ == '__main__':
    main()

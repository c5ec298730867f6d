"""MD steps/s on the 98,304-atom water box (32,768 H2O, L = 99.34 A) with the seeded
se_a model (embed 25/50/100, fit 240x3, axis 16, rcut 6 A), NVE velocity Verlet at
dt = 0.5 fs, 1 A neighbour skin, one MI355X.

`value` is device-resident throughput over the configured run: from the initial state
(water grid, 300 K velocities) and `--warmup` untimed steps, `--steps` (default 2,000) timed
NVE steps in the reference's chunks, each step the drift test + conditional neighbour
rebuild, the integrator, and one full energy/force/virial evaluation; positions, momenta,
neighbour lists and model stay in HBM.  The BF16 / FP16 forms repeat the same run.  The
Simulation.run rate (host trajectory copies every chunk) is reported beside it as
`simulation_run_steps_per_s`; the precision errors over 16 seeded 98,304-atom frames
against the FP64 C++ restatement and the NVE energy-conservation study are further legs.

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
# same counters read for a 256 MiB elementwise copy (x2.0 and x1.0): 1,073.6 MiB read +
# 447.0 MiB written (profiles/embed_fit_nlist_water98k_trace.md).  Counters cannot be read
# without the profiler, so the figure is carried here and re-measured with the kernel.
EMBED_BWD_TRAFFIC_FP32 = (1073.6 + 447.0) * 2 ** 20


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
    net, hand-written backward, one thread per centre block) driving velocity-Verlet MD steps on
    this host's cores, on bounded samples:
      - 4,096 H2O (12,288 atoms): NVE steps (dt 0.5 fs; neighbour list at rcut + 1 A from the
        KD-tree restatement, rebuilt on the drift test), steps/s extrapolated x8 in N to the
        98,304-atom box (the work is per atom);
      - the 64-H2O box of the CPU-runnable configuration (192 atoms): 1,000 NVE steps.
    The step loop (integrator, drift test, rebuilds) is host NumPy; the forces are the C++ code."""
    sys.path.insert(0, ROOT)
    from oracle import dp_oracle as O, dp_cpu
    from deepmd_jax import synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault('OMP_NUM_THREADS', str(ncores))
    if not os.path.exists(dp_cpu.LIB):
        dp_cpu.build()
    pk = dp_cpu.Packed(params, var)

    def md(nmol, nsteps, budget, lat_skin):
        coord, box, types = syn.water_box(nmol, seed=0)
        m = (np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)[:, None]
        x, pm = coord.copy(), syn.maxwell_boltzmann(types, 300.0, seed=0) * m
        rc_skin = params['rcut'] + lat_skin
        thr = lat_skin / 2

        def build(x):
            nb, _, _ = O.typed_nlist(np.mod(x, box), box, types, rc_skin, buffer_ratio=1.5)
            return dp_cpu.neighbour_rows(nb, types) + (x.copy(),)
        rows, ks, x0 = build(x)
        F = dp_cpu.energy_force(pk, x, box, types, rows, ks, 'f32')[1]
        t0 = time.time()
        done, rebuilds = 0, 0
        while done < nsteps and time.time() - t0 < budget:
            pm += 0.25 * F
            x += 0.5 * pm / m
            d = x - x0
            d -= box * np.round(d / box)
            if np.sqrt((d ** 2).sum(1)).max() > thr:
                rows, ks, x0 = build(x)
                rebuilds += 1
            F = dp_cpu.energy_force(pk, x, box, types, rows, ks, 'f32')[1]
            pm += 0.25 * F
            done += 1
        return done, time.time() - t0, rebuilds, len(types)

    n_big, t_big, rb_big, N_big = md(4096, 1000, seconds_budget, 1.0)
    n_small, t_small, rb_small, N_small = md(64, 1000, seconds_budget, 0.2)
    rate = n_big / t_big / (98304 / N_big)
    return {'value': rate, 'unit': 'MD steps/s', 'cores': dp_cpu.threads(), 'kind': 'port',
            'sample': '%d NVE steps of the 12,288-atom water box (4,096 H2O, seed 0) in %.1f s (%d list rebuilds): '
                      'C++/OpenMP FP32 restatement (oracle/dp_cpu.cpp) forces, host NumPy integrator and drift '
                      'test, KD-tree list rebuilds; steps/s x1/8 for 98,304 atoms; host cores %d'
                      % (n_big, t_big, rb_big, ncores),
            'config0_192_atoms': {'steps': n_small, 'wall_s': t_small, 'steps_per_s': n_small / t_small,
                                  'list_rebuilds': rb_small}}


def accuracy(params, var, model, frames=16, nmol=32768, budget_s=240.0):
    """Energy/atom, force RMSE / max error and virial error of fp32 / bf16 / fp16 against the
    FP64 build of the C++ restatement (oracle/dp_cpu.cpp, hand-written backward, pinned to the
    torch oracle at 1e-15) on `frames` seeded water frames of `nmol` molecules (seeds 0, 1, ...),
    within a wall-clock budget for the CPU side (frames beyond it are reported as not done).
    The CPU evaluation takes the device's neighbour rows at rcut + 0.5 A (bit-exact against the
    oracle's list in tests/test_gpu_nlist.py); it recomputes every distance in FP64.
    Errors: the worst frame's value and the mean over frames."""
    sys.path.insert(0, ROOT)
    from oracle import dp_cpu
    from deepmd_jax import _lib, synthetic as syn
    ncores = len(os.sched_getaffinity(0))
    os.environ.setdefault('OMP_NUM_THREADS', str(ncores))
    if not os.path.exists(dp_cpu.LIB):
        dp_cpu.build()
    pk = dp_cpu.Packed(params, var)
    keys = ('energy_per_atom_abs_err', 'force_rmse_eV_per_A', 'force_max_abs_err', 'virial_max_rel_err')
    per = {prec: {k: [] for k in keys} for prec in ('fp32', 'bf16', 'fp16', 'bf16x3', 'bf16-fit32')}
    frms = []
    s = None
    t0 = time.time()
    done = 0
    for fr in range(frames):
        if time.time() - t0 > budget_s:
            break
        coord, box, types = syn.water_box(nmol, seed=fr)
        N = len(types)
        if s is None:
            s = _lib.DeviceSystem(model, types, 1)
            s.set_box(box)
        s.set_positions(coord)
        s.nlist_allocate(params['rcut'] + 0.5, 1.2)
        dev = {prec: s.compute(prec) for prec in per}
        off = np.concatenate([[0], np.cumsum(np.bincount(types))])
        ex = [s.nlist_export(j)[0] for j in range(2)]
        nb = [[ex[j][off[i]:off[i + 1]] for j in range(2)] for i in range(2)]
        rows, ks = dp_cpu.neighbour_rows(nb, types)
        E, F, W, _ = dp_cpu.energy_force(pk, coord, box, types, rows, ks, 'f64')
        frms.append(float(np.sqrt((F ** 2).mean())))
        for prec, (e, f, v, _) in dev.items():
            vals = (abs(e[0] - E) / N, float(np.sqrt(((f[0] - F) ** 2).mean())),
                    float(np.abs(f[0] - F).max()), float(np.abs(v[0] - W).max() / np.abs(W).max()))
            for k, x in zip(keys, vals):
                per[prec][k].append(x)
        done += 1
        print('# accuracy frame %d done (%.0f s)' % (fr, time.time() - t0), file=sys.stderr, flush=True)
    res = {'frames': '%d of %d requested x %d H2O (%d atoms), seeds 0..%d' % (done, frames, nmol, 3 * nmol,
                                                                            max(done - 1, 0)),
           'frames_done': done, 'reference': 'FP64 C++ restatement (oracle/dp_cpu.cpp), %d host threads'
                                              % dp_cpu.threads(),
           'cpu_wall_s': time.time() - t0}
    for prec in per:
        res[prec] = {k: float(np.max(v)) for k, v in per[prec].items()}
        res[prec]['mean'] = {k: float(np.mean(v)) for k, v in per[prec].items()}
    res['force_rms_eV_per_A'] = float(np.mean(frms))
    return res


def nve_conservation(model, params, var, nmol=32768, seeds=(0, 1, 2), clean_steps=240, total_steps=600,
                     sample=20, small_nmol=216, small_steps=400):
    """NVE energy conservation of the FP32 and BF16 forms (velocity Verlet, dt 0.5 fs, skin 1 A).

    The seeded random-init potential has no repulsive core: once the box has warmed (after
    ~250 fs from the water grid) pairs fall into steep wells where forces spike to 50-200 eV/A,
    and velocity Verlet at 0.5 fs then gains energy in jumps -- the FP64 oracle does the same on
    the same trajectory.  So two windows are reported:
      clean: steps [0, clean_steps] from the grid, before the first close encounter;
      full:  steps [0, total_steps], encounters included.
    (1) `nmol`-molecule box, for each seed s (positions seed s, 300 K velocities seed s + 1):
        H = KE + PE every `sample` steps; slope (eV/atom/ps) of a least-squares line per window,
        mean and standard deviation over seeds.  For the BF16 trajectory H is reported with its own
        (BF16) energy and with the FP32 energy of the same positions (the Hamiltonian the
        dynamics approximates; the BF16 energy carries a ~2e-4 eV/atom evaluation error that
        wanders with the configuration).
    (2) FP64 reference: on a `small_nmol`-molecule box the FP64 C++ restatement
        (oracle/dp_cpu.cpp) drives the same velocity Verlet from the same state for `small_steps`
        steps; the device FP32 and BF16 runs are compared with its H(t)."""
    sys.path.insert(0, ROOT)
    from oracle import dp_oracle as O, dp_cpu
    from deepmd_jax import _lib, synthetic as syn
    from deepmd_jax.constants import MASS_UNIT_CONVERSION
    kB = 8.617333262e-5

    def run_device(coord, box, types, vel, prec, nsteps):
        N = len(types)
        m = np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION
        s = _lib.DeviceSystem(model, types, 1)
        s.set_box(box)
        s.set_masses(m)
        s.set_positions(coord)
        s.set_velocities(vel)
        s.nlist_allocate(params['rcut'] + 1.0, 1.3)
        p = _lib.MDParams()
        p.routine, p.dt, p.precision = 0, 0.5, _lib.PREC[prec]
        k, H, H32, Fm = 0, [], [], []
        while True:
            e, f, _, _ = s.compute(prec)
            ke = float(s.kinetic_energy()[0])
            H.append((float(e[0]) + ke) / N)
            H32.append((float(s.compute('fp32')[0][0]) + ke) / N if prec != 'fp32' else H[-1])
            Fm.append(float(np.abs(f).max()))
            if k >= nsteps:
                break
            _, _, err = s.md_run(p, sample, record=False)
            if err & 2:
                s.nlist_allocate(params['rcut'] + 1.0, 1.4)
                continue
            if err & 16:
                break
            k += sample
        T = 2 * ke / (3 * N) / kB
        return np.array(H), np.array(H32), np.array(Fm), T, k

    def slope(H, n):
        t = np.arange(len(H)) * sample * 0.5e-3
        kk = t <= n * 0.5e-3 + 1e-12
        return float(np.polyfit(t[kk], H[kk], 1)[0])

    out = {'dt_fs': 0.5, 'sample_every': sample, 'clean_window_steps': clean_steps,
           'full_window_steps': total_steps, 'atoms': 3 * nmol, 'seeds': list(seeds)}
    forms = ('fp32', 'bf16', 'bf16x3', 'bf16-fit32')
    per = {f: {'clean': [], 'full': [], 'range_clean': [], 'maxF_clean': [], 'maxF_full': [], 'T_end': []}
           for f in forms}
    for f in forms[1:]:
        per[f].update({'clean_fp32_energy': [], 'full_fp32_energy': []})
    for sd in seeds:
        coord, box, types = syn.water_box(nmol, seed=sd)
        vel = syn.maxwell_boltzmann(types, 300.0, seed=sd + 1)
        for prec in forms:
            H, H32, Fm, T, k = run_device(coord, box, types, vel, prec, total_steps)
            nc = clean_steps // sample + 1
            d = per[prec]
            d['clean'].append(slope(H, clean_steps))
            d['full'].append(slope(H, total_steps))
            d['range_clean'].append(float(H[:nc].max() - H[:nc].min()))
            d['maxF_clean'].append(float(Fm[:nc].max()))
            d['maxF_full'].append(float(Fm.max()))
            d['T_end'].append(T)
            if prec != 'fp32':
                d['clean_fp32_energy'].append(slope(H32, clean_steps))
                d['full_fp32_energy'].append(slope(H32, total_steps))
            print('# conservation seed %d %s: clean %.3e full %.3e eV/atom/ps, max|F| %.1f / %.1f' % (
                sd, prec, d['clean'][-1], d['full'][-1], d['maxF_clean'][-1], d['maxF_full'][-1]),
                file=sys.stderr, flush=True)
    for prec, d in per.items():
        out[prec] = {}
        for k_, v in d.items():
            out[prec][k_] = {'per_seed': v, 'mean': float(np.mean(v)), 'std': float(np.std(v))}
    # "no worse drift" per 16-bit form: |mean drift of its trajectories (FP32 energy)| <= |FP32's| +
    # 2 sigma of the mean difference over the seeds
    f32c = per['fp32']['clean']
    out['no_worse_drift_by_form'] = {}
    parts = []
    for f in forms[1:]:
        c = per[f]['clean_fp32_energy']
        tol = 2 * np.sqrt((np.var(f32c) + np.var(c)) / len(seeds))
        holds = bool(abs(np.mean(c)) <= abs(np.mean(f32c)) + tol)
        out['no_worse_drift_by_form'][f] = holds
        parts.append('%s %.2e +- %.1e (%s)' % (f, np.mean(c), np.std(c), 'no worse' if holds else 'worse'))
    out['no_worse_drift'] = out['no_worse_drift_by_form']['bf16']
    out['statement'] = (
        'clean window (%d steps, %.2f ps), drift of H = KE + PE with the FP32 energy of each trajectory: FP32 %.2e +- '
        '%.1e eV/atom/ps; %s (no worse: |form| <= |FP32| + 2 sigma of the mean difference, %d seeds).  BF16 with '
        'its own BF16 energy %.2e +- %.1e.  Full window (%d steps, close encounters included): FP32 %.2e +- %.1e, '
        'BF16 %.2e +- %.1e, bf16x3 %.2e +- %.1e.' % (
            clean_steps, clean_steps * 0.5e-3, np.mean(f32c), np.std(f32c), '; '.join(parts), len(seeds),
            np.mean(per['bf16']['clean']), np.std(per['bf16']['clean']), total_steps,
            np.mean(per['fp32']['full']), np.std(per['fp32']['full']),
            np.mean(per['bf16']['full_fp32_energy']), np.std(per['bf16']['full_fp32_energy']),
            np.mean(per['bf16x3']['full_fp32_energy']), np.std(per['bf16x3']['full_fp32_energy'])))
    # (2) FP64 reference trajectory on the small box
    coord, box, types = syn.water_box(small_nmol, seed=0)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=1)
    N = len(types)
    m = (np.array(syn.MASSES)[types] * MASS_UNIT_CONVERSION)[:, None]
    if not os.path.exists(dp_cpu.LIB):
        dp_cpu.build()
    pk = dp_cpu.Packed(params, var)

    def rows_for(x):
        nb, _, _ = O.typed_nlist(np.mod(x, box), box, types, params['rcut'] + 1.0, buffer_ratio=1.5)
        return dp_cpu.neighbour_rows(nb, types) + (x.copy(),)
    t0 = time.time()
    x, pm = coord.copy(), vel * m
    rows, ks, x0 = rows_for(x)
    E, F, _, _ = dp_cpu.energy_force(pk, x, box, types, rows, ks, 'f64')
    Href = [(E + 0.5 * (pm ** 2 / m).sum()) / N]
    for k in range(small_steps):
        pm += 0.25 * F
        x += 0.5 * pm / m
        d = x - x0
        d -= box * np.round(d / box)
        if np.sqrt((d ** 2).sum(1)).max() > 0.5:
            rows, ks, x0 = rows_for(x)
        E, F, _, _ = dp_cpu.energy_force(pk, x, box, types, rows, ks, 'f64')
        pm += 0.25 * F
        if (k + 1) % sample == 0:
            Href.append((E + 0.5 * (pm ** 2 / m).sum()) / N)
    Href = np.array(Href)
    ref = {'atoms': N, 'steps': small_steps, 'cpu_wall_s': time.time() - t0,
           'fp64_drift_eV_per_atom_per_ps': slope(Href, small_steps),
           'fp64_H_range_eV_per_atom': float(Href.max() - Href.min())}
    for prec in ('fp32', 'bf16', 'bf16x3'):
        H, H32, Fm, T, k = run_device(coord, box, types, vel, prec, small_steps)
        n = min(len(H), len(Href))
        ref[prec] = {'drift_eV_per_atom_per_ps': slope(H32, small_steps),
                     'max_abs_H_minus_fp64_eV_per_atom': float(np.abs(H32[:n] - Href[:n]).max()),
                     'H_range_eV_per_atom': float(H32.max() - H32.min())}
    out['fp64_reference_small_box'] = ref
    print('# conservation: %s' % out['statement'], file=sys.stderr, flush=True)
    return out


def config0(model, params, var, steps=1000):
    """The CPU-runnable configuration on the device: 64 H2O (192 atoms, L = 12.42 A, seed 0),
    one E/F/virial evaluation against the FP64 C++ restatement, then 1,000 velocity-Verlet
    steps at dt 0.5 fs from 300 K velocities through Simulation (the box is below
    2 (rcut + skin), so the skin shrinks to fit the neighbour-list regime, as
    deepmd_jax/md.py:360-370 decides between its list and lattice paths)."""
    import contextlib
    import io
    sys.path.insert(0, ROOT)
    from oracle import dp_oracle as O, dp_cpu
    from deepmd_jax import synthetic as syn
    from deepmd_jax.dpmodel import DPModel
    from deepmd_jax.md import Simulation
    coord, box, types = syn.water_box(64, seed=0)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=0)
    with contextlib.redirect_stdout(io.StringIO()):
        sim = Simulation(None, box, types, syn.MASSES, 'NVE', 0.5, coord, initial_velocity=vel,
                         report_interval=10 ** 9, model=(DPModel(params), var, model))
        E, F = sim.getEnergy(), sim.getForce()
        t0 = time.perf_counter()
        traj = sim.run(steps)
        el = time.perf_counter() - t0
    if not os.path.exists(dp_cpu.LIB):
        dp_cpu.build()
    nb, _, _ = O.typed_nlist(coord, box, types, params['rcut'] + 0.2)
    rows, ks = dp_cpu.neighbour_rows(nb, types)
    Er, Fr, _, _ = dp_cpu.energy_force(dp_cpu.Packed(params, var), coord, box, types, rows, ks, 'f64')
    N = len(types)
    return {'workload': '64 H2O (192 atoms, L=%.2f A), Simulation NVE dt 0.5 fs, %d steps' % (box[0], steps),
            'steps_per_s': steps / el, 'wall_s': el, 'finite': bool(np.isfinite(traj['position']).all()),
            'energy_per_atom_abs_err_vs_fp64': abs(E - Er) / N,
            'force_max_abs_err_vs_fp64': float(np.abs(F - Fr).max())}


def pimd(model, params, var, steps=1000, warmup=10, precisions=('fp32', 'bf16'), acc_beads=4, npt_steps=200):
    """Ring-polymer NVT (PILE-L, 300 K, dt 0.25 fs): 32 beads x 4,096 H2O (12,288 atoms per
    bead, L = 49.7 A), per-bead +-0.05 A uniform jitter from seed 3, all beads in one
    system (one launch per kernel for all 393,216 atoms), the configured 1,000 steps per
    precision.  Device-resident bead*steps/s per precision, the path-integral NPT rate on
    `npt_steps`, and the centroid-force error against the FP64 C++ restatement over the first
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
    run(npt_steps)
    s.synchronize()
    el = time.perf_counter() - t0
    out['npt_fp32'] = {'steps_per_s': npt_steps / el, 'bead_steps_per_s': P * npt_steps / el,
                       'ms_per_step': 1e3 * el / npt_steps, 'steps': npt_steps,
                       'box_rel_change': float(s.get_boxes()[0][0] / box[0] - 1.0)}
    p.routine, p.gamma = _lib.ROUTINE['PIMD'], 0.01
    if acc_beads:
        sys.path.insert(0, ROOT)
        from oracle import dp_cpu
        if not os.path.exists(dp_cpu.LIB):
            dp_cpu.build()
        pk = dp_cpu.Packed(params, var)
        reset()
        # the device's rows (bit-exact against the oracle's list) at rcut + 1 A; FP64 distances
        off = np.concatenate([[0], np.cumsum(np.bincount(types))])
        ex = [s.nlist_export(j) for j in range(2)]
        Fo = []
        for b in range(acc_beads):
            nb = [[ex[j][b][off[i]:off[i + 1]] for j in range(2)] for i in range(2)]
            rows, ks = dp_cpu.neighbour_rows(nb, types)
            Fo.append(dp_cpu.energy_force(pk, x[b], box, types, rows, ks, 'f64')[1])
        Fc = np.mean(Fo, 0)
        out['centroid_force'] = {'beads': acc_beads, 'oracle': 'FP64 C++ restatement, per bead',
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


def probe_issue(blocks=512, iters=20000, reps=5):
    """Issue-overlap probe: does matrix-core work overlap FP32 VALU work on this GPU?

    Each case runs 2 blocks per CU; every wave issues the same chains in every case (4
    independent MFMA accumulators, 8 independent FMA chains x 4 per iteration).  If MFMA and
    VALU work overlap, "separate waves" takes about the longer of the two alone; if they share
    the issue path it takes about their sum."""
    import ctypes
    path = os.path.join(ROOT, 'deepmd-jax_amd', 'deepmd_jax', 'libdpmd_probe.so')
    if not os.path.exists(path):
        raise SystemExit('libdpmd_probe.so is not built (python deepmd-jax_amd/build.py)')
    lib = ctypes.CDLL(path)
    lib.dpmd_probe_issue.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]
    names = ['f32_mfma', 'valu', 'f32_mfma+valu_same_wave', 'f32_mfma|valu_separate_waves_halves',
             'bf16_mfma', 'bf16_mfma+valu_same_wave', 'bf16_mfma|valu_separate_waves_halves',
             'f32_mfma|valu_separate_waves_parity', 'bf16_mfma|valu_separate_waves_parity',
             # the same with the MFMA accumulators in AccVGPRs (asm constraint) instead of ArchVGPRs
             'f32agpr_mfma', 'f32agpr_mfma+valu_same_wave', 'f32agpr_mfma|valu_separate_waves_parity',
             'bf16agpr_mfma', 'bf16agpr_mfma+valu_same_wave', 'bf16agpr_mfma|valu_separate_waves_parity',
             'f32agpr_mfma|valu_separate_waves_halves', 'bf16agpr_mfma|valu_separate_waves_halves']
    out = {}
    for which, name in enumerate(names):
        ms = ctypes.c_float()
        if lib.dpmd_probe_issue(which, blocks, iters, reps, ctypes.byref(ms)) != 0:
            raise SystemExit('probe case %s failed' % name)
        out[name] = round(ms.value, 4)
        print('probe %s: %.4f ms' % (name, ms.value), file=sys.stderr, flush=True)
    m, v = out['f32_mfma'], out['valu']
    mb = out['bf16_mfma']
    for kind, mk in (('f32', m), ('bf16', mb), ('f32agpr', out['f32agpr_mfma']), ('bf16agpr', out['bf16agpr_mfma'])):
        same = out['%s_mfma+valu_same_wave' % kind]
        sep = min(out['%s_mfma|valu_separate_waves_halves' % kind], out['%s_mfma|valu_separate_waves_parity' % kind])
        out['%s_same_wave_over_sum' % kind] = round(same / (mk + v), 3)
        out['%s_separate_over_sum' % kind] = round(sep / (mk + v), 3)
        out['%s_separate_over_max' % kind] = round(sep / max(mk, v), 3)
    out['config'] = {'blocks': blocks, 'iters': iters, 'launches': reps, 'mfma_per_wave_iter': 4,
                     'valu_fma_per_lane_iter': 32}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2000)
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
    ap.add_argument('--frames', type=int, default=16, help='seeded 98,304-atom frames for the precision errors')
    ap.add_argument('--acc-budget', type=float, default=240.0, help='CPU seconds for the FP64 reference frames')
    ap.add_argument('--no-conservation', action='store_true', help='skip the NVE energy-conservation leg')
    ap.add_argument('--no-config0', action='store_true', help='skip the 192-atom configuration leg')
    ap.add_argument('--deterministic', action='store_true',
                    help='the timed system gathers pair gradients in a fixed order instead of force atomics')
    ap.add_argument('--pimd-steps', type=int, default=1000)
    ap.add_argument('--pimd-acc-beads', type=int, default=2)
    ap.add_argument('--serving-rates', default='20,200,1000',
                    help='comma-separated Poisson arrival rates (requests/s)')
    ap.add_argument('--serving-requests', type=int, default=256)
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='run a plain 256 MiB elementwise copy (read 256 MiB, write 256 MiB) five times, for '
                         'scaling FETCH_SIZE / WRITE_SIZE counters')
    ap.add_argument('--probe-issue', action='store_true',
                    help='time FP32 / BF16 MFMA chains and FP32 VALU chains alone, in one wave and in separate '
                         'waves of one block (libdpmd_probe.so), and exit')
    args = ap.parse_args()
    if args.probe_issue:
        probe_issue()
        return
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
        print('bench.py runs on one GPU; --gpus %d is not supported' % args.gpus, file=sys.stderr)
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
    sys_.set_deterministic(args.deterministic)
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
    def warm():
        """The initial state and `--warmup` untimed steps (includes the first force evaluation)."""
        reset()
        done = 0
        while done < args.warmup:
            k = min(50, args.warmup - done)
            _, _, err = sys_.md_run(p, k, record=False)
            if err & 2:
                sys_.nlist_allocate(params['rcut'] + 1.0, 1.25)
                continue
            done += k
        sys_.synchronize()

    warm()

    # chunks of the reference's run protocol (deepmd_jax/md.py:874-916): max(10, min(100, 100000 // N))
    # steps per device call; an overflowing chunk is re-run after the reallocation
    chunk = max(10, min(100, 100000 // N))

    def run_steps(n, pair_samples=None):
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
            if pair_samples is not None:
                pair_samples.append(sys_.pair_count())
        sys_.synchronize()
        return realloc

    t0 = time.perf_counter()
    realloc = run_steps(args.steps)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # per-kernel HIP-event timing on the library's stream over a second pass of the same
    # configured run (same initial state, warmup and steps; the value pass above carries no
    # events), so the averages cover the whole timed workload, cold lattice to heated box; the
    # counted FLOPs use the pair count within rcut averaged over the pass (sampled every chunk)
    warm()
    ntimed = args.steps
    sys_.timing(on=True)
    samples = []
    run_steps(ntimed, pair_samples=samples)
    kt = {KERNELS[w]: sys_.timing(which=w) for w in KERNELS}
    sys_.timing(on=False)
    pairs = float(np.mean(samples)) if samples else sys_.pair_count()
    pairs_end = sys_.pair_count()
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
                   'atoms': N, 'precision': args.precision, 'deterministic_forces': bool(args.deterministic)},
        'atoms_steps_per_s': steps_per_s * N,
        'ns_per_day': steps_per_s * 0.5 * 86400 * 1e-6,
        # the reference's run-profile unit (deepmd_jax/md_utils.py:172-179): 1e-6 steps natoms / s
        'step_per_us_per_atom': 1e-6 * steps_per_s * N,
        'first_chunk_s': first_chunk_s,
        'model_tflops': ftot * steps_per_s / 1e12,
        'pairs_within_rcut_per_atom': pairs,
        'pairs_within_rcut_per_atom_end': pairs_end,
        'kernel_ms_per_step': {k: v[0] / ntimed for k, v in kt.items()},
        'nlist_reallocations': realloc,
        'roofline': roof,
    }
    # precision forms: the same configured run in BF16; FP16 (5 exponent bits) on the shorter
    # window, since its operands overflow once the seeded potential's close encounters drive
    # descriptor entries past 65,504 (the run then ends with error bit 16, reported as such)
    if args.precision == 'fp32' and not args.no_half:
        for prec, nst in (('bf16', args.steps), ('bf16x3', args.steps), ('fp16', min(args.steps, 200))):
            p.precision = _lib.PREC[prec]
            reset()
            try:
                run_steps(args.warmup)
                t0 = time.perf_counter()
                run_steps(nst)
                out['%s_steps_per_s' % prec] = nst / (time.perf_counter() - t0)
                out['%s_speedup_vs_fp32' % prec] = out['%s_steps_per_s' % prec] / steps_per_s
                out['%s_timed_steps' % prec] = nst
            except RuntimeError as e:
                out['%s_steps_per_s' % prec] = None
                out['%s_error' % prec] = '%s (within %d steps from the initial state)' % (e, nst + args.warmup)
            print('# %s: %s steps/s' % (prec, out['%s_steps_per_s' % prec]), file=sys.stderr, flush=True)
        p.precision = _lib.PREC['fp32']
    # the secondary legs (compressed model, NPT, DPLR, the published-figure model) time their own
    # shorter windows so that the whole default run stays within a few minutes
    leg_steps = min(args.steps, 200)
    if not args.no_sim:
        from deepmd_jax.md import Simulation
        from deepmd_jax.dpmodel import DPModel
        import io
        import contextlib
        ns = max(20, leg_steps)
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
        out['precision_errors_vs_fp64_oracle'] = accuracy(params, var, model, args.frames, nmol=args.nmol,
                                                          budget_s=args.acc_budget)
    if not args.no_conservation:
        out['nve_energy_conservation'] = nve_conservation(model, params, var, nmol=args.nmol)
    if not args.no_compressed:
        out['compressed'] = compressed(params, var, coord, box, types, vel, leg_steps, args.warmup,
                                       acc=not args.no_acc)
    if not args.no_npt:
        out['npt'] = npt(model, params, coord, box, types, vel, leg_steps, args.warmup)
    if not args.no_dplr:
        out['dplr'] = dplr(model, params, coord, box, types, vel, leg_steps, args.warmup)
    if not args.no_published and args.nmol == 32768:
        out['published_model'] = published_model(coord, box, types, vel, leg_steps, args.warmup)
    if not args.no_config0:
        out['config0'] = config0(model, params, var)
    if not args.no_pimd:
        out['pimd'] = pimd(model, params, var, steps=args.pimd_steps, acc_beads=0 if args.no_acc else args.pimd_acc_beads)
    if not args.no_serving:
        rates = [float(x) for x in args.serving_rates.split(',') if x]
        out['serving'] = serving(model, rates, ('fp32', 'bf16'), args.serving_requests)
    # the precision target per 16-bit form: >= 1.5x the FP32 form's steps/s over the configured
    # run, force RMSE <= 1e-3 eV/A against the FP64 restatement (worst frame) and NVE drift no
    # worse than FP32's (clean window, FP32 energy of each trajectory)
    if args.precision == 'fp32' and not args.no_half:
        acc = out.get('precision_errors_vs_fp64_oracle') or {}
        cons = (out.get('nve_energy_conservation') or {}).get('no_worse_drift_by_form', {})
        tgt = {}
        for f in ('bf16', 'bf16x3', 'fp16'):
            sp = out.get('%s_speedup_vs_fp32' % f)
            rm = (acc.get(f) or {}).get('force_rmse_eV_per_A')
            nwd = cons.get(f)
            tgt[f] = {'speedup_vs_fp32': sp, 'force_rmse_eV_per_A': rm, 'no_worse_drift': nwd,
                      'meets': bool(sp is not None and sp >= 1.5 and rm is not None and rm <= 1e-3 and nwd)}
        out['precision_target'] = tgt
    print(json.dumps(out))


if __name__ == '__main__':
    main()

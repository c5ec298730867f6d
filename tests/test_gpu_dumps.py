"""Simulation.run(dump_prefix=...) on the device: the XYZ frames equal the in-memory trajectory
of an identical run rounded to float32, within 1e-6 of the largest value (the force sums use
atomics, so two runs agree to rounding rather than bit for bit: 2e-8 A/fs apart after 60 steps).  The frame text rules are tested in
tests/test_dumps.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _read_xyz(path):
    lines = open(path).read().splitlines()
    frames, steps, i = [], [], 0
    while i < len(lines):
        n = int(lines[i])
        steps.append(int(lines[i + 1].split('step=')[1].split()[0]))
        frames.append(np.array([[float(v) for v in ln.split()[1:]] for ln in lines[i + 2:i + 2 + n]]))
        i += 2 + n
    return steps, np.array(frames)


def test_nve_dumps_match_trajectory(model_pkl, tmp_path):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case dumps nve', flush=True)
    coord, box, types = syn.water_box(64, seed=8)
    vel = syn.maxwell_boltzmann(types, 300.0, seed=8)
    kw = dict(initial_velocity=vel, report_interval=25, type_symbols=['O', 'H'])
    ref = Simulation(model_pkl, box, types, syn.MASSES, 'NVE', 0.5, coord, **kw).run(60)
    sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVE', 0.5, coord, **kw)
    prefix = str(tmp_path / 'w')
    assert sim.run(60, dump_interval=20, dump_prefix=prefix, dump_content='all') is None
    for kind in ('position', 'velocity'):
        steps, fr = _read_xyz('%s_%s.xyz' % (prefix, kind))
        assert steps == [0, 20, 40, 60]
        want = ref[kind][steps].astype(np.float32).astype(np.float64)
        assert np.abs(fr - want).max() <= 5e-9 + 1e-6 * np.abs(want).max()
    head = open(prefix + '_position.xyz').read().splitlines()[:3]
    assert head[0] == '192' and head[2].startswith('O  ')
    assert 'Lattice="%.8f 0.00000000' % np.float32(box[0]) in head[1]
    # a second run appends nothing unless asked: overwrite mode restarts the file
    sim.run(20, dump_prefix=prefix)
    steps, _ = _read_xyz(prefix + '_position.xyz')
    assert steps == [75]


def test_pimd_dumps(model_pkl, tmp_path):
    from deepmd_jax import synthetic as syn
    from deepmd_jax.md import Simulation
    print('case dumps pimd', flush=True)
    coord, box, types = syn.water_box(64, seed=9)
    sim = Simulation(model_pkl, box, types, syn.MASSES, 'NVT', 0.5, coord, temperature=300.0, n_bead=4,
                     report_interval=10, type_symbols=['O', 'H'], seed=1)
    prefix = str(tmp_path / 'pi')
    sim.run(20, dump_prefix=prefix)
    steps, cen = _read_xyz(prefix + '_centroid.xyz')
    assert steps == [0, 10, 20]
    beads = [_read_xyz('%s_position_bead%d.xyz' % (prefix, b))[1] for b in range(4)]
    np.testing.assert_allclose(beads[0][-1], sim.getPosition()[0], atol=1e-5)
    np.testing.assert_allclose(cen[-1], sim.getCentroidPosition(), atol=1e-5)

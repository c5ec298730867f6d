"""XYZ dump rules and frame text (deepmd_jax/md_utils.py:20-256), on the CPU.

The expected frame text below is written out by hand from the reference's format:
`%-2s %.8f %.8f %.8f` per atom for float32 frames, the comment line
`Lattice="<cell row by row>" Properties=species:S:1:pos:R:3 step=<n> pbc="T T T"`
(deepmd_jax/md_utils.py:115-136; the orthorhombic cell of `set_cell([L, 90, 90, 90])`).
"""
import os

import numpy as np
import pytest

from deepmd_jax import dumps as D

FRAME = ('2\n'
         'Lattice="10.00000000 0.00000000 0.00000000 0.00000000 11.00000000 0.00000000 0.00000000 0.00000000 '
         '12.00000000" Properties=species:S:1:pos:R:3 step=5 pbc="T T T"\n'
         'O  1.50000000 2.25000000 3.00000000\n'
         'H  0.10000000 -0.20000000 10.00000000\n')


def test_frame_text():
    x = np.array([[1.5, 2.25, 3.0], [0.1, -0.2, 10.0]], np.float32)
    assert D.format_frame(['O', 'H'], x, np.array([10, 11, 12], np.float32), 5) == FRAME
    # float64 frames carry 15 decimals; no step -> no step field
    t = D.format_frame(['O', 'H'], x.astype(np.float64), np.array([10., 11., 12.]))
    assert 'step=' not in t
    assert t.splitlines()[2] == 'O  1.500000000000000 2.250000000000000 3.000000000000000'
    assert t.splitlines()[3].startswith('H  0.100000001490116 -0.200000002980232')
    # a general (3,3) cell is written row by row (deepmd_jax/md_utils.py:236-239)
    H = np.array([[10, 0, 0], [2.5, 11, 0], [-1, 3, 12]], np.float32)
    t = D.format_frame(['O', 'H'], x, H, 5)
    assert t.splitlines()[1].startswith('Lattice="10.00000000 0.00000000 0.00000000 2.50000000 11.00000000 '
                                        '0.00000000 -1.00000000 3.00000000 12.00000000" ')


def test_content_rules():
    assert D.normalize_dump_content('positions', 1) == ('position',)
    assert D.normalize_dump_content('positions', 4) == ('position', 'centroid')
    assert D.normalize_dump_content('all', 1) == ('position', 'velocity')
    assert D.normalize_dump_content('all', 2) == ('position', 'velocity', 'centroid')
    assert D.normalize_dump_content(['velocity', 'position'], 1) == ('position', 'velocity')
    for bad, nb in ((['centroid'], 1), (['pos'], 1), ([], 1), ([3], 1), (5, 1)):
        with pytest.raises(ValueError):
            D.normalize_dump_content(bad, nb)
    with pytest.raises(ValueError):
        D.normalize_dump_mode('replace')
    with pytest.raises(ValueError):
        D.requested_paths(None, 'velocity', 1)
    with pytest.raises(ValueError):
        D.requested_paths('', 'positions', 1)
    assert D.requested_paths(None, 'all', 1) == {}
    assert D.requested_paths('out/t', 'all', 1) == {'position': 'out/t_position.xyz', 'velocity': 'out/t_velocity.xyz'}
    assert D.bead_path('out/t_position.xyz', 3) == 'out/t_position_bead3.xyz'


def test_prepare_and_write(tmp_path):
    prefix = str(tmp_path / 'traj')
    types = np.array([0, 1, 1])
    with pytest.raises(ValueError):
        D.prepare_dumps(types, None, 1, 10, dump_prefix=prefix)
    with pytest.raises(ValueError):
        D.prepare_dumps(types, ['O', 'H'], 1, 10, dump_interval=0, dump_prefix=prefix)
    # overwrite truncates, append keeps
    open(prefix + '_position.xyz', 'w').write('old\n')
    dumps, sym, interval = D.prepare_dumps(types, ['O', 'H'], 1, 10, dump_prefix=prefix, dump_content='all')
    assert interval == 10 and sym == ['O', 'H', 'H']
    assert not os.path.exists(prefix + '_position.xyz')
    w = D.DumpWriter(dumps, sym)
    x = np.arange(9, dtype=np.float32).reshape(3, 3)
    for step in (0, 10):
        w.put(x + step, -x, None, np.array([5, 5, 5], np.float32), step)
    w.close()
    lines = open(prefix + '_position.xyz').read().splitlines()
    assert len(lines) == 10 and 'step=0' in lines[1] and 'step=10' in lines[6]
    assert lines[7] == 'O  10.00000000 11.00000000 12.00000000'
    assert open(prefix + '_velocity.xyz').read().splitlines()[2] == 'O  -0.00000000 -1.00000000 -2.00000000'
    dumps, sym, _ = D.prepare_dumps(types, ['O', 'H'], 1, 10, dump_prefix=prefix, dump_mode='append')
    w = D.DumpWriter(dumps, sym)
    w.put(x, None, None, np.array([5, 5, 5], np.float32), 20)
    w.close()
    assert len(open(prefix + '_position.xyz').read().splitlines()) == 15
    # a missing source is reported by close()
    w = D.DumpWriter(dumps, sym)
    w.put(None, None, None, np.array([5, 5, 5], np.float32), 30)
    with pytest.raises(ValueError):
        w.close()


def test_pimd_bead_files(tmp_path):
    prefix = str(tmp_path / 'pi')
    for b in range(5):
        open('%s_position_bead%d.xyz' % (prefix, b), 'w').write('stale\n')
    dumps, sym, _ = D.prepare_dumps(np.array([0, 1]), ['O', 'H'], 3, 7, dump_prefix=prefix)
    assert [d[1] for d in dumps] == ['%s_position_bead%d.xyz' % (prefix, b) for b in range(3)] + [prefix + '_centroid.xyz']
    assert not any(os.path.exists('%s_position_bead%d.xyz' % (prefix, b)) for b in range(5))
    w = D.DumpWriter(dumps, sym)
    pos = np.arange(18, dtype=np.float32).reshape(3, 2, 3)
    w.put(pos, None, pos.mean(0), np.array([9, 9, 9], np.float32), 7)
    w.close()
    assert open(prefix + '_position_bead2.xyz').read().splitlines()[2] == 'O  12.00000000 13.00000000 14.00000000'
    assert open(prefix + '_centroid.xyz').read().splitlines()[3] == 'H  9.00000000 10.00000000 11.00000000'

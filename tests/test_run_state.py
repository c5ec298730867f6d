"""The run loop's error-code state machine (deepmd_jax/md.py:616-676, 874-964) on a stub
device that returns scripted codes: bit 2 reallocates and sets bit 4 for the rest of the
run, so a later, non-consecutive overflow grows the buffer ratio; bit 1 (NPT box shrink)
widens the skin by 0.4 A and reallocates; bit 16 stops the run after recording the chunk;
an overflowing chunk is run again."""
import numpy as np

from deepmd_jax.md import Simulation


class _StubDev:
    def __init__(self, box):
        self.box = box

    def box_trajectory(self, n):
        return np.broadcast_to(self.box, (n, 1, 3))


class _StubSys:
    def __init__(self, codes, natoms, box):
        self.codes = list(codes)
        self.calls = []
        self.natoms = natoms
        self.dev = _StubDev(box)

    def md_run(self, params, nsteps, record=True, out=None):
        err = self.codes.pop(0) if self.codes else 0
        self.calls.append((nsteps, err))
        return out[0], out[1], err


def _sim(codes, routine='NVE', box=30.0):
    s = Simulation.__new__(Simulation)
    s._sys = _StubSys(codes, 3, np.full(3, box))
    s._routine, s._is_pimd, s._n_bead, s._natoms = routine, False, 1, 3
    s._params, s._dt = None, 0.5
    s.report_interval, s._step_chunk_size = 10, 5
    s.step, s._is_initial_state = 0, False
    s._box = np.full(3, box)
    s._rcut, s._neighbor_skin, s._neighbor_buffer_ratio = 6.0, 0.5, 1.2
    s.allocs = []
    s._allocate_nlist = lambda knbr=None: s.allocs.append((s._neighbor_skin, s._neighbor_buffer_ratio))
    s._print_report = lambda: None
    s.getBox = lambda: s._box.copy()
    s.getPosition = s.getVelocity = lambda: np.zeros((3, 3))
    return s


def test_overflow_bit4_persists_within_a_run():
    s = _sim([0, 2, 0, 0, 2, 0])
    s._run(20, None, None)
    # chunks: ok, overflow (re-run), ok, ok, overflow (re-run), ok
    assert [c[1] for c in s._sys.calls] == [0, 2, 0, 0, 2, 0]
    assert s.allocs == [(0.5, 1.2), (0.5, 1.25)]    # the second overflow found bit 4 still set
    assert s.step == 20 and s._error_code == 4


def test_bit4_resets_at_the_next_run():
    s = _sim([2, 0, 0, 0, 0, 2, 0])
    s._run(20, None, None)
    s._run(15, None, None)
    assert s.allocs == [(0.5, 1.2), (0.5, 1.2)]


def test_npt_box_shrink_widens_the_skin():
    s = _sim([0, 1, 0, 0], routine='NPT')
    s._run(15, None, None)
    assert s.allocs == [(0.9, 1.2)]
    assert s._error_code == 0 and s.step == 15


def test_npt_box_shrink_below_the_list_regime_raises():
    import pytest
    s = _sim([1], routine='NPT', box=13.5)
    with pytest.raises(NotImplementedError):
        s._run(5, None, None)


def test_nan_stops_after_recording_the_chunk():
    s = _sim([0, 16, 0])
    s._run(20, None, None)
    assert len(s._sys.calls) == 2 and s.step == 10 and s._error_code & 16

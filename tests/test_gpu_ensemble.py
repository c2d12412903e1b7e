"""Ensemble runner (cryo_ensemble_run_steps): the tanks of an ensemble give the
bitwise same particles and diagnostics as the same tanks stepped one by one
with cryo_solver_run_steps, including tanks with phase change (where the
particle count changes and each step ends with a host round trip).  With
step graphs on, the runner replays each member's step as a captured HIP
graph (captured on the first two steps, one per persistent buffer set), so
runs of 1 to 6 steps check the eager result against 0 to 4 replays.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import cabi
from cabi import HipSolver

pytestmark = pytest.mark.gpu
sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def small_tanks():
    import bench

    tanks = bench.c4_tanks(cabi, count=12, seed=7)
    tanks.sort(key=lambda t: len(t["ps"]))
    return tanks[:4]


def make(tanks, graphs=False):
    out = []
    for t in tanks:
        s = HipSolver(t["cfg"], t["fluid"])
        s.set_particles(t["ps"], t["lo"], t["hi"])
        s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
        if graphs:
            s.set_step_graphs(True)
        out.append(s)
    return out


@pytest.mark.parametrize("graphs,steps", [(False, 6), (True, 1), (True, 2), (True, 3), (True, 6)])
def test_ensemble_matches_per_tank_loop(graphs, steps):
    print("test_ensemble_matches_per_tank_loop", graphs, steps, flush=True)
    tanks = small_tanks()
    print([t["desc"] for t in tanks], flush=True)
    a, b = make(tanks), make(tanks, graphs)
    for s in a:
        s.run_steps(steps)
    cabi.hip_ensemble_run_steps(b, steps)
    for sa, sb in zip(a, b):
        assert sa.step_count == sb.step_count == steps
        assert sa.num_particles == sb.num_particles
        assert sa.time == sb.time
        for f in ("x", "v", "m", "T", "id", "phase"):
            fa, fb = sa.field(f), sb.field(f)
            assert np.array_equal(fa.view(np.uint8), fb.view(np.uint8)), f
        da, db = sa.diagnostics(), sb.diagnostics()
        assert bytes(da) == bytes(db)


def test_ensemble_rejects_bad_arguments():
    tanks = small_tanks()[:1]
    s = make(tanks)
    with pytest.raises(RuntimeError):
        cabi.hip_ensemble_run_steps([s[0], s[0]], 1)
    with pytest.raises(RuntimeError):
        cabi.hip_ensemble_run_steps(s, -1)


def _same(sa, sb):
    assert sa.step_count == sb.step_count
    assert sa.num_particles == sb.num_particles
    assert sa.time == sb.time
    for f in ("x", "v", "m", "T", "id", "phase", "rho", "a"):
        fa, fb = sa.field(f), sb.field(f)
        assert np.array_equal(fa.view(np.uint8), fb.view(np.uint8)), f
    assert bytes(sa.diagnostics()) == bytes(sb.diagnostics())


def test_ensemble_calls_in_a_row_and_plain_steps_between():
    """Graphs live for one call: plain steps between runner calls, a member
    switched back to plain launches, and graph replay inside run_steps keep
    the loop's bits."""
    print("test_ensemble_calls_in_a_row_and_plain_steps_between", flush=True)
    tanks = small_tanks()[:2]
    a, b = make(tanks), make(tanks, graphs=True)
    for s in a:
        s.run_steps(9)
    cabi.hip_ensemble_run_steps(b, 4)
    for s in b:
        s.run_steps(1)
    b[0].set_step_graphs(False)
    cabi.hip_ensemble_run_steps(b, 2)
    for s in b:
        s.run_steps(2)  # b[1] replays graphs in run_steps too
    for sa, sb in zip(a, b):
        _same(sa, sb)


def test_launch_variants_give_the_same_bits_and_autotune_restores_the_state():
    """Every launch variant of the neighbour passes (LDS cell tiles or list
    gathers, for forces and interface) steps a tank to the same bits; the
    autotuner's trial steps leave no trace in the solver it tunes."""
    print("test_launch_variants_give_the_same_bits_and_autotune_restores_the_state", flush=True)
    tanks = small_tanks()[:2]
    for t in tanks:
        print(t["desc"], flush=True)
        ref = make([t])[0]
        ref.run_steps(6)
        for v in (1, 2, 3):
            s = make([t])[0]
            s.set_launch_variant(v)
            assert s.launch_variant() == v
            s.run_steps(6)
            _same(ref, s)
        s = make([t])[0]
        s.run_steps(2)
        v, ms = s.autotune(2)
        print("autotune", v, ms, flush=True)
        assert v in (0, 1, 2, 3) and s.launch_variant() == v and min(ms) > 0
        assert ms[v] == min(ms)
        assert s.step_count == 2
        s.run_steps(4)
        _same(ref, s)
    with pytest.raises(ValueError):
        make(tanks[:1])[0].set_launch_variant(4)

"""Ensemble runner (cryo_ensemble_run_steps): the tanks of an ensemble give the
bitwise same particles and diagnostics as the same tanks stepped one by one
with cryo_solver_run_steps, including tanks with phase change (where the
particle count changes and each step ends with a host round trip).
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


def make(tanks):
    out = []
    for t in tanks:
        s = HipSolver(t["cfg"], t["fluid"])
        s.set_particles(t["ps"], t["lo"], t["hi"])
        s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
        out.append(s)
    return out


def test_ensemble_matches_per_tank_loop():
    print("test_ensemble_matches_per_tank_loop", flush=True)
    tanks = small_tanks()
    print([t["desc"] for t in tanks], flush=True)
    a, b = make(tanks), make(tanks)
    for s in a:
        s.run_steps(6)
    cabi.hip_ensemble_run_steps(b, 6)
    for sa, sb in zip(a, b):
        assert sa.step_count == sb.step_count == 6
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

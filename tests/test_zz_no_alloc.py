"""cryo_solver_probe_pressure_noalloc, the allocation-free probe: through the
C ABI its results equal cryo_solver_probe_pressure's bit for bit, on two
calls in a row (the second with fewer points, so the fixed buffer still holds
the first call's contents beyond them), sizes beyond its limit and an
uninitialised solver are refused."""
import ctypes as C

import numpy as np
import pytest

import cabi

pytestmark = pytest.mark.gpu
LIMIT = 65536


def probe(s, fn, pts):
    out = np.full(len(pts), np.nan, np.float32)
    p = np.ascontiguousarray(pts, np.float32)
    rc = getattr(s.lib, fn)(s.h, C.c_int32(len(p)), p.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return rc, out


def test_noalloc_probe_equals_the_allocating_probe():
    print("test_noalloc_probe_equals_the_allocating_probe", flush=True)
    fluid = cabi.hip_fluid_preset("LN2")
    cfg = cabi.default_config(dx=0.004, gravity=(0, 0, -9.81e-4), heater_power=20.0, thermo_mode=1)
    s = cabi.HipSolver(cfg, fluid)
    s.build_tank(cabi.cylinder(0.05, 0.12), cabi.default_fill(T_liquid=fluid.T_sat + 2.0))
    rc, _ = probe(s, "cryo_solver_probe_pressure_noalloc", np.zeros((4, 3)))
    assert rc == cabi.ERR_RUNTIME and b"not initialised" in s.lib.cryo_last_error()
    s.run_steps(3)
    rng = np.random.default_rng(11)
    for n in (5000, 37):  # the second call is smaller: stale entries past 37 must not leak
        pts = rng.uniform((-0.06, -0.06, -0.07), (0.06, 0.06, 0.07), (n, 3))
        pts[: n // 10] = (1.0, 1.0, 1.0)  # far outside: no liquid in range
        rc0, ref = probe(s, "cryo_solver_probe_pressure", pts)
        rc1, got = probe(s, "cryo_solver_probe_pressure_noalloc", pts)
        assert rc0 == rc1 == 0
        assert np.isnan(ref).sum() >= n // 10 and np.isfinite(ref).sum() > 0
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), n
        s.run_steps(2)
    rc, _ = probe(s, "cryo_solver_probe_pressure_noalloc", np.zeros((LIMIT + 1, 3)))
    assert rc == cabi.ERR_OUT_OF_RANGE and b"65536" in s.lib.cryo_last_error()
    rc, out = probe(s, "cryo_solver_probe_pressure_noalloc", np.zeros((0, 3)))
    assert rc == 0

"""The drop-in boundary without a GPU: libcryosph.so loads, exports every entry
point include/cryosph.h declares, and its host-side parts (config defaults,
fluid presets, Clausius-Clapeyron, geometry, lattice generator, manoeuvre
table) agree with the reference build bit for bit.  The CryoSPH module imports
and mirrors the reference's pybind surface (bindings.cpp:133-499).
"""
import ctypes as C
import math
import re
import subprocess

import numpy as np
import pytest

import cabi
from cabi import HIP_LIB, REF_LIB, ROOT

HAVE_REF = REF_LIB[32].exists()
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="reference CPU build absent (make -C oracle ref)")


def declared_symbols():
    text = (ROOT / "include" / "cryosph.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cryo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    out = subprocess.run(["nm", "-D", "--defined-only", str(HIP_LIB)], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s+(cryo_\w+)", out.stdout))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    lib = C.CDLL(str(HIP_LIB))
    for s in declared_symbols():
        assert hasattr(lib, s)


def test_no_device_is_a_loud_error():
    lib = cabi.HipSolver.load()
    if lib.cryo_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        cabi.HipSolver(cabi.default_config(), cabi.hip_fluid_preset("LN2"))
    # argument validation comes first (sph_solver.cu:273-276)
    with pytest.raises(ValueError):
        cabi.HipSolver(cabi.default_config(contact_angle_deg=200.0), cabi.hip_fluid_preset("LN2"))
    with pytest.raises(ValueError):
        cabi.HipSolver(cabi.default_config(dx=0.0), cabi.hip_fluid_preset("LN2"))


@needs_ref
def test_presets_and_defaults_match_reference():
    for name in ("LH2", "LN2", "LOX", "LCH4", "H2O", "custom"):
        assert bytes(cabi.fluid_preset(name)) == bytes(cabi.hip_fluid_preset(name)), name
    assert bytes(cabi.hip_config_default()) == bytes(cabi.default_config())
    lib = cabi.HipSolver.load()
    ref = cabi.CpuSolver.load("ref", 32)
    for name in ("LN2", "LH2"):
        f = cabi.hip_fluid_preset(name)
        for p in (5e4, 101325.0, 2.5e5):
            assert lib.cryo_fluid_saturation_temperature(C.byref(f), p) == np.float32(ref.ref_saturation_temperature(
                C.byref(f), C.c_double(np.float32(p))))
        for T in (f.T_sat - 2, f.T_sat, f.T_sat + 3):
            assert lib.cryo_fluid_saturation_pressure(C.byref(f), T) == np.float32(ref.ref_saturation_pressure(
                C.byref(f), C.c_double(np.float32(T))))


GEOMETRIES = {
    "cylinder": (lambda: cabi.cylinder(0.05, 0.12), 0.004, dict()),
    "box": (lambda: cabi.box(0.04, 0.04, 0.04), 0.004, dict(T_liquid=80.0)),
    "sphere": (lambda: cabi.sphere(0.1), 0.008, dict(fill_fraction=0.3)),
    "capsule_bubble_heater": (lambda: cabi.capsule(0.05, 0.1), 0.005,
                              dict(mode=1, heater_zmin=-0.02, heater_zmax=0.02)),
    "full_box": (lambda: cabi.box(0.1, 0.06, 0.08), 0.01, dict(mode=2)),
    "csg": (lambda: cabi.geo_nodes([(cabi.GEO_BOX, -1, -1, (0.1, 0.1, 0.1)),
                                    (cabi.GEO_TRANSLATED, 0, -1, (0.05, 0.0, 0.05)),
                                    (cabi.GEO_BOX, -1, -1, (0.02, 0.02, 0.02)),
                                    (cabi.GEO_TRANSLATED, 2, -1, (0.05, 0.0, 0.01)),
                                    (cabi.GEO_DIFFERENCE, 1, 3, ())]), 0.005, dict(settle_direction=(1, 0, -1))),
}


@needs_ref
@pytest.mark.parametrize("name", sorted(GEOMETRIES))
def test_lattice_generator_matches_reference(name):
    make, dx, fill = GEOMETRIES[name]
    fluid = cabi.hip_fluid_preset("LN2")
    r = cabi.cpu_build_tank_particles("ref", make(), fluid, dx, cabi.default_fill(**fill))
    h = cabi.hip_build_tank_particles(make(), fluid, dx, cabi.default_fill(**fill))
    assert len(r) == len(h) > 0
    for f in ("x", "v", "nrm", "m", "rho", "T", "phase", "flags", "id"):
        assert np.array_equal(getattr(r, f).view(np.uint8), getattr(h, f).view(np.uint8)), f


def test_c1_lattice_count():
    """LN2 cylinder tank r=0.05 h=0.12 dx=4 mm: 26,360 particles = 14,520 fluid + 11,840 wall."""
    ps = cabi.hip_build_tank_particles(cabi.cylinder(0.05, 0.12), cabi.hip_fluid_preset("LN2"), 0.004,
                                       cabi.default_fill())
    assert len(ps) == 26360 and int((ps.phase == cabi.WALL).sum()) == 11840


def test_maneuver_profile_interpolation():
    """test_maneuver.cpp:34-53."""
    lib = cabi.HipSolver.load()
    t = np.array([0.0, 1.0, 3.0])
    acc = np.array([[0, 0, 0], [2, 0, 0], [2, 0, 0]], np.float32)
    om = np.array([[0, 0, 0], [0, 0, 1], [0, 0, 0]], np.float32)
    a = (C.c_float * 3)()
    w = (C.c_float * 3)()
    wd = (C.c_float * 3)()

    def ev(tt):
        assert lib.cryo_maneuver_evaluate(3, t.ctypes.data_as(C.c_void_p), acc.ctypes.data_as(C.c_void_p),
                                          om.ctypes.data_as(C.c_void_p), C.c_double(tt), a, w, wd) == 0
        return list(a), list(w), list(wd)

    A, W, WD = ev(0.5)
    assert abs(A[0] - 1.0) < 1e-6 and abs(W[2] - 0.5) < 1e-6 and abs(WD[2] - 1.0) < 1e-6
    A, W, WD = ev(2.0)
    assert abs(W[2] - 0.5) < 1e-6 and abs(WD[2] + 0.5) < 1e-6
    A, W, WD = ev(10.0)
    assert abs(A[0] - 2.0) < 1e-6 and WD[2] == 0.0
    bad = np.array([0.0, 0.0, 1.0])
    assert lib.cryo_maneuver_evaluate(3, bad.ctypes.data_as(C.c_void_p), acc.ctypes.data_as(C.c_void_p), None,
                                      C.c_double(0.5), a, w, wd) == -1


def test_python_module_surface_without_gpu():
    """test_bindings.py:25-54 parts that need no device."""
    import CryoSPH as cs
    assert cs.__version__ and isinstance(cs.cuda_backend, bool)
    ln2 = cs.FluidProperties.liquid_nitrogen()
    assert ln2.name == "LN2" and ln2.rho_liquid / ln2.rho_vapor > 100
    assert ln2.saturation_temperature(ln2.p_sat) == pytest.approx(ln2.T_sat, rel=1e-6)
    p2 = ln2.saturation_pressure(ln2.T_sat + 2.0)
    assert p2 > ln2.p_sat and ln2.saturation_temperature(p2) == pytest.approx(ln2.T_sat + 2.0, rel=1e-5)
    tank = cs.CylinderTank(radius=0.1, height=0.3)
    assert tank.sdf(0, 0, 0) == pytest.approx(-0.1)
    assert tank.volume() == pytest.approx(math.pi * 0.01 * 0.3)
    lo, hi = tank.bounds
    assert lo[2] == pytest.approx(-0.15) and hi[0] == pytest.approx(0.1)
    cfg = cs.SolverConfig()
    cfg.contact_angle_deg = 12.5
    cfg.surface_tension = cs.SurfaceTensionModel.Akinci
    cfg.gravity = (0.0, 0.0, 0.0)
    assert cfg.gravity == (0.0, 0.0, 0.0) and cfg.contact_angle_deg == 12.5
    assert cfg.surface_tension == cs.SurfaceTensionModel.Akinci
    box = cs.Translated(cs.BoxTank(0.1, 0.1, 0.1), (0.05, 0.0, 0.05))
    block = cs.Translated(cs.BoxTank(0.02, 0.02, 0.02), (0.05, 0.0, 0.01))
    geo = cs.Difference(box, block)
    assert geo.sdf(0.05, 0.0, 0.01) > 0 and geo.sdf(0.02, 0.0, 0.05) < 0
    assert cs.Union(box, block).sdf(0.05, 0.0, 0.01) < 0
    prof = cs.ManeuverProfile([0.0, 1.0], np.array([[0, 0, 0], [2.0, 0, 0]]), np.array([[0, 0, 0], [0, 0, 1.0]]))
    a, w, wd = prof.evaluate(0.5)
    assert a == pytest.approx((1.0, 0.0, 0.0)) and w[2] == pytest.approx(0.5) and wd[2] == pytest.approx(1.0)
    with pytest.raises(ValueError):
        cs.ManeuverProfile([0.0, 0.0], np.zeros((2, 3)))
    with pytest.raises(ValueError):
        cs.Solver(cfg if False else _bad_angle(cs), ln2)


def _bad_angle(cs):
    c = cs.SolverConfig()
    c.contact_angle_deg = 200.0
    return c


def test_device_lattice_rejects_too_deep_csg_before_any_device_work():
    """cryo_build_tank_particles_device walks the CSG tree with a bounded stack:
    a deeper tree is an invalid argument raised on the host, before any device
    call; the host generator takes the same tree."""
    fluid = cabi.hip_fluid_preset("LN2")
    nodes = [(cabi.GEO_BOX, -1, -1, (0.04, 0.04, 0.04))]
    nodes += [(cabi.GEO_TRANSLATED, k, -1, (0.0, 0.0, 0.0)) for k in range(40)]
    with pytest.raises(RuntimeError, match="too deep"):
        cabi.hip_build_tank_particles(cabi.geo_nodes(nodes), fluid, 0.004, cabi.default_fill(), device=True)
    ps = cabi.hip_build_tank_particles(cabi.geo_nodes(nodes), fluid, 0.004, cabi.default_fill())
    ref = cabi.hip_build_tank_particles(cabi.box(0.04, 0.04, 0.04), fluid, 0.004, cabi.default_fill())
    assert np.array_equal(ps.x, ref.x) and np.array_equal(ps.phase, ref.phase)

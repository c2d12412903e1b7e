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


def icosphere(radius, levels=1):
    """Closed triangle mesh of a sphere: the icosahedron, each face split in
    four `levels` times, vertices pushed out to `radius`."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(v, float) / np.linalg.norm(v) for v in verts]
    for _ in range(levels):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]
        nxt = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nxt
    return np.array(verts) * radius, np.array(faces, np.int32)


@needs_ref
@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.013, -0.007, 0.021)])
def test_mesh_lattice_matches_reference(shift):
    """A mesh tank (80-triangle icosphere: nearest-triangle distances and the
    winding-number sign) alone and translated inside a union with a box, settled
    30 % fill along a slanted direction: every field bit for bit."""
    verts, tris = icosphere(0.05)
    nodes = [(cabi.GEO_MESH, -1, -1, ())]
    if any(shift):
        nodes += [(cabi.GEO_TRANSLATED, 0, -1, shift), (cabi.GEO_BOX, -1, -1, (0.03, 0.03, 0.14)),
                  (cabi.GEO_UNION, 1, 2, ())]
    arr = cabi.geo_nodes(nodes)
    arr[0].mesh_id = 0
    fluid = cabi.hip_fluid_preset("LN2")
    fill = cabi.default_fill(fill_fraction=0.3, settle_direction=(0.3, 0.2, -1.0))
    r = cabi.cpu_build_tank_particles("ref", arr, fluid, 0.006, fill, meshes=[(verts, tris)])
    h = cabi.hip_build_tank_particles(arr, fluid, 0.006, fill, meshes=[(verts, tris)])
    assert len(r) == len(h) > 1000
    assert int((h.phase == cabi.WALL).sum()) > 0 and int((h.phase == cabi.VAPOR).sum()) > 0
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


def test_python_module_agrees_with_the_c_abi_without_gpu():
    """The CryoSPH module's device-free surface is the C ABI's: every
    SolverConfig default, the fluid presets and their saturation curve, the
    geometry classes' distances / bounds / volumes at random points, and
    ManeuverProfile.evaluate, value for value."""
    import CryoSPH as cs
    lib = cabi.HipSolver.load()
    assert cs.__version__ == lib.cryo_version().decode()
    cfg, ref = cs.SolverConfig(), cabi.hip_config_default()
    for name, _ in cabi._CONFIG_FIELDS:
        got = getattr(cfg, name)
        want = getattr(ref, name)
        if name == "gravity":
            assert got == pytest.approx(want.tup()), name
        else:
            assert float(int(got) if not isinstance(got, float) else got) == pytest.approx(float(want)), name
    for make, name in ((cs.FluidProperties.liquid_nitrogen, "LN2"), (cs.FluidProperties.liquid_hydrogen, "LH2"),
                       (cs.FluidProperties.liquid_oxygen, "LOX"), (cs.FluidProperties.water, "H2O")):
        mod, c = make(), cabi.hip_fluid_preset(name)
        for k in ("rho_liquid", "rho_vapor", "mu_liquid", "k_liquid", "cp_vapor", "sigma", "h_fg", "T_sat", "p_sat"):
            assert np.float32(getattr(mod, k)) == getattr(c, k), (name, k)
        for T in (c.T_sat - 3.0, c.T_sat + 0.5, c.T_sat + 4.0):
            p = mod.saturation_pressure(T)
            assert np.float32(p) == lib.cryo_fluid_saturation_pressure(C.byref(c), np.float32(T))
            assert mod.saturation_temperature(p) == pytest.approx(T, rel=2e-5)
    box = cs.Translated(cs.BoxTank(0.1, 0.08, 0.12), (0.05, -0.01, 0.06))
    block = cs.Translated(cs.SphereTank(0.03), (0.06, 0.0, 0.02))
    shapes = {
        "cylinder": (cs.CylinderTank(radius=0.07, height=0.2), cabi.cylinder(0.07, 0.2), math.pi * 0.07 ** 2 * 0.2),
        "capsule": (cs.CapsuleTank(0.04, 0.1), cabi.capsule(0.04, 0.1),
                    math.pi * 0.04 ** 2 * 0.1 + 4 / 3 * math.pi * 0.04 ** 3),
        "difference": (cs.Difference(box, block), cabi.geo_nodes([
            (cabi.GEO_BOX, -1, -1, (0.1, 0.08, 0.12)), (cabi.GEO_TRANSLATED, 0, -1, (0.05, -0.01, 0.06)),
            (cabi.GEO_SPHERE, -1, -1, (0.03,)), (cabi.GEO_TRANSLATED, 2, -1, (0.06, 0.0, 0.02)),
            (cabi.GEO_DIFFERENCE, 1, 3, ())]), None),
        "union": (cs.Union(box, block), cabi.geo_nodes([
            (cabi.GEO_BOX, -1, -1, (0.1, 0.08, 0.12)), (cabi.GEO_TRANSLATED, 0, -1, (0.05, -0.01, 0.06)),
            (cabi.GEO_SPHERE, -1, -1, (0.03,)), (cabi.GEO_TRANSLATED, 2, -1, (0.06, 0.0, 0.02)),
            (cabi.GEO_UNION, 1, 3, ())]), None),
    }
    pts = np.random.default_rng(5).uniform(-0.15, 0.2, (64, 3))
    for name, (geo, nodes, vol) in shapes.items():
        out = C.c_double()
        for q in pts:
            assert lib.cryo_geo_sdf(len(nodes), nodes, 0, None, *map(C.c_double, q), C.byref(out)) == 0
            assert geo.sdf(*q) == out.value, name
        lo3, hi3 = (C.c_float * 3)(), (C.c_float * 3)()
        assert lib.cryo_geo_bounds(len(nodes), nodes, 0, None, lo3, hi3) == 0
        lo, hi = geo.bounds
        assert np.array_equal(np.float32(lo), np.array(lo3)) and np.array_equal(np.float32(hi), np.array(hi3))
        if vol is not None:
            assert geo.volume() == pytest.approx(vol, rel=1e-12)
    times = np.array([0.0, 0.4, 1.5, 2.0])
    acc = np.array([[0, 0, 0], [0.3, -0.1, 0.0], [0.3, 0.2, 0.05], [0.0, 0.0, 0.0]], np.float32)
    om = np.array([[0, 0, 0], [0, 0.01, 0], [0, 0.02, 0], [0, 0.02, 0]], np.float32)
    prof = cs.ManeuverProfile(times, acc, om)
    for t in (-1.0, 0.0, 0.2, 0.4, 1.0, 1.7, 2.0, 5.0):
        a, w, wd = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
        assert lib.cryo_maneuver_evaluate(len(times), times.ctypes.data_as(C.c_void_p), acc.ctypes.data_as(C.c_void_p),
                                          om.ctypes.data_as(C.c_void_p), C.c_double(t), a, w, wd) == 0
        ma, mw, mwd = prof.evaluate(t)
        assert np.array_equal(np.float32(ma), np.array(a)) and np.array_equal(np.float32(mw), np.array(w))
        assert np.array_equal(np.float32(mwd), np.array(wd)), t


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

"""File formats of the drop-in: the C ABI's VTK and JSON-frame writers
(cryo_write_vtk, cryo_write_json_frame, cryo_frame_speed) against the
reference's own writers (io.cpp:52-142, compiled in oracle/_ref), byte for
byte, on the same particle arrays; and a decode of each file back into the
arrays it was written from.  No device is needed."""
import ctypes as C
import json
import struct

import numpy as np
import pytest

import cabi

needs_ref = pytest.mark.skipif(not cabi.REF_LIB[32].exists(), reason="reference build (oracle/_ref) not present")


def particle_arrays(n, seed, with_odd_values=True):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 0.3, (n, 3)).astype(np.float32)
    v = rng.normal(0.0, 2.0, (n, 3)).astype(np.float32)
    T = rng.uniform(60.0, 90.0, n).astype(np.float32)
    m = rng.uniform(1e-7, 1e-5, n).astype(np.float32)
    rho = rng.uniform(1.0, 900.0, n).astype(np.float32)
    phase = rng.integers(0, 3, n).astype(np.uint8)
    p = rng.normal(0.0, 5e4, n).astype(np.float32)
    if with_odd_values and n >= 8:
        T[1], p[2], p[3], v[4] = np.nan, np.inf, -np.inf, (-0.0, 0.0, -0.0)
        x[5] = (1e-30, -3.4e38, 123456.789)
    return x, v, T, m, rho, phase, p


def vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def hip_lib():
    lib = cabi.HipSolver.load()
    lib.cryo_write_vtk.argtypes = [C.c_char_p, C.c_int32] + [C.c_void_p] * 7
    lib.cryo_write_json_frame.argtypes = [C.c_char_p, C.c_double, C.c_int64, C.c_int32] + [C.c_void_p] * 5 + \
        [C.c_int32]
    lib.cryo_frame_speed.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
    lib.cryo_frame_speed.restype = None
    return lib


def ref_lib():
    lib = cabi.CpuSolver.load("ref", 32)
    lib.ref_write_vtk.argtypes = [C.c_char_p, C.c_int] + [C.c_void_p] * 7
    lib.ref_write_json_frame.argtypes = [C.c_char_p, C.c_double, C.c_longlong, C.c_int] + [C.c_void_p] * 5 + \
        [C.c_int]
    return lib


def write_both_vtk(tmp_path, arrays, with_pressure):
    x, v, T, m, rho, phase, p = arrays
    p = p if with_pressure else None
    ours, theirs = tmp_path / "ours.vtk", tmp_path / "ref.vtk"
    assert hip_lib().cryo_write_vtk(str(ours).encode(), len(T), vp(x), vp(v), vp(T), vp(m), vp(rho), vp(phase),
                                    vp(p)) == 0
    assert ref_lib().ref_write_vtk(str(theirs).encode(), len(T), vp(x), vp(v), vp(T), vp(m), vp(rho), vp(phase),
                                   vp(p)) == 0
    return ours.read_bytes(), theirs.read_bytes()


@needs_ref
@pytest.mark.parametrize("n,with_pressure", [(0, True), (1, False), (517, True), (517, False), (20000, True)])
def test_vtk_bytes_match_reference_writer(tmp_path, n, with_pressure):
    ours, theirs = write_both_vtk(tmp_path, particle_arrays(n, seed=n), with_pressure)
    assert ours == theirs


@needs_ref
@pytest.mark.parametrize("n,include_walls", [(0, False), (3, True), (517, False), (517, True), (20000, False)])
def test_json_frame_bytes_match_reference_writer(tmp_path, n, include_walls):
    x, v, T, m, rho, phase, p = particle_arrays(n, seed=100 + n)
    lib = hip_lib()
    speed = np.empty(n, np.float32)
    lib.cryo_frame_speed(n, vp(v), vp(speed))
    ours, theirs = tmp_path / "ours.json", tmp_path / "ref.json"
    time, step = 0.123456789012, 98765
    assert lib.cryo_write_json_frame(str(ours).encode(), time, step, n, vp(x), vp(T), vp(p), vp(speed), vp(phase),
                                     int(include_walls)) == 0
    assert ref_lib().ref_write_json_frame(str(theirs).encode(), time, step, n, vp(x), vp(v), vp(T), vp(p),
                                          vp(phase), int(include_walls)) == 0
    assert ours.read_bytes() == theirs.read_bytes()


def test_vtk_decodes_to_its_arrays(tmp_path):
    """The legacy binary layout read back field by field (big-endian)."""
    x, v, T, m, rho, phase, p = particle_arrays(300, seed=7, with_odd_values=False)
    path = tmp_path / "f.vtk"
    assert hip_lib().cryo_write_vtk(str(path).encode(), 300, vp(x), vp(v), vp(T), vp(m), vp(rho), vp(phase),
                                    vp(p)) == 0
    raw = path.read_bytes()
    at = 0

    def line():
        nonlocal at
        end = raw.index(b"\n", at)
        s = raw[at:end].decode()
        at = end + 1
        return s

    def floats(count):
        nonlocal at
        out = np.frombuffer(raw, ">f4", count, at).astype(np.float32)
        at += 4 * count
        return out
    assert [line() for _ in range(4)] == ["# vtk DataFile Version 3.0", "CryoSPH-CUDA particles", "BINARY",
                                          "DATASET POLYDATA"]
    assert line() == "POINTS 300 float"
    assert np.array_equal(floats(900).reshape(-1, 3), x)
    assert line() == "" and line() == "VERTICES 300 600"
    cells = np.frombuffer(raw, ">i4", 600, at)
    at += 2400
    assert np.array_equal(cells[0::2], np.ones(300)) and np.array_equal(cells[1::2], np.arange(300))
    assert line() == "" and line() == "POINT_DATA 300"
    for name, ref in (("temperature", T), ("mass", m), ("density", rho), ("phase", phase.astype(np.float32)),
                      ("pressure", p)):
        assert line() == f"SCALARS {name} float 1" and line() == "LOOKUP_TABLE default"
        assert np.array_equal(floats(300), ref), name
        assert line() == ""
    assert line() == "VECTORS velocity float"
    assert np.array_equal(floats(900).reshape(-1, 3), v)
    assert line() == "" and at == len(raw)


def test_json_frame_decodes_to_its_arrays(tmp_path):
    """Walls left out by default; 6 significant digits; non-finite values as 0;
    speed is the float32 norm of the velocity."""
    x, v, T, m, rho, phase, p = particle_arrays(400, seed=11)
    lib = hip_lib()
    speed = np.empty(400, np.float32)
    lib.cryo_frame_speed(400, vp(v), vp(speed))
    assert np.array_equal(speed, np.sqrt((v * v).sum(1, dtype=np.float32), dtype=np.float32))
    path = tmp_path / "f.json"
    assert lib.cryo_write_json_frame(str(path).encode(), 1.5, 42, 400, vp(x), vp(T), vp(p), vp(speed), vp(phase),
                                     0) == 0
    frame = json.loads(path.read_text())
    keep = phase != cabi.WALL
    assert frame["time"] == 1.5 and frame["step"] == 42 and frame["n"] == int(keep.sum())
    assert frame["phase"] == phase[keep].tolist()
    np.testing.assert_allclose(np.array(frame["position"]).reshape(-1, 3), x[keep], rtol=1e-5)
    clean = lambda a: np.where(np.isfinite(a), a, 0.0)  # noqa: E731
    np.testing.assert_allclose(frame["temperature"], clean(T[keep]), rtol=1e-5)
    np.testing.assert_allclose(frame["pressure"], clean(p[keep]), rtol=1e-5, atol=1e-30)
    np.testing.assert_allclose(frame["speed"], speed[keep], rtol=1e-5)


def test_writers_report_errors(tmp_path):
    lib = hip_lib()
    bad = str(tmp_path / "no_such_dir" / "f.vtk").encode()
    x, v, T, m, rho, phase, p = particle_arrays(4, seed=1)
    assert lib.cryo_write_vtk(bad, 4, vp(x), vp(v), vp(T), vp(m), vp(rho), vp(phase), vp(p)) == cabi.ERR_RUNTIME
    assert b"cannot open" in lib.cryo_last_error()
    assert lib.cryo_write_vtk(b"x.vtk", 4, None, vp(v), vp(T), vp(m), vp(rho), vp(phase), vp(p)) == \
        cabi.ERR_INVALID_ARGUMENT
    assert struct.calcsize("f") == 4

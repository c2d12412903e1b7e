"""Pins the CPU oracle (oracle/cryo_oracle.cpp) before it is trusted as the checker.

* against the golden vectors written by the reference's own CPU path
  (tests/golden/make_golden.py): sort permutation, cell tables, neighbour
  counts/id sums, every field after start-up and after 20 steps;
* against the live reference build (oracle/_ref), bit for bit, over the
  configurations the reference's tests exercise (maneuvers, continuity,
  Akinci, SPS, shifting, wetting models, venting, splitting, relax), in FP32
  and FP64;
* against closed forms: kernel normalisation and derivative
  (test_kernels.cpp:15-63), stable-sort and brute-force neighbour sets
  (test_spatial_hash.cpp:25-93).
"""
import ctypes as C
import math

import numpy as np
import pytest

from cabi import (FIELDS, REF_LIB, WALL, CpuSolver, ParticleSet, box, capsule, cpu_build_tank_particles, cylinder,
                  default_config, default_fill, fluid_preset, sphere)
from pathlib import Path

GOLD = Path(__file__).resolve().parent / "golden"
HAVE_REF = REF_LIB[32].exists()
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="reference CPU build absent (make -C oracle ref)")


def threads(n=8):
    CpuSolver.load("orc", 32).orc_set_threads(n)
    CpuSolver.load("orc", 64).orc_set_threads(n)
    if HAVE_REF:
        CpuSolver.load("ref", 32).ref_set_threads(n)
        CpuSolver.load("ref", 64).ref_set_threads(n)


def same(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint8),
                                                 np.ascontiguousarray(b).view(np.uint8))


def ln2():
    return fluid_preset("LN2") if HAVE_REF else _ln2_literal()


def _ln2_literal():
    from cabi import Fluid
    return Fluid(name=b"LN2", rho_liquid=806.1, rho_vapor=4.61, mu_liquid=1.61e-4, mu_vapor=5.4e-6, k_liquid=0.1446,
                 k_vapor=0.0074, cp_liquid=2042.0, cp_vapor=1123.0, sigma=8.85e-3, h_fg=199.2e3, T_sat=77.36,
                 p_sat=101325.0, R_gas=296.8)


# ------------------------------------------------------------------ golden
def test_oracle_hash_sort_matches_reference_golden():
    g = np.load(GOLD / "hash_sort.npz")
    lib = CpuSolver.load("orc", 32)
    x = g["x"]
    for tag in ("alive", "dead"):
        ph = g[f"{tag}_phase"]
        n = len(x)
        perm = np.zeros(n, np.uint32)
        cnt = np.zeros(n, np.int32)
        sums = np.zeros(n, np.uint64)
        na = C.c_int()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        lo = np.zeros(3, np.float32)
        hi = np.ones(3, np.float32)
        assert lib.orc_hash_sort(C.c_int(n), vp(x), vp(ph), vp(lo), vp(hi), C.c_double(np.float32(0.08)), vp(perm),
                                 C.byref(na), C.c_double(np.float32(0.08)), vp(cnt), vp(sums)) == 0
        k = na.value
        assert k == int(g[f"{tag}_n_alive"])
        assert np.array_equal(perm[:k], g[f"{tag}_perm"])
        assert np.array_equal(cnt[:k], g[f"{tag}_counts"])
        assert np.array_equal(sums[:k], g[f"{tag}_id_sums"])


def _golden_box_solver(bits=32):
    g = np.load(GOLD / "box_tank.npz")
    ps = ParticleSet(*(g["in_" + k] for k in ("x", "v", "nrm", "m", "rho", "T", "phase", "flags", "id")))
    cfg = default_config(dx=0.004, gravity=(0, 0, -0.5), contact_angle_deg=10, velocity_scale=1.0, thermo_mode=1,
                         lee_evaporation=5.0, lee_condensation=5.0)
    s = CpuSolver("orc", cfg, ln2(), bits)
    s.set_particles(ps, g["lo"], g["hi"])
    return g, s


def test_oracle_box_tank_matches_reference_golden():
    threads(8)
    g, s = _golden_box_solver()
    s.next_dt()
    for f in FIELDS:
        assert same(s.field(f), g["init_" + f]), f"start-up field {f}"
    diag = []
    for _ in range(20):
        s.step()
        d = s.diagnostics()
        diag.append([d.time, d.mass_liquid, d.mass_vapor, d.enthalpy, d.evaporated_mass, d.p_ullage, d.T_sat,
                     d.max_speed, d.slosh_force.x, d.slosh_force.y, d.slosh_force.z])
    for f in ("x", "v", "m", "T", "rho", "id", "phase", "a", "p"):
        assert same(s.field(f), g["step20_" + f]), f"step-20 field {f}"
    assert np.array_equal(np.array(diag), g["diag"])


def test_oracle_c1_grid_matches_reference_golden():
    g = np.load(GOLD / "c1_grid.npz")
    n = len(g["in_x"])
    fluid = ln2()
    cfg = default_config(dx=0.004, gravity=(0, 0, -9.81e-4), heater_power=20.0, thermo_mode=1)
    zeros = np.zeros((n, 3), np.float32)
    ps = ParticleSet(g["in_x"], zeros, zeros, np.ones(n), np.full(n, 800.0), np.full(n, 77.36), g["in_phase"],
                     None, None)
    s = CpuSolver("orc", cfg, fluid, 32)
    s.set_particles(ps, g["lo"], g["hi"])
    s.next_dt()
    assert np.array_equal(s.field("id"), g["sorted_id"])
    cs, ce = s.grid_tables()
    assert np.array_equal(cs, g["cell_start"]) and np.array_equal(ce, g["cell_end"])
    cnt, sums = s.neighbor_stats(float(g["support"]))
    assert np.array_equal(cnt, g["counts"]) and np.array_equal(sums, g["id_sums"])


# -------------------------------------------------------- closed forms
@pytest.mark.parametrize("ktype", [0, 1])
def test_kernel_closed_forms(ktype):
    lib = CpuSolver.load("orc", 64)
    W = C.c_double()
    dW = C.c_double()

    def ev(h, r):
        lib.orc_kernel_eval(C.c_int(ktype), C.c_double(h), C.c_double(r), C.byref(W), C.byref(dW))
        return W.value, dW.value

    # lattice partition of unity (test_kernels.cpp:15-27)
    dx = 0.01
    h = 1.3 * dx
    s = sum(ev(h, dx * math.sqrt(i * i + j * j + k * k))[0] for i in range(-4, 5) for j in range(-4, 5)
            for k in range(-4, 5))
    assert abs(s * dx ** 3 - 1.0) < 0.02
    # continuous normalisation 4 pi int r^2 W dr = 1 (:30-40)
    steps = 20000
    dr = 2 * 0.5 / steps
    acc = sum(4 * math.pi * ((q + 0.5) * dr) ** 2 * ev(0.5, (q + 0.5) * dr)[0] * dr for q in range(steps))
    assert abs(acc - 1.0) < 1e-4
    # dW/dr against central differences, non-positive (:42-50)
    norm = ev(1.0, 0.0)[0] if ktype == 1 else 1.0 / math.pi
    for r in np.arange(0.05, 1.95, 0.1):
        fd = (ev(1.0, r + 1e-3)[0] - ev(1.0, r - 1e-3)[0]) / 2e-3
        assert abs(ev(1.0, r)[1] - fd) < 1e-3 * norm + 1e-4
        assert ev(1.0, r)[1] <= 0
    # compact support (:52-63)
    assert ev(0.1, 0.2) == (0.0, 0.0) and ev(0.1, 0.25)[0] == 0.0


@pytest.mark.parametrize("ktype", [0, 1])
def test_adjoint_kernel_functions_are_the_kernel_derivatives(ktype):
    """The adjoint's kernel functions (csrc/adjoint_pair.hpp: W, s = W'/r and
    s'/r, shared by the GPU and FP64 adjoints) against the forward kernel: W
    equal, W' and s' by central differences, across both branches of the
    cubic spline (q < 1 and 1 <= q < 2)."""
    lib = CpuSolver.load("orc", 64)
    W, dW, s, spr = (C.c_double() for _ in range(4))
    h = 0.013

    def fwd(r):
        lib.orc_kernel_eval(C.c_int(ktype), C.c_double(h), C.c_double(r), C.byref(W), C.byref(dW))
        return W.value, dW.value

    def adj(r):
        lib.orc_adjoint_kernel_eval(C.c_int(ktype), C.c_double(h), C.c_double(r), C.byref(W), C.byref(s),
                                    C.byref(spr))
        return W.value, s.value, spr.value

    scale = abs(adj(0.5 * h)[1])
    for q in np.arange(0.05, 1.96, 0.05):
        if abs(q - 1.0) < 1e-9:
            continue
        r = q * h
        w_f, dw_f = fwd(r)
        w_a, s_a, spr_a = adj(r)
        assert w_a == pytest.approx(w_f, rel=1e-13, abs=1e-300)
        assert s_a * r == pytest.approx(dw_f, rel=1e-12, abs=1e-12 * scale * h)
        e = 1e-6 * h
        ds = (adj(r + e)[1] - adj(r - e)[1]) / (2 * e)
        assert spr_a * r == pytest.approx(ds, rel=1e-6, abs=1e-6 * scale / h)


def test_oracle_sort_is_stable_and_neighbours_match_brute_force():
    """Stable sort by cell and brute-force neighbour sets (test_spatial_hash.cpp:25-93)."""
    lib = CpuSolver.load("orc", 32)
    rng = np.random.default_rng(11)
    n = 2000
    x = rng.random((n, 3), dtype=np.float32)
    ph = np.zeros(n, np.uint8)
    ph[::7] = 3
    perm = np.zeros(n, np.uint32)
    cnt = np.zeros(n, np.int32)
    sums = np.zeros(n, np.uint64)
    na = C.c_int()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    r = np.float32(0.1)
    assert lib.orc_hash_sort(C.c_int(n), vp(x), vp(ph), vp(np.zeros(3, np.float32)), vp(np.ones(3, np.float32)),
                             C.c_double(r), vp(perm), C.byref(na), C.c_double(r), vp(cnt), vp(sums)) == 0
    k = na.value
    assert k == n - int((ph == 3).sum())
    # stable sort: keys non-decreasing, ties in original order
    origin = np.float32(0) - r
    inv = np.float32(1) / r
    dims = [max(1, int(math.ceil(np.float32(1) / r)) + 2)] * 3
    c = np.clip(np.floor((x - origin) * inv).astype(np.int64), 0, dims[0] - 1)
    key = (c[:, 2] * dims[1] + c[:, 1]) * dims[0] + c[:, 0]
    alive = np.nonzero(ph != 3)[0]
    expect = alive[np.argsort(key[alive], kind="stable")]
    assert np.array_equal(perm[:k], expect)
    # brute force with the same float32 arithmetic
    xs = x[perm[:k]]
    for i in range(0, k, 37):
        d = xs - xs[i]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        inside = r2 < r * r
        assert cnt[i] == int(inside.sum())
        assert sums[i] == int(perm[:k][inside].astype(np.uint64).sum())


# --------------------------------------------------- vs reference build
def _tank(nodes, dx, fill, seed=0):
    ps = cpu_build_tank_particles("ref", nodes, ln2(), dx, fill)
    rng = np.random.default_rng(seed)
    fl = ps.phase != WALL
    ps.x[fl] += rng.uniform(-0.05 * dx, 0.05 * dx, size=(int(fl.sum()), 3)).astype(np.float32)
    return ps, ps.x.min(0) - np.float32(4 * dx), ps.x.max(0) + np.float32(4 * dx)


SCENARIOS = {
    "closed_pc": dict(cfg=dict(dx=0.004, gravity=(0, 0, -0.5), contact_angle_deg=10, velocity_scale=1.0,
                               thermo_mode=1, lee_evaporation=5.0, lee_condensation=5.0, split_ratio=1.05),
                      geo=lambda: box(0.04, 0.04, 0.04), fill=dict(T_liquid=80.36), steps=40),
    "vented_heater": dict(cfg=dict(dx=0.004, gravity=(0, 0, -1e-4), contact_angle_deg=5, velocity_scale=0.5,
                                   heater_power=5.0, lee_evaporation=0.5, lee_condensation=0.5, vent_relaxation=0.2),
                          geo=lambda: cylinder(0.03, 0.06), fill=dict(T_liquid=77.86, heater_zmin=-0.02,
                                                                      heater_zmax=0.0), steps=30),
    "continuity_sps": dict(cfg=dict(dx=0.005, density_mode=1, sps_turbulence=1, surface_tension=0, thermal=0,
                                    phase_change=0, artificial_viscosity=0.05),
                           geo=lambda: box(0.06, 0.04, 0.05), fill=dict(include_vapor=0), steps=30),
    "akinci": dict(cfg=dict(dx=0.005, surface_tension=2, thermal=0, phase_change=0, gravity=(0, 0, -0.3),
                            contact_angle_deg=60),
                   geo=lambda: sphere(0.03), fill=dict(include_vapor=0, fill_fraction=0.6), steps=25),
    "wetting_both_smooth": dict(cfg=dict(dx=0.004, wetting_model=2, csf_smooth_normals=1, csf_balanced=1,
                                         contact_angle_deg=30, gravity=(0, 0, -0.2), velocity_scale=0.5,
                                         wall_thermal_bc=1, thermo_mode=1),
                                geo=lambda: capsule(0.02, 0.03), fill=dict(T_liquid=78.0), steps=25),
    "shifting_tvf_stress": dict(cfg=dict(dx=0.005, particle_shifting=1, transport_stress=1, thermal=0,
                                         phase_change=0, surface_tension=1, velocity_scale=1.0),
                                geo=lambda: box(0.05, 0.05, 0.05), fill=dict(), steps=25),
    # the cubic-spline kernel (SPH_kernels.cuh:34-58), the Tait EOS with
    # gamma != 7 (its pow path) and the negative-pressure clamp
    # (sim_params.cuh:100-119), in a closed tank with phase change
    "cubic_tait5_clamp": dict(cfg=dict(dx=0.005, kernel=0, tait_gamma=5.0, clamp_negative_pressure=1,
                                       thermo_mode=1, lee_evaporation=5.0, lee_condensation=5.0,
                                       gravity=(0, 0, -0.3), velocity_scale=0.5),
                              geo=lambda: box(0.05, 0.05, 0.05), fill=dict(T_liquid=79.0), steps=25),
    # the transport velocity in the bulk too (transport_interface_only = 0)
    # and walls without the extrapolated colour (wall_colour = 0)
    "tvf_bulk_no_wall_colour": dict(cfg=dict(dx=0.004, transport_interface_only=0, wall_colour=0,
                                             contact_angle_deg=30, gravity=(0, 0, -0.2), velocity_scale=0.5,
                                             thermo_mode=1),
                                    geo=lambda: capsule(0.02, 0.03), fill=dict(T_liquid=78.0), steps=25),
}


@needs_ref
@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("name", sorted(SCENARIOS))
def test_oracle_bit_exact_with_reference(name, bits):
    threads(8)
    sc = SCENARIOS[name]
    fluid = ln2()
    cfg = default_config(**sc["cfg"])
    ps, lo, hi = _tank(sc["geo"](), sc["cfg"]["dx"], default_fill(**sc["fill"]))
    sims = [CpuSolver(k, cfg, fluid, bits) for k in ("ref", "orc")]
    for s in sims:
        s.set_particles(ps, lo, hi)
        s.next_dt()
    if name == "closed_pc":
        for s in sims:
            s.set_maneuver([0.0, 0.01, 0.02], [[0, 0, 0], [0.5, 0, 0], [0.5, 0.2, 0]], [[0, 0, 0], [0, 0, 1], [0, 0, 2]],
                           (0.0, 0.0, -0.01))
    for k in range(sc["steps"]):
        for s in sims:
            s.step()
    for f in FIELDS:
        assert same(sims[0].field(f), sims[1].field(f)), f"{name}/{bits}: field {f}"
    assert sims[0].diagnostics().as_dict() == sims[1].diagnostics().as_dict()


@needs_ref
def test_oracle_relax_and_run_until_match_reference():
    threads(8)
    fluid = ln2()
    cfg = default_config(dx=0.005, gravity=(0, 0, 0), velocity_scale=1.0, thermo_mode=1, lee_evaporation=2.0)
    ps, lo, hi = _tank(box(0.04, 0.04, 0.04), 0.005, default_fill(T_liquid=79.36))
    sims = [CpuSolver(k, cfg, fluid, 32) for k in ("ref", "orc")]
    for s in sims:
        s.set_particles(ps, lo, hi)
        s.relax(0.01)
        s.run_until(0.004)
        s.damp_velocities(0.5)
        s.set_heater_power(3.0)
        s.set_contact_angle(40.0)
        s.run_steps(5)
    for f in FIELDS:
        assert same(sims[0].field(f), sims[1].field(f)), f
    assert sims[0].diagnostics().as_dict() == sims[1].diagnostics().as_dict()
    assert sims[0].time == sims[1].time and sims[0].step_count == sims[1].step_count


def test_fp32_slosh_force_error_comes_from_float_positions():
    """Why an FP32 run's tank load differs from FP64's by 1e-4..1e-3 at
    microgravity (the batched-ensemble tanks): the first drift moves particles
    by a few 1e-8 m, a few float ulps of their 0.1 m coordinates, and the
    load follows those displacements.  The same FP64 run with only the
    positions rounded to float after each drift (orc_round_positions) shows an
    error of the same size, while one FP32 evaluation on identical input
    agrees with FP64 to ~1e-6.  Ensemble member 0 of bench.c4_tanks(seed=0)."""
    import bench
    import cabi
    threads(8)
    t = bench.c4_tanks(cabi, count=1, seed=0)[0]

    def loads(bits, rounded):
        lib = CpuSolver.load("orc", 64)
        lib.orc_round_positions(1 if rounded else 0)
        try:
            s = CpuSolver("orc", t["cfg"], t["fluid"], bits)
            s.set_particles(t["ps"], t["lo"], t["hi"])
            s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
            s.next_dt()
            out = [np.array(s.diagnostics().slosh_force.tup())]
            s.run_steps(5)
            return out + [np.array(s.diagnostics().slosh_force.tup())]
        finally:
            lib.orc_round_positions(0)

    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    f64, f64r, f32 = loads(64, False), loads(64, True), loads(32, False)
    assert rel(f32[0], f64[0]) < 1e-5 and rel(f64r[0], f64[0]) == 0.0
    e32, e_pos = rel(f32[1], f64[1]), rel(f64r[1], f64[1])
    assert e32 > 1e-4 and e_pos > 1e-4 and 0.2 < e_pos / e32 < 5.0, (e32, e_pos)

"""The `CryoSPH` Python module (the drop-in for the reference's module,
bindings.cpp:133-499) on the gfx950 build.

What is checked here, through the module only:
* a solver driven through the module is the oracle's solver, bit for bit, after
  start-up, relax and steps of a two-phase closed tank with phase change;
* the module hands out copies, with the reference's array shapes and dtypes;
* telemetry frames, VTK and JSON files carry exactly the downloaded state;
* bad inputs raise the Python exception the reference maps them to;
* physics read off the module's diagnostics and probes: closed-tank mass and
  energy budgets, the tank load M (g - a), the hydrostatic gradient through
  probes, conduction's energy balance.
"""
import json
import math

import numpy as np
import pytest

import cabi
import CryoSPH as cs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def configs(**fields):
    """One set of SolverConfig values as the module's object and the C ABI
    struct the oracle takes (enums given as their integer values)."""
    mod = cs.SolverConfig()
    for k, v in fields.items():
        if k == "gravity":
            mod.gravity = tuple(v)
        else:
            cur = getattr(mod, k)
            setattr(mod, k, type(cur)(v) if not isinstance(cur, (bool, int, float)) else v)
    return mod, cabi.default_config(**fields)


def fills(**fields):
    mod = cs.FillSpec()
    for k, v in fields.items():
        cur = getattr(mod, k)
        if k in ("settle_direction", "bubble_center"):
            v = tuple(v)
        elif not isinstance(cur, (bool, int, float)):
            v = type(cur)(v)
        setattr(mod, k, v)
    return mod, cabi.default_fill(**fields)


# particles() key -> oracle field
STATE = {"position": "x", "velocity": "v", "normal": "nrm", "mass": "m", "density": "rho", "temperature": "T",
         "phase": "phase", "flags": "flags", "id": "id"}


def test_device_identity():
    say("test_device_identity")
    assert "gfx950" in cs.device_name() and cs.device_count() >= 1
    assert cs.cuda_backend is False and cs.double_precision is False
    s = cs.Solver(cs.SolverConfig(), cs.FluidProperties.water())
    assert "gfx950" in s.device and s.num_particles == 0 and s.step_count == 0


def test_module_solver_is_the_oracle_bit_for_bit():
    """LN2 box tank, closed, Lee 5 with 1.5 K superheat, a 2 W heater band,
    contact angle 40 deg: build_tank, relax(5 ms), 15 steps."""
    say("test_module_solver_is_the_oracle_bit_for_bit")
    fields = dict(dx=0.004, gravity=(0.0, 0.0, -0.05), velocity_scale=1.0, thermo_mode=1, lee_evaporation=5.0,
                  lee_condensation=5.0, heater_power=2.0, contact_angle_deg=40.0)
    mcfg, ccfg = configs(**fields)
    ln2 = cs.FluidProperties.liquid_nitrogen()
    mfill, cfill = fills(fill_fraction=0.45, T_liquid=ln2.T_sat + 1.5, heater_zmin=-0.02, heater_zmax=-0.01)
    mod = cs.Solver(mcfg, ln2)
    mod.build_tank(cs.BoxTank(0.04, 0.04, 0.04), mfill)
    # the oracle takes the same lattice (the generator is pinned to the
    # reference's elsewhere) in Solver::build_tank's padded domain
    fluid = cabi.hip_fluid_preset("LN2")
    ps = cabi.hip_build_tank_particles(cabi.box(0.04, 0.04, 0.04), fluid, 0.004, cfill)
    lo, hi = cabi.tank_domain(ps, (-0.02, -0.02, -0.02), (0.02, 0.02, 0.02), 0.004)
    orc = cabi.CpuSolver("orc", ccfg, fluid, 32)
    orc.set_particles(ps, lo, hi)
    for s in (mod, orc):
        s.relax(0.005)
        s.run_steps(15)
    assert mod.num_particles == orc.num_particles and mod.step_count == orc.step_count
    assert mod.time == orc.time
    got = mod.particles()
    for key, f in STATE.items():
        assert np.array_equal(got[key], orc.field(f).reshape(got[key].shape)), key
    assert np.array_equal(mod.pressure(), orc.field("p"))
    assert np.array_equal(mod.curvature(), orc.field("kappa"))
    d, do = mod.diagnostics(), orc.diagnostics()
    assert d.n_total == do.n_total and d.n_vapor == do.n_vapor and d.splits == do.splits
    assert d.dt == do.dt and mod.last_dt == d.dt
    for k in ("mass_total", "enthalpy", "evaporated_mass", "condensed_mass", "p_ullage", "T_sat", "heat_input"):
        assert getattr(d, k) == pytest.approx(getattr(do, k), rel=1e-9, abs=1e-300), k
    assert mod.sound_speed == pytest.approx(orc.params().c0) and mod.smoothing_length == pytest.approx(0.0052)


def test_particle_arrays_are_copies_with_reference_dtypes():
    say("test_particle_arrays_are_copies_with_reference_dtypes")
    mcfg, _ = configs(dx=0.005, gravity=(0, 0, 0), thermal=False, phase_change=False, surface_tension=0)
    s = cs.Solver(mcfg, cs.FluidProperties.water())
    fill, _ = fills(mode=2)
    s.build_tank(cs.SphereTank(0.03), fill)
    n = s.num_particles
    p = s.particles()
    shapes = {"position": (n, 3), "velocity": (n, 3), "normal": (n, 3), "mass": (n,), "density": (n,),
              "temperature": (n,), "phase": (n,), "flags": (n,), "id": (n,)}
    dtypes = {"phase": np.uint8, "flags": np.uint8, "id": np.uint32}
    for k, shape in shapes.items():
        assert p[k].shape == shape and p[k].dtype == dtypes.get(k, np.float32), k
    assert sorted(np.unique(p["phase"]).tolist()) == [0, 2]
    assert np.array_equal(np.sort(p["id"]), np.arange(n, dtype=np.uint32))
    p["position"][:] = 1e9
    p["mass"][:] = -1
    again = s.particles()
    assert np.abs(again["position"]).max() < 1.0 and again["mass"].min() > 0
    assert s.pressure().shape == (n,) and s.curvature().shape == (n,)


def test_module_launch_tuning_keeps_the_results():
    """The module's extensions for launch shapes (launch_variant, tile_shape,
    autotune, autotune_tiles) change no particle: a tuned solver and an untuned
    one stay bit for bit equal."""
    say("test_module_launch_tuning_keeps_the_results")
    mcfg, _ = configs(dx=0.005, gravity=(0, 0, -0.3), thermal=False, phase_change=False, surface_tension=2)
    fill, _ = fills(include_vapor=False, fill_fraction=0.6)
    a, b = (cs.Solver(mcfg, cs.FluidProperties.liquid_nitrogen()) for _ in range(2))
    for s in (a, b):
        s.build_tank(cs.SphereTank(0.03), fill)
        s.run_steps(2)
    assert b.tile_shape == (3, 3, 3)
    b.tile_shape = (4, 2, 3)
    assert b.tile_shape == (4, 2, 3)
    with pytest.raises(ValueError):
        b.tile_shape = (3, 4, 3)
    shape, ms = b.autotune_tiles(1)
    assert len(ms) == 6 and min(ms) > 0 and b.tile_shape == tuple(shape)
    shape, ms = b.autotune_tiles(1, shapes=[(1, 1, 1), (2, 2, 2)])
    assert len(ms) == 2 and tuple(shape) in ((1, 1, 1), (2, 2, 2)) and b.tile_shape == tuple(shape)
    v, ms4 = b.autotune(1)
    assert b.launch_variant == v and len(ms4) == 4
    for s in (a, b):
        s.run_steps(3)
    pa, pb = a.particles(), b.particles()
    for k in pa:
        assert np.array_equal(pa[k].view(np.uint8), pb[k].view(np.uint8)), k


def test_module_device_counts_and_step_graphs_keep_the_results():
    """The module's set_device_counts and set_step_graphs: a closed tank with
    phase change and splitting, run in calls of several steps with the count
    on the device and the steps replayed as graphs, stays bit for bit the
    solver that reads the count back every step."""
    say("test_module_device_counts_and_step_graphs_keep_the_results")
    ln2 = cs.FluidProperties.liquid_nitrogen()
    mcfg, _ = configs(dx=0.004, gravity=(0.0, 0.0, -0.5), velocity_scale=1.0, thermo_mode=1, lee_evaporation=20.0,
                      lee_condensation=20.0, split_ratio=1.05, merge_ratio=0.9)
    fill, _ = fills(T_liquid=ln2.T_sat + 3.0)
    a, b = (cs.Solver(mcfg, ln2) for _ in range(2))
    for s in (a, b):
        s.build_tank(cs.BoxTank(0.04, 0.04, 0.04), fill)
    b.set_device_counts(True)
    b.set_step_graphs(True)
    with pytest.raises(ValueError):
        b.set_device_counts(True, -1)
    for n in (30, 15, 10):
        for s in (a, b):
            s.run_steps(n)
        pa, pb = a.particles(), b.particles()
        for k in pa:
            assert np.array_equal(pa[k].view(np.uint8), pb[k].view(np.uint8)), (n, k)
    d = a.diagnostics()
    say("  splits", d.splits, "merges", d.merges, "restarts", b.device_count_restarts)
    assert d.splits > 0 and d.merges > 0 and b.device_count_restarts >= 0


def test_closed_tank_mass_and_energy_budgets():
    """Closed LN2 capsule, superheated liquid, Lee 10, 120 steps: mass to
    2.5e-6, enthalpy change = heater input within 1 % of the latent heat
    moved, the ullage pressure rises with the evaporated mass."""
    say("test_closed_tank_mass_and_energy_budgets")
    ln2 = cs.FluidProperties.liquid_nitrogen()
    mcfg, _ = configs(dx=0.004, gravity=(0.0, 0.0, -0.01), velocity_scale=1.0, thermo_mode=1, heater_power=3.0)
    fill, _ = fills(fill_fraction=0.6, T_liquid=ln2.T_sat + 2.5)
    s = cs.Solver(mcfg, ln2)
    s.build_tank(cs.CapsuleTank(0.025, 0.03), fill)
    d0 = s.diagnostics()
    s.run_steps(120)
    d1 = s.diagnostics()
    assert d1.step == 120 and d1.time > d0.time
    assert abs(d1.mass_total - d0.mass_total) <= 2.5e-6 * d0.mass_total
    assert d1.evaporated_mass > 0 and d1.p_ullage > d0.p_ullage and d1.T_sat > d0.T_sat
    latent = (d1.evaporated_mass + d1.condensed_mass) * ln2.h_fg
    roundoff = 1e-6 * d0.mass_total * ln2.cp_liquid * ln2.T_sat
    assert abs((d1.enthalpy - d0.enthalpy) - d1.heat_input) < 0.01 * latent + roundoff
    assert d1.heat_input == pytest.approx(3.0 * (d1.time - d0.time), rel=1e-5)
    assert set(d1.as_dict()) >= {"mass_total", "enthalpy", "p_ullage", "slosh_force", "slosh_moment"}


def small_water_box(n_side=7, dx=0.005):
    mcfg, _ = configs(dx=dx, gravity=(0.0, 0.0, 0.0), velocity_scale=1.0, thermal=False, phase_change=False,
                      surface_tension=0)
    water = cs.FluidProperties.water()
    s = cs.Solver(mcfg, water)
    g = (np.arange(n_side) - n_side / 2 + 0.5) * dx
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    rng = np.random.default_rng(3)
    vel = rng.normal(0.0, 0.2, (n, 3))
    rho = np.full(n, water.rho_liquid)
    s.set_particles(pos, rho * dx ** 3, rho, np.full(n, 293.0), np.zeros(n, np.uint8), domain_lo=(-0.05,) * 3,
                    domain_hi=(0.05,) * 3, velocity=vel)
    return s


@pytest.mark.parametrize("mode", ["AsyncCopy", "ZeroCopy"])
def test_telemetry_frames_are_the_downloaded_state(mode):
    say("test_telemetry_frames_are_the_downloaded_state", mode)
    s = small_water_box()
    s.configure_telemetry(getattr(cs.TelemetryMode, mode))
    assert s.telemetry_mode == getattr(cs.TelemetryMode, mode)
    assert s.fetch_telemetry() is None
    for rounds in range(2):
        s.run_steps(4)
        assert s.request_telemetry() and not s.request_telemetry()
        fr = s.fetch_telemetry()
        p = s.particles()
        assert fr["step"] == s.step_count == 4 * (rounds + 1) and fr["time"] == s.time
        assert np.array_equal(fr["position"], p["position"])
        assert np.array_equal(fr["temperature"], p["temperature"])
        assert np.array_equal(fr["pressure"], s.pressure())
        v = p["velocity"]
        assert np.array_equal(fr["speed"], np.sqrt((v * v).sum(1, dtype=np.float32), dtype=np.float32))
        assert np.array_equal(fr["phase"], p["phase"])
        assert s.fetch_telemetry() is None and s.telemetry_ready()


def test_vtk_and_json_files_are_the_downloaded_state(tmp_path):
    say("test_vtk_and_json_files_are_the_downloaded_state")
    ln2 = cs.FluidProperties.liquid_nitrogen()
    mcfg, _ = configs(dx=0.005, gravity=(0.0, 0.0, -0.1), velocity_scale=1.0)
    s = cs.Solver(mcfg, ln2)
    s.build_tank(cs.CylinderTank(0.02, 0.04))
    s.run_steps(5)
    p, pres = s.particles(), s.pressure()
    n = s.num_particles
    s.write_vtk(str(tmp_path / "f.vtk"))
    raw = (tmp_path / "f.vtk").read_bytes()
    head = f"# vtk DataFile Version 3.0\nCryoSPH-CUDA particles\nBINARY\nDATASET POLYDATA\nPOINTS {n} float\n"
    assert raw.startswith(head.encode())
    pts = np.frombuffer(raw, ">f4", 3 * n, len(head)).reshape(n, 3)
    assert np.array_equal(pts, p["position"])
    at = raw.index(b"SCALARS pressure float 1\nLOOKUP_TABLE default\n") + 46
    assert np.array_equal(np.frombuffer(raw, ">f4", n, at), pres)
    at = raw.index(b"VECTORS velocity float\n") + 23
    assert np.array_equal(np.frombuffer(raw, ">f4", 3 * n, at).reshape(n, 3), p["velocity"])
    for walls in (False, True):
        s.write_json(str(tmp_path / "f.json"), include_walls=walls)
        fr = json.loads((tmp_path / "f.json").read_text())
        keep = np.ones(n, bool) if walls else p["phase"] != 2
        assert fr["n"] == int(keep.sum()) and fr["step"] == 5 and fr["time"] == pytest.approx(s.time, rel=1e-8)
        assert fr["phase"] == p["phase"][keep].tolist()
        np.testing.assert_allclose(fr["temperature"], p["temperature"][keep], rtol=1e-5)
        np.testing.assert_allclose(np.reshape(fr["position"], (-1, 3)), p["position"][keep], rtol=1e-5, atol=1e-9)


def test_bad_inputs_raise_reference_exceptions():
    say("test_bad_inputs_raise_reference_exceptions")
    ln2 = cs.FluidProperties.liquid_nitrogen()
    for field, bad in (("contact_angle_deg", 181.0), ("contact_angle_deg", -0.5), ("dx", -1.0), ("dx", 0.0)):
        c = cs.SolverConfig()
        setattr(c, field, bad)
        with pytest.raises(ValueError):
            cs.Solver(c, ln2)
    s = cs.Solver(cs.SolverConfig(), ln2)
    with pytest.raises(RuntimeError, match="no particles loaded"):  # sph_solver.cu:441
        s.step()
    args = dict(domain_lo=(-1, -1, -1), domain_hi=(1, 1, 1))
    one = np.ones(3)
    with pytest.raises(ValueError):  # positions must be (N, 3)
        s.set_particles(np.zeros((3, 4)), one, one, one, np.zeros(3, np.uint8), **args)
    with pytest.raises(ValueError):  # length mismatch
        s.set_particles(np.zeros((3, 3)), np.ones(2), one, one, np.zeros(3, np.uint8), **args)
    with pytest.raises(ValueError, match="no particles"):  # sph_solver.cu:408
        s.set_particles(np.zeros((0, 3)), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0, np.uint8), **args)
    with pytest.raises(ValueError):
        s.set_contact_angle(190.0)
    with pytest.raises(ValueError):  # times must increase
        cs.ManeuverProfile([0.0, 1.0, 1.0], np.zeros((3, 3)))
    with pytest.raises(ValueError):  # one acceleration row per time
        cs.ManeuverProfile([0.0, 1.0], np.zeros((3, 3)))
    ak = cs.SolverConfig()
    ak.surface_tension = cs.SurfaceTensionModel.Akinci
    with pytest.raises(ValueError):  # Akinci with a vapour phase (sph_solver.cu:410-413)
        cs.Solver(ak, ln2).build_tank(cs.SphereTank(0.02))
    with pytest.raises(ValueError):  # fill fraction outside [0, 1] (geometry.cpp:245-246)
        f = cs.FillSpec()
        f.fill_fraction = 1.5
        cs.Solver(cs.SolverConfig(), ln2).build_tank(cs.SphereTank(0.02), f)


def water_tank(extent, dx, gravity, geometry=None, **fields):
    mcfg, _ = configs(dx=dx, gravity=gravity, velocity_scale=1.0, thermal=False, phase_change=False,
                      surface_tension=0, **fields)
    s = cs.Solver(mcfg, cs.FluidProperties.water())
    fill, _ = fills(mode=2)
    s.build_tank(geometry if geometry is not None else cs.BoxTank(*extent), fill)
    return s


def settle(s, cycles, damping=0.5):
    for _ in range(cycles):
        s.run_until(s.time + 0.02)
        s.damp_velocities(damping)
    s.run_until(s.time + 0.05)


def test_probes_read_the_hydrostatic_head():
    """A full water box with a block cut from its floor (Translated and
    Difference), settled under 1 g: two probes 6 cm apart in height read
    rho g dz within 10 %; a probe outside the tank reads NaN."""
    say("test_probes_read_the_hydrostatic_head")
    tank = cs.Translated(cs.BoxTank(0.1, 0.1, 0.12), (0.05, 0.0, 0.06))
    block = cs.Translated(cs.BoxTank(0.03, 0.03, 0.02), (0.08, 0.0, 0.01))
    s = water_tank(None, 0.01, (0.0, 0.0, -9.81), cs.Difference(tank, block), artificial_viscosity=0.1)
    settle(s, 20)
    probes = np.array([[0.03, 0.0, 0.02], [0.03, 0.0, 0.08], [0.5, 0.5, 0.5]])
    p = s.probe_pressure(probes)
    assert p.shape == (3,) and np.isnan(p[2]) and np.isfinite(p[:2]).all()
    assert (p[0] - p[1]) / (1000.0 * 9.81 * 0.06) == pytest.approx(1.0, abs=0.1)


def test_tank_load_is_mass_times_g_minus_a():
    """test_maneuver.cpp:55-76 restated: the settled water in a full box loads
    the tank with M (g - a), first under gravity, then under a frame
    acceleration alone."""
    say("test_tank_load_is_mass_times_g_minus_a")
    s = water_tank((0.1, 0.1, 0.1), 0.01, (0.0, 0.0, -9.81), artificial_viscosity=0.1)
    settle(s, 30)
    d = s.diagnostics()
    M = d.mass_liquid
    assert d.slosh_force[2] / (-M * 9.81) == pytest.approx(1.0, abs=0.03)
    assert math.hypot(d.slosh_force[0], d.slosh_force[1]) < 0.01 * M * 9.81
    s.set_gravity((0.0, 0.0, 0.0))
    s.set_frame_motion(acceleration=(2.0, 0.0, 0.0))
    settle(s, 30)
    d = s.diagnostics()
    assert d.frame_acceleration == pytest.approx((2.0, 0.0, 0.0))
    assert d.slosh_force[0] / (-M * 2.0) == pytest.approx(1.0, abs=0.03)


def test_hydrostatic_pressure_gradient():
    """test_hydrostatic.cpp:13-58 restated: dp/dz = -rho g within 5 % over the
    liquid column of a settled full box."""
    say("test_hydrostatic_pressure_gradient")
    s = water_tank((0.1, 0.1, 0.2), 0.01, (0.0, 0.0, -9.81), artificial_viscosity=0.1)
    settle(s, 30)
    parts, p = s.particles(), s.pressure()
    x = parts["position"]
    core = (parts["phase"] == 0) & (np.abs(x[:, 0]) < 0.03) & (np.abs(x[:, 1]) < 0.03)
    slope = np.polyfit(x[core, 2].astype(np.float64), p[core].astype(np.float64), 1)[0]
    assert slope / (-1000.0 * 9.81) == pytest.approx(1.0, abs=0.05)


def test_conduction_between_two_slabs_keeps_energy():
    """test_conservation.cpp:26-66 restated: a cold and a warmer LN2 slab in
    contact, conduction x 2e4, 300 steps: mass to 1e-6, enthalpy to 2e-3 of
    its scale, mean temperature unchanged, the slabs' temperatures approach."""
    say("test_conduction_between_two_slabs_keeps_energy")
    ln2 = cs.FluidProperties.liquid_nitrogen()
    mcfg, _ = configs(dx=0.004, gravity=(0.0, 0.0, 0.0), contact_angle_deg=5.0, thermo_mode=1, velocity_scale=1.0,
                      phase_change=False, conductivity_scale=2e4)
    s = cs.Solver(mcfg, ln2)
    dx = 0.004
    g = (np.arange(10) + 0.5) * dx
    lower = np.stack(np.meshgrid(g, g, g[:5], indexing="ij"), -1).reshape(-1, 3)
    upper = lower + (0.0, 0.0, 5 * dx)
    pos = np.vstack([lower, upper])
    n = len(pos)
    T = np.r_[np.full(len(lower), ln2.T_sat - 4.0), np.full(len(upper), ln2.T_sat - 2.0)]
    rho = np.full(n, ln2.rho_liquid)
    s.set_particles(pos, rho * dx ** 3, rho, T, np.zeros(n, np.uint8), domain_lo=(-0.01,) * 3, domain_hi=(0.05,) * 3)
    d0 = s.diagnostics()
    s.run_steps(300)
    d1 = s.diagnostics()
    assert d1.mass_total == pytest.approx(d0.mass_total, rel=1e-6)
    assert abs(d1.enthalpy - d0.enthalpy) < 2e-3 * d0.mass_total * ln2.cp_liquid * 4.0
    assert d1.T_liquid_mean == pytest.approx(d0.T_liquid_mean, abs=0.01)
    parts = s.particles()
    low = parts["id"] < len(lower)
    gap = parts["temperature"][~low].mean() - parts["temperature"][low].mean()
    assert 0.0 < gap < 2.0

"""The CryoSPH module as a drop-in: the reference's Python tests
(tests/python/test_bindings.py:25-166) and physics checks of its C++ suite
(test_maneuver.cpp, test_conservation.cpp, test_hydrostatic.cpp), run on the
gfx950 build.
"""
import json
import math

import numpy as np
import pytest

import CryoSPH as cs

pytestmark = pytest.mark.gpu


def small_config(**overrides):
    cfg = cs.SolverConfig()
    cfg.dx = 0.005
    cfg.gravity = (0.0, 0.0, 0.0)
    cfg.velocity_scale = 1.0
    for k, v in overrides.items():
        setattr(cfg, k, v)
    return cfg


def test_module_metadata_on_device():
    print("test_module_metadata_on_device", flush=True)
    assert cs.__version__ and isinstance(cs.cuda_backend, bool)
    assert "gfx950" in cs.device_name()
    assert cs.device_count() >= 1


def test_tank_run_conserves_mass(tmp_path):
    print("test_tank_run_conserves_mass", flush=True)
    ln2 = cs.FluidProperties.liquid_nitrogen()
    solver = cs.Solver(small_config(thermo_mode=cs.TankThermoMode.Closed, lee_evaporation=2.0), ln2)
    fill = cs.FillSpec()
    fill.fill_fraction = 0.5
    fill.T_liquid = ln2.T_sat + 2.0
    solver.build_tank(cs.BoxTank(0.04, 0.04, 0.04), fill)
    solver.relax(0.01)
    d0 = solver.diagnostics()
    solver.run_steps(100)
    d1 = solver.diagnostics()
    assert d1.step == 100
    assert d1.mass_total == pytest.approx(d0.mass_total, rel=1e-5)
    assert d1.evaporated_mass > 0
    assert d1.p_ullage > d0.p_ullage
    parts = solver.particles()
    n = solver.num_particles
    assert parts["position"].shape == (n, 3)
    assert parts["phase"].dtype == np.uint8
    assert set(np.unique(parts["phase"])) <= {0, 1, 2}
    assert solver.pressure().shape == (n,)
    assert np.isfinite(parts["temperature"]).all()
    path = tmp_path / "frame.json"
    solver.write_json(str(path))
    frame = json.loads(path.read_text())
    assert frame["step"] == 100 and len(frame["position"]) == 3 * frame["n"]
    assert 2 not in frame["phase"]
    solver.write_vtk(str(tmp_path / "frame.vtk"))
    assert (tmp_path / "frame.vtk").stat().st_size > 0


def test_custom_particles_and_telemetry():
    print("test_custom_particles_and_telemetry", flush=True)
    ln2 = cs.FluidProperties.liquid_nitrogen()
    solver = cs.Solver(small_config(thermal=False, phase_change=False,
                                    surface_tension=cs.SurfaceTensionModel.NoTension), ln2)
    dx = 0.005
    g = (np.arange(6) + 0.5) * dx
    pos = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    n = len(pos)
    rho = np.full(n, ln2.rho_liquid)
    solver.set_particles(pos, rho * dx ** 3, rho, np.full(n, ln2.T_sat), np.zeros(n, np.uint8),
                         domain_lo=(-0.01, -0.01, -0.01), domain_hi=(0.04, 0.04, 0.04))
    assert solver.num_particles == n
    for mode in (cs.TelemetryMode.AsyncCopy, cs.TelemetryMode.ZeroCopy):
        solver.configure_telemetry(mode)
        solver.run_steps(3)
        assert solver.request_telemetry()
        assert not solver.request_telemetry()
        frame = solver.fetch_telemetry()
        assert frame["position"].shape == (n, 3)
        assert frame["step"] == solver.step_count
        assert np.array_equal(frame["position"], solver.particles()["position"])
        assert solver.fetch_telemetry() is None


def test_bad_inputs_raise():
    print("test_bad_inputs_raise", flush=True)
    ln2 = cs.FluidProperties.liquid_nitrogen()
    with pytest.raises(ValueError):
        cs.Solver(small_config(contact_angle_deg=200.0), ln2)
    solver = cs.Solver(small_config(), ln2)
    with pytest.raises(ValueError):
        solver.set_particles(np.zeros((4, 2)), np.ones(4), np.ones(4), np.ones(4), np.zeros(4, np.uint8),
                             domain_lo=(0, 0, 0), domain_hi=(1, 1, 1))
    with pytest.raises(ValueError, match="no particles"):  # empty input (sph_solver.cu:408)
        solver.set_particles(np.zeros((0, 3)), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0, np.uint8),
                             domain_lo=(0, 0, 0), domain_hi=(1, 1, 1))
    with pytest.raises(RuntimeError, match="no particles loaded"):  # stepping an empty solver (sph_solver.cu:441)
        cs.Solver(small_config(), ln2).run_steps(1)
    with pytest.raises(ValueError):
        solver.set_contact_angle(-5.0)
    akinci = cs.Solver(small_config(surface_tension=cs.SurfaceTensionModel.Akinci), ln2)
    with pytest.raises(ValueError):
        akinci.build_tank(cs.BoxTank(0.04, 0.04, 0.04))  # vapour + Akinci (sph_solver.cu:410-413)


def test_csg_geometry_and_probes():
    print("test_csg_geometry_and_probes", flush=True)
    tank = cs.Translated(cs.BoxTank(0.1, 0.1, 0.1), (0.05, 0.0, 0.05))
    block = cs.Translated(cs.BoxTank(0.02, 0.02, 0.02), (0.05, 0.0, 0.01))
    geo = cs.Difference(tank, block)
    water = cs.FluidProperties.water()
    cfg = small_config(surface_tension=cs.SurfaceTensionModel.NoTension, thermal=False, phase_change=False,
                       gravity=(0.0, 0.0, -9.81))
    cfg.dx = 0.01
    solver = cs.Solver(cfg, water)
    fill = cs.FillSpec()
    fill.mode = cs.FillMode.Full
    solver.build_tank(geo, fill)
    solver.run_steps(5)
    p = solver.probe_pressure(np.array([[0.02, 0.0, 0.02], [0.02, 0.0, 0.08], [0.5, 0.5, 0.5]]))
    assert p.shape == (3,)
    assert np.isfinite(p[0]) and np.isfinite(p[1])
    assert np.isnan(p[2])
    assert p[0] > p[1]


def test_maneuver_profile_and_slosh_loads():
    print("test_maneuver_profile_and_slosh_loads", flush=True)
    water = cs.FluidProperties.water()
    cfg = small_config(surface_tension=cs.SurfaceTensionModel.NoTension, thermal=False, phase_change=False)
    cfg.dx = 0.01
    solver = cs.Solver(cfg, water)
    fill = cs.FillSpec()
    fill.mode = cs.FillMode.Full
    solver.build_tank(cs.BoxTank(0.06, 0.06, 0.06), fill)
    solver.set_frame_motion(acceleration=(1.0, 0.0, 0.0))
    solver.run_steps(20)
    d = solver.diagnostics()
    assert d.frame_acceleration == pytest.approx((1.0, 0.0, 0.0))
    assert d.slosh_force[0] < 0
    assert set(d.as_dict()) >= {"slosh_force", "slosh_moment"}


def settle(solver, cycles):
    for _ in range(cycles):
        solver.run_until(solver.time + 0.02)
        solver.damp_velocities(0.5)
    solver.run_until(solver.time + 0.05)


def test_slosh_force_equals_mass_times_g_minus_a():
    """test_maneuver.cpp:55-76: a full settled water box loads the tank with M (g - a)."""
    print("test_slosh_force_equals_mass_times_g_minus_a", flush=True)
    water = cs.FluidProperties.water()
    cfg = small_config(surface_tension=cs.SurfaceTensionModel.NoTension, thermal=False, phase_change=False,
                       artificial_viscosity=0.1)
    cfg.dx = 0.01
    solver = cs.Solver(cfg, water)
    fill = cs.FillSpec()
    fill.mode = cs.FillMode.Full
    solver.build_tank(cs.BoxTank(0.1, 0.1, 0.1), fill)
    solver.set_gravity((0.0, 0.0, -9.81))
    settle(solver, 30)
    d = solver.diagnostics()
    M = d.mass_liquid
    assert d.slosh_force[2] / (-M * 9.81) == pytest.approx(1.0, abs=0.03)
    assert abs(d.slosh_force[0]) < 0.01 * M * 9.81
    solver.set_gravity((0.0, 0.0, 0.0))
    solver.set_frame_motion(acceleration=(2.0, 0.0, 0.0))
    settle(solver, 30)
    d = solver.diagnostics()
    assert d.slosh_force[0] / (-M * 2.0) == pytest.approx(1.0, abs=0.03)


def test_hydrostatic_pressure_gradient():
    """test_hydrostatic.cpp:13-58: dp/dz = -rho g within 5 %."""
    print("test_hydrostatic_pressure_gradient", flush=True)
    water = cs.FluidProperties.water()
    cfg = small_config(surface_tension=cs.SurfaceTensionModel.NoTension, thermal=False, phase_change=False,
                       gravity=(0.0, 0.0, -9.81), artificial_viscosity=0.1)
    cfg.dx = 0.01
    solver = cs.Solver(cfg, water)
    fill = cs.FillSpec()
    fill.mode = cs.FillMode.Full
    solver.build_tank(cs.BoxTank(0.1, 0.1, 0.2), fill)
    settle(solver, 30)
    parts = solver.particles()
    p = solver.pressure()
    liq = (parts["phase"] == 0) & (np.abs(parts["position"][:, 0]) < 0.03) & (np.abs(parts["position"][:, 1]) < 0.03)
    z = parts["position"][liq, 2].astype(np.float64)
    slope = np.polyfit(z, p[liq].astype(np.float64), 1)[0]
    assert slope / (-water.rho_liquid * 9.81) == pytest.approx(1.0, abs=0.05)


def test_heat_conduction_conserves_energy():
    """test_conservation.cpp:26-66 (mass 1e-6, enthalpy 2e-3 of scale, no net drift of the mean)."""
    print("test_heat_conduction_conserves_energy", flush=True)
    ln2 = cs.FluidProperties.liquid_nitrogen()
    cfg = cs.SolverConfig()
    cfg.dx = 0.004
    cfg.gravity = (0.0, 0.0, 0.0)
    cfg.contact_angle_deg = 5
    cfg.thermo_mode = cs.TankThermoMode.Closed
    cfg.velocity_scale = 1.0
    cfg.phase_change = False
    cfg.conductivity_scale = 2e4
    solver = cs.Solver(cfg, ln2)
    dx = 0.004
    g = (np.arange(10) + 0.5) * dx
    gz = (np.arange(5) + 0.5) * dx
    lo = np.array(np.meshgrid(g, g, gz, indexing="ij")).reshape(3, -1).T
    hi = lo + np.array([0, 0, 0.02])
    pos = np.vstack([lo, hi])
    n = len(pos)
    T = np.concatenate([np.full(len(lo), ln2.T_sat - 4), np.full(len(hi), ln2.T_sat - 2)])
    rho = np.full(n, ln2.rho_liquid)
    solver.set_particles(pos, rho * dx ** 3, rho, T, np.zeros(n, np.uint8), domain_lo=(-0.01, -0.01, -0.01),
                         domain_hi=(0.05, 0.05, 0.05))
    d0 = solver.diagnostics()
    solver.run_steps(300)
    d1 = solver.diagnostics()
    assert d1.mass_total == pytest.approx(d0.mass_total, rel=1e-6)
    scale = d0.mass_total * ln2.cp_liquid * 4.0
    assert abs(d1.enthalpy - d0.enthalpy) < 2e-3 * scale
    assert d1.T_liquid_mean == pytest.approx(d0.T_liquid_mean, abs=0.01)

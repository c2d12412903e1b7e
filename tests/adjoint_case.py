"""Inputs of the adjoint tests: a small closed box tank, half liquid / half
vapour LN2, hydrodynamics and optionally CSF surface tension with a contact
angle and the transport velocity, optionally conduction and Lee phase change
(superheated liquid, subcooled vapour, vented tank, vapour rebalance, heated
walls, conductivity scaled up), fixed dt, jittered positions, random velocities, and a
share of the fluid particles next to the walls moved inside the near-wall guard
distance (0.6 dx) and sent wallward, so that every branch of the wall terms
(guard repulsion, guard damping, no-slip shear, wall artificial viscosity) is
on the differentiated path.  The parameters are per-step frame accelerations.
"""
import numpy as np

import cabi
from cabi import WALL


def adjoint_case(steps=12, seed=1, dx=0.004, L=0.04, dt=2.5e-4, acc_scale=0.3, csf=False, contact_angle=30.0,
                 tvf=False, lee=None, superheat=3.0, vented=True, rebalance=1.0, conduct=2000.0, heater=500.0,
                 iso_walls=False, kernel=1):
    fluid = cabi.hip_fluid_preset("LN2")
    if csf:  # a surface tension large enough to weigh in the wall loads
        fluid.sigma = 0.2
    # strong artificial viscosity and gravity: their terms weigh in the gradient
    cfg = cabi.default_config(dx=dx, gravity=(0.0, 0.0, -5.0), thermal=0, phase_change=0,
                              surface_tension=1 if csf else 0, contact_angle_deg=contact_angle,
                              transport_velocity=1 if tvf else 0, velocity_scale=0.3, fixed_dt=dt,
                              artificial_viscosity=0.5,
                              artificial_viscosity_vapor=1.0, kernel=kernel)
    if lee is not None:  # conduction and Lee phase change with both directions of transfer
        cfg.thermal = 1
        cfg.phase_change = 1
        cfg.lee_evaporation = cfg.lee_condensation = lee
        cfg.thermo_mode = 0 if vented else 1
        cfg.vapor_mass_rebalance = rebalance
        # conduction and the heater scaled up so that they move the loads within a few steps
        cfg.conductivity_scale = conduct
        cfg.heater_power = heater
        cfg.wall_thermal_bc = 1 if iso_walls else 0
    ps = cabi.hip_build_tank_particles(cabi.box(L, L, L), fluid, dx, cabi.default_fill())
    rng = np.random.default_rng(seed)
    fl = np.flatnonzero(ps.phase != WALL)
    wl = np.flatnonzero(ps.phase == WALL)
    ps.x[fl] += rng.uniform(-0.05 * dx, 0.05 * dx, size=(len(fl), 3)).astype(np.float32)
    ps.v[fl] = rng.normal(0.0, 0.03, size=(len(fl), 3)).astype(np.float32)
    # nearest wall particle of each fluid particle (brute force over a small tank)
    xw = ps.x[wl].astype(np.float64)
    near = []
    for i in fl:
        d = np.linalg.norm(xw - ps.x[i].astype(np.float64), axis=1)
        k = int(np.argmin(d))
        if d[k] < 1.1 * dx:
            near.append((i, k, d[k]))
    for i, k, d in near[::5]:
        u = (xw[k] - ps.x[i]) / d
        ps.x[i] += (u * 0.45 * dx).astype(np.float32)
        ps.v[i] = (u * 0.02).astype(np.float32)
    if lee is not None:  # superheated liquid, subcooled vapour
        liq = ps.phase[fl] == 0
        ps.T[fl] = np.where(liq, fluid.T_sat + rng.uniform(0.0, superheat, len(fl)),
                            fluid.T_sat - rng.uniform(0.0, 0.5 * superheat, len(fl))).astype(np.float32)
    lo = ps.x.min(0) - np.float32(4 * dx)
    hi = ps.x.max(0) + np.float32(4 * dx)
    acc = rng.normal(0.0, acc_scale, size=(steps, 3)).astype(np.float32)
    return fluid, cfg, ps, lo, hi, acc


def loss_of(make_solver, acc):
    """Forward-only loss of a fresh solver on the given accelerations."""
    return make_solver().adjoint_slosh(acc, grad=False)[0]

"""The FP64 CPU adjoint (oracle/cryo_oracle.cpp, Solver::adjoint_slosh) against
central differences of the FP64 forward step, every parameter of the case, with
and without CSF surface tension (then the contact angle is a parameter too).

It pins the hand-written per-pair derivatives (csrc/adjoint_pair.hpp) that the
GPU adjoint uses: dropping any one term of them (the guard repulsion or
damping, no-slip shear, wall or fluid artificial viscosity, the Shepard wall
pressure's gravity term; with CSF the curvature denominator, the neighbour
normals, the wall-normal sum, the tangent projection of the rotation, the
blend weight, the wall colour) moves the acceleration gradient by 3e-6 to 5e-3
of its largest component or the contact-angle gradient by 6e-5 or more, while
the central differences of this case agree with the correct adjoint to 2e-7
and 1e-8.
"""
import numpy as np
import pytest

from adjoint_case import adjoint_case
from cabi import CpuSolver

STEP = 1e-3  # absolute FD step on each acceleration component (m/s^2)


def oracle(cfg, fluid, ps, lo, hi):
    s = CpuSolver("orc", cfg, fluid, 64)
    s.set_particles(ps, lo, hi)
    return s


def test_fp64_adjoint_matches_central_differences():
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=12)
    loss, grad, load, _ = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    assert loss > 0 and grad.shape == acc.shape
    # the forward pass alone gives the same loss and loads
    loss_f, none, load_f, _ = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)
    assert none is None and loss_f == loss and np.array_equal(load_f, load)
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])  # the step the float32 inputs hold
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    assert err <= 1e-6, err


def test_loss_is_the_integrated_wall_load():
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=5)
    s = oracle(cfg, fluid, ps, lo, hi)
    loss, _, load, _ = s.adjoint_slosh(acc, grad=False)
    assert loss == pytest.approx(cfg.fixed_dt * float((load[:, :2] ** 2).sum()), rel=1e-15)
    # the negated slosh force of the diagnostics (reported as float) is the last step's load
    d = s.diagnostics()
    assert np.allclose(-np.array(d.slosh_force.tup()), load[-1], rtol=1e-6, atol=0)
    assert s.step_count == 5


def test_adjoint_rejects_undifferentiated_physics():
    for csf, field, value in ((False, "surface_tension", 2), (False, "sps_turbulence", 1), (False, "fixed_dt", 0.0),
                              (False, "particle_shifting", 1)):
        fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=2, csf=csf)
        setattr(cfg, field, value)
        with pytest.raises(ValueError):
            oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)


def test_fp64_adjoint_with_csf_matches_central_differences():
    """CSF surface tension with the contact-angle rotation: every acceleration
    component and the contact angle (set before start-up, so the start-up
    rates' share counts)."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=12, csf=True, contact_angle=30.0)
    loss, grad, _, dtheta = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    assert np.abs(grad - fd).max() / np.abs(fd).max() <= 1e-6
    d = 0.01  # degrees
    out = []
    for th in (30.0 + d, 30.0 - d):
        cfg.contact_angle_deg = th
        out.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
    cfg.contact_angle_deg = 30.0
    fd_theta = (out[0] - out[1]) / (np.float32(30.0 + d) - np.float32(30.0 - d))
    assert dtheta != 0.0
    assert abs(dtheta - fd_theta) <= 1e-5 * abs(fd_theta), (dtheta, fd_theta)


def test_fp64_adjoint_with_transport_velocity_matches_central_differences():
    """The transport velocity on (interface band, with CSF): its acceleration is
    read by the next step's drift, so its cotangent is part of the state.
    Dropping that cotangent, or only its wall-pressure share, moves the
    gradient by 1e-2 of its largest component."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=12, csf=True, tvf=True)
    loss, grad, _, _ = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    assert np.abs(grad - fd).max() / np.abs(fd).max() <= 1e-6


@pytest.mark.parametrize("lee", [None, 50.0])
def test_fp64_adjoint_with_transport_stress_matches_central_differences(lee):
    """The transport-velocity stress of momentum (fluid and wall pairs read
    both velocities and both transport velocities, vt = v_half + ab dt/2), the
    transport velocity active in the bulk (not only near the interface), with
    CSF, without and with phase change."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=10, csf=True, tvf=True, lee=lee)
    cfg.transport_stress = 1
    cfg.transport_interface_only = 0
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    grad, dtheta = out[1], out[3]
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    print("transport stress: acc gradient against central differences", err)
    # measured: 3.7e-8 (hydrodynamics and CSF), 1.2e-10 (with phase change)
    assert err <= (3e-7 if lee is None else 2e-9)
    # the contact-angle derivative: the loss has kinks near this angle (the
    # viscosity and guard switches), so the central difference approaches the
    # adjoint as its step shrinks (with phase change 2.4e-5, 2.0e-5, 1.3e-5 at
    # 0.01, 0.0025, 0.00125 degrees); measured at 0.0025 degrees
    d, th0 = 0.0025, float(cfg.contact_angle_deg)
    losses = []
    for th in (th0 + d, th0 - d):
        cfg.contact_angle_deg = th
        losses.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
    cfg.contact_angle_deg = th0
    fd_theta = (losses[0] - losses[1]) / (float(np.float32(th0 + d)) - float(np.float32(th0 - d)))
    assert abs(dtheta - fd_theta) <= 5e-5 * abs(fd_theta), (dtheta, fd_theta)


@pytest.mark.parametrize("lee", [None, 50.0])
def test_fp64_adjoint_with_balanced_csf_matches_central_differences(lee):
    """The balanced CSF force (a += G sigma kappa / m, G a colour-weighted
    neighbour sum): acceleration and contact-angle gradients against central
    differences of the FP64 forward, without and with conduction and phase
    change (masses and densities then carry cotangents of their own)."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=10, csf=True, lee=lee)
    cfg.csf_balanced = 1
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    grad, dtheta = out[1], out[3]
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    # measured: 6.9e-8 (hydrodynamics only), 1.4e-10 (with phase change)
    assert np.abs(grad - fd).max() / np.abs(fd).max() <= (3e-7 if lee is None else 2e-9)
    d, th0 = 0.01, float(cfg.contact_angle_deg)
    losses = []
    for th in (th0 + d, th0 - d):
        cfg.contact_angle_deg = th
        losses.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
    cfg.contact_angle_deg = th0
    fd_theta = (losses[0] - losses[1]) / (np.float32(th0 + d) - np.float32(th0 - d))
    assert dtheta != 0.0
    assert abs(dtheta - fd_theta) <= 1e-5 * abs(fd_theta), (dtheta, fd_theta)
    if lee is not None:  # the Lee derivatives run through the masses (a += G sigma kappa / m)
        for which, name in enumerate(("lee_evaporation", "lee_condensation")):
            vals = []
            for sgn in (1, -1):
                setattr(cfg, name, lee * (1 + sgn * 1e-3))
                vals.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
            setattr(cfg, name, lee)
            fd_lee = (vals[0] - vals[1]) / (2 * lee * 1e-3)
            assert out.lee[which] != 0 and abs(out.lee[which] / fd_lee - 1) <= 2e-4, (name, out.lee[which], fd_lee)


@pytest.mark.parametrize("lee", [None, 50.0])
def test_fp64_adjoint_with_smoothed_normals_matches_central_differences(lee):
    """CSF with smoothed normals (each fluid particle's normal the normalised
    kernel-weighted sum of its live neighbours' unit colour gradients, itself
    included): acceleration, contact-angle and Lee gradients against central
    differences of the FP64 forward."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=10, csf=True, lee=lee)
    cfg.csf_smooth_normals = 1
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    grad, dtheta = out[1], out[3]
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    print("acc", err)
    assert err <= (3e-7 if lee is None else 2e-9), err
    # the smoothed loss bends more in theta: central differences with a step of
    # 0.04, 0.01 and 0.0025 degrees are 1e-3, 5e-5 and 1.5e-7 from the adjoint
    d, th0 = 0.0025, float(cfg.contact_angle_deg)
    losses = []
    for th in (th0 + d, th0 - d):
        cfg.contact_angle_deg = th
        losses.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
    cfg.contact_angle_deg = th0
    fd_theta = (losses[0] - losses[1]) / (np.float32(th0 + d) - np.float32(th0 - d))
    print("theta", dtheta, fd_theta, abs(dtheta / fd_theta - 1))
    assert dtheta != 0.0
    assert abs(dtheta - fd_theta) <= 1e-5 * abs(fd_theta), (dtheta, fd_theta)
    if lee is not None:
        for which, name in enumerate(("lee_evaporation", "lee_condensation")):
            vals = []
            for sgn in (1, -1):
                setattr(cfg, name, lee * (1 + sgn * 1e-3))
                vals.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
            setattr(cfg, name, lee)
            fd_lee = (vals[0] - vals[1]) / (2 * lee * 1e-3)
            print(name, out.lee[which], fd_lee)
            assert out.lee[which] != 0 and abs(out.lee[which] / fd_lee - 1) <= 2e-4, (name, out.lee[which], fd_lee)


@pytest.mark.parametrize("wetting,smooth", [(1, 0), (2, 0), (2, 1)])
def test_fp64_adjoint_with_the_contact_line_force_matches_central_differences(wetting, smooth):
    """The contact-line force of fluid particles near a wall (wetting model 1;
    2: with the normal rotation too), a += t f with t the wall-tangent part of
    the colour gradient's direction and f proportional to (cos theta - cos of
    the current angle), the gradient's size and the wall-indicator gradient:
    acceleration and contact-angle gradients against central differences of the
    FP64 forward, with conduction and phase change (and with smoothed normals)."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=10, csf=True, lee=50.0)
    cfg.wetting_model = wetting
    cfg.csf_smooth_normals = smooth
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    grad, dtheta = out[1], out[3]
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    print("acc", err)
    assert err <= 2e-9, err
    d, th0 = 0.0025, float(cfg.contact_angle_deg)
    losses = []
    for th in (th0 + d, th0 - d):
        cfg.contact_angle_deg = th
        losses.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
    cfg.contact_angle_deg = th0
    fd_theta = (losses[0] - losses[1]) / (np.float32(th0 + d) - np.float32(th0 - d))
    print("theta", dtheta, fd_theta, abs(dtheta / fd_theta - 1))
    assert dtheta != 0.0
    assert abs(dtheta - fd_theta) <= 1e-5 * abs(fd_theta), (dtheta, fd_theta)
    # Lee: with both wetting terms the loss bends more; a relative step of 1e-3
    # is 2.8e-4 from the adjoint, 3e-4 is 3e-5 and 3e-5 is 1.1e-5
    rel = 3e-4
    for which, name in enumerate(("lee_evaporation", "lee_condensation")):
        vals = []
        for sgn in (1, -1):
            setattr(cfg, name, 50.0 * (1 + sgn * rel))
            vals.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
        setattr(cfg, name, 50.0)
        fd_lee = (vals[0] - vals[1]) / (float(np.float32(50.0 * (1 + rel))) - float(np.float32(50.0 * (1 - rel))))
        print(name, out.lee[which], fd_lee)
        assert out.lee[which] != 0 and abs(out.lee[which] / fd_lee - 1) <= 2e-4, (name, out.lee[which], fd_lee)


def test_fp64_adjoint_with_the_cubic_spline_kernel_matches_central_differences():
    """The cubic-spline kernel (both of its branches) under every pair term of
    the case: CSF with the contact angle, the transport velocity, conduction and
    Lee phase change; acceleration and contact-angle gradients against central
    differences of the FP64 forward."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=10, csf=True, tvf=True, lee=50.0, kernel=0)
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    loss, grad, dtheta = out[0], out[1], out[3]
    assert loss > 0
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    assert np.abs(grad - fd).max() / np.abs(fd).max() <= 1e-6
    d, th0 = 0.01, float(cfg.contact_angle_deg)
    losses = []
    for th in (th0 + d, th0 - d):
        cfg.contact_angle_deg = th
        losses.append(oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
    cfg.contact_angle_deg = th0
    fd_theta = (losses[0] - losses[1]) / (np.float32(th0 + d) - np.float32(th0 - d))
    assert dtheta != 0.0
    assert abs(dtheta - fd_theta) <= 1e-5 * abs(fd_theta), (dtheta, fd_theta)


@pytest.mark.parametrize("iso_walls,vented", [(False, True), (True, True), (False, False)])
def test_fp64_adjoint_with_heat_and_phase_change_matches_central_differences(iso_walls, vented):
    """Conduction (scaled up), the heater, Lee evaporation and condensation,
    receiver normalisation with the vapour density cap, apply, vapour
    rebalance and venting on: masses and temperatures are then part of the
    state.  Both the acceleration gradient and the two Lee-coefficient
    derivatives are checked.  Dropping one derivative term moves the
    acceleration gradient by 3e-9 to 2e-5 of its largest component (the
    force-over-mass term, the sound speed's temperature, conduction pairs,
    the interface fraction, rebalance pairs, the vapour EOS temperature,
    rho = m sigma of vapour, apply and normalisation volumes, the heater
    weights and their normalisation) or a Lee derivative by 9e-3 or more
    (venting).  Run with adiabatic and with isothermal walls (wall conduction
    pairs), and for a closed tank, where each step's saturation temperature
    comes from the previous step's liquid mass and vapour m T sums (dropping
    that feedback moves the Lee derivatives by 4e-2 and 5e-3)."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=12, lee=50.0, iso_walls=iso_walls, vented=vented)
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    grad = out[1]
    fd = np.zeros_like(grad)
    for k in range(acc.shape[0]):
        for c in range(3):
            a1, a2 = acc.copy(), acc.copy()
            a1[k, c] += STEP
            a2[k, c] -= STEP
            h = float(a1[k, c]) - float(a2[k, c])
            fd[k, c] = (oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a1, grad=False)[0] -
                        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(a2, grad=False)[0]) / h
    assert np.abs(grad - fd).max() / np.abs(fd).max() <= 2e-9
    for which, name in enumerate(("lee_evaporation", "lee_condensation")):
        vals = []
        for sgn in (1, -1):
            c2 = adjoint_case(steps=12, lee=50.0, iso_walls=iso_walls, vented=vented)[1]
            setattr(c2, name, 50.0 * (1 + sgn * 1e-3))
            vals.append(oracle(c2, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)[0])
        fd_lee = (vals[0] - vals[1]) / (2 * 50.0 * 1e-3)
        assert out.lee[which] != 0 and abs(out.lee[which] / fd_lee - 1) <= 2e-4


def test_phase_change_adjoint_rejects_splits():
    # every particle above the split ratio: the first step splits them
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=4, lee=50.0)
    cfg.split_ratio = 0.5
    with pytest.raises(ValueError, match="merge or split"):
        oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc)
    oracle(cfg, fluid, ps, lo, hi).adjoint_slosh(acc, grad=False)  # the forward alone runs


def test_fp64_adjoint_in_a_rotating_frame_matches_central_differences():
    """Per-step frame angular velocity and its rate (Euler, centrifugal and
    Coriolis terms about the frame origin): the acceleration, angular-velocity
    and angular-acceleration gradients against central differences (steps
    1e-4, 1e-4 and 1e-3: the rotating case's loads curve faster, and 1e-3 on
    the acceleration leaves a truncation error of 7e-7)."""
    fluid, cfg, ps, lo, hi, acc = adjoint_case(steps=8)
    rng = np.random.default_rng(3)
    w = rng.normal(0.0, 2.0, size=acc.shape).astype(np.float32)
    wd = rng.normal(0.0, 20.0, size=acc.shape).astype(np.float32)
    out = oracle(cfg, fluid, ps, lo, hi).adjoint_slosh_frame(acc, w, wd)

    def loss(a, b, c):
        return oracle(cfg, fluid, ps, lo, hi).adjoint_slosh_frame(a, b, c, grad=False)[0]

    for which, arr, g, step in ((0, acc, out[1], 1e-4), (1, w, out.omega, 1e-4), (2, wd, out.omega_dot, 1e-3)):
        fd = np.zeros_like(g)
        for k in range(arr.shape[0]):
            for c in range(3):
                a1, a2 = arr.copy(), arr.copy()
                a1[k, c] += step
                a2[k, c] -= step
                h = float(a1[k, c]) - float(a2[k, c])
                p1, p2 = [acc, w, wd], [acc, w, wd]
                p1[which], p2[which] = a1, a2
                fd[k, c] = (loss(*p1) - loss(*p2)) / h
        err = np.abs(g - fd).max() / np.abs(fd).max()
        print(which, err)
        assert err <= (1e-8 if which == 0 else 2e-7), which
    # the plain entry point rejects a rotating frame set beforehand
    s = oracle(cfg, fluid, ps, lo, hi)
    s.set_frame_motion(acc=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 1.0))
    with pytest.raises(ValueError):
        s.adjoint_slosh(acc)

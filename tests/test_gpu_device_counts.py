"""Device counts (cryo_solver_set_device_counts): a call's phase-change steps
queued back to back, with the particle count kept on the device (splits append
there, the next cell sort drops merged particles there), give the same bits as
the same steps with the count read back after each one.  Covered on a tank
that splits and merges particles by the thousand and then grows several times
over: the default bound, the bound passed on purpose (room 1: the call restores
its start state and runs step by step), step graphs replayed with device
counts, and the ensemble runner.
"""
import numpy as np
import pytest

import cabi
from cabi import HipSolver

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "vt", "m", "m0", "rho", "T", "id", "phase", "a")


def say(*a):
    print(*a, flush=True)


def churn_tank(device_counts=False, room=0, graphs=False, **extra):
    """A closed 4 cm LN2 box, liquid 3 K superheated, Lee 20 s^-1 both ways,
    merge ratio 0.9 and split ratio 1.05: quiet for about 40 steps, then
    vapour particles split and light ones merge by the thousand every ten
    steps, and from about step 60 the count grows several times over (the
    particle arrays grow, and device-count calls pass their launch bound)."""
    fluid = cabi.hip_fluid_preset("LN2")
    cfg = cabi.default_config(dx=0.004, gravity=(0, 0, -0.5), contact_angle_deg=10, velocity_scale=1.0,
                              thermo_mode=1, lee_evaporation=20.0, lee_condensation=20.0, split_ratio=1.05,
                              merge_ratio=0.9)
    for k, v in extra.items():
        setattr(cfg, k, v)
    s = HipSolver(cfg, fluid)
    s.build_tank(cabi.box(0.04, 0.04, 0.04), cabi.default_fill(T_liquid=fluid.T_sat + 3.0))
    if device_counts:
        s.set_device_counts(True, room)
    if graphs:
        s.set_step_graphs(graphs)
    return s


def same(sa, sb, where):
    assert sa.num_particles == sb.num_particles, where
    assert sa.step_count == sb.step_count and sa.time == sb.time, where
    for f in FIELDS:
        fa, fb = sa.field(f), sb.field(f)
        assert np.array_equal(fa.view(np.uint8), fb.view(np.uint8)), (where, f)
    assert bytes(sa.diagnostics()) == bytes(sb.diagnostics()), where


@pytest.mark.parametrize("room,graphs", [(0, False), (1, False), (0, True)])
def test_device_counts_give_the_same_bits(room, graphs):
    say("test_device_counts_give_the_same_bits room", room, "graphs", graphs)
    ref = churn_tank()
    dev = churn_tank(True, room, graphs)
    n0 = ref.num_particles
    done = 0
    for n in (30, 10, 10, 10, 4, 3, 3):
        ref.run_steps(n)
        dev.run_steps(n)
        done += n
        same(ref, dev, f"after {done} steps")
        d = ref.diagnostics()
        say(f"  step {done}: particles {ref.num_particles} splits {d.splits} merges {d.merges} "
            f"restarts {dev.device_count_restarts()}")
    d = ref.diagnostics()
    assert d.splits > 0 and d.merges > 0 and ref.num_particles > 2 * n0
    if room == 1:  # one particle of room: every call with two or more splits restarts
        assert dev.device_count_restarts() > 0


@pytest.mark.parametrize("extra", [dict(compensated_sums=1), dict(particle_shifting=1, shifting_strength=0.5),
                                   dict(thermo_mode=0, vent_relaxation=0.05)],
                         ids=["precision_form", "shifting", "vented"])
def test_device_counts_other_step_paths(extra):
    """The precision form (double-float positions in the saved state), particle
    shifting (its second source pass) and a vented tank (venting launched over
    the bound): device counts keep the bits through the churn."""
    say("test_device_counts_other_step_paths", extra)
    ref, dev = churn_tank(**extra), churn_tank(True, **extra)
    done = 0
    for n in (30, 15, 15):
        ref.run_steps(n)
        dev.run_steps(n)
        done += n
        same(ref, dev, f"{extra} after {done} steps")
    d = ref.diagnostics()
    say(f"  particles {ref.num_particles} splits {d.splits} merges {d.merges} vented {d.vented_mass} "
        f"restarts {dev.device_count_restarts()}")
    if "thermo_mode" in extra:  # the open tank vents its vapour instead of splitting it
        assert d.vented_mass > 0
    else:
        assert d.splits > 0 and d.merges > 0


def test_device_counts_single_steps_and_calls_in_a_row():
    """Plain step() calls between device-count calls, and one-step calls
    (which read the count back as usual), keep the bits."""
    say("test_device_counts_single_steps_and_calls_in_a_row")
    ref, dev = churn_tank(), churn_tank(True)
    for k, n in enumerate((20, 1, 20, 2, 11)):
        ref.run_steps(n)
        dev.run_steps(n)
        dev.step()
        ref.step()
        same(ref, dev, f"call {k}")
    d = ref.diagnostics()
    assert d.splits > 0 and d.merges > 0


def test_ensemble_with_device_counts_matches_the_per_tank_loop():
    say("test_ensemble_with_device_counts_matches_the_per_tank_loop")
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import bench

    tanks = sorted(bench.c4_tanks(cabi, count=12, seed=7), key=lambda t: len(t["ps"]))[:3]

    def members(dc, graphs=False):
        out = []
        for t in tanks:
            s = HipSolver(t["cfg"], t["fluid"])
            s.set_particles(t["ps"], t["lo"], t["hi"])
            s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
            if dc:
                s.set_device_counts(True)
            if graphs:
                s.set_step_graphs(2)
            out.append(s)
        out.append(churn_tank(dc, 0, 2 if graphs else False))
        return out

    ref = members(False)
    for s in ref:
        s.run_steps(56)
    for graphs in (False, True):
        ens = members(True, graphs)
        # with graphs (mode 2) every call of two or more steps runs the whole
        # ensemble as one captured graph; 15 ends on an eager step, 1 is a plain step
        cabi.hip_ensemble_run_steps(ens, 40)
        cabi.hip_ensemble_run_steps(ens, 15)
        cabi.hip_ensemble_run_steps(ens, 1)
        for k, (sa, sb) in enumerate(zip(ref, ens)):
            same(sa, sb, f"member {k} graphs {graphs}")
    d = ref[-1].diagnostics()
    assert d.splits > 0 and d.merges > 0


def test_one_graph_ensemble_without_phase_change_matches_the_per_tank_loop():
    """Members without phase change and with graphs on: the ensemble runs as
    one captured graph per two steps (an odd call ends on an eager step) and
    stays bitwise the per-tank loop."""
    say("test_one_graph_ensemble_without_phase_change_matches_the_per_tank_loop")
    import copy
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import bench

    tanks = sorted(bench.c4_tanks(cabi, count=12, seed=7), key=lambda t: len(t["ps"]))[:3]

    def members(graphs):
        out = []
        for t in tanks:
            cfg = copy.copy(t["cfg"])
            cfg.phase_change = 0
            s = HipSolver(cfg, t["fluid"])
            s.set_particles(t["ps"], t["lo"], t["hi"])
            s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
            s.set_step_graphs(2 if graphs else 0)
            out.append(s)
        return out

    ref, ens = members(False), members(True)
    for s in ref:
        s.run_steps(13)
    cabi.hip_ensemble_run_steps(ens, 6)
    cabi.hip_ensemble_run_steps(ens, 7)
    for k, (sa, sb) in enumerate(zip(ref, ens)):
        same(sa, sb, f"member {k}")


def test_device_counts_reject_negative_room():
    s = churn_tank()
    with pytest.raises(ValueError):
        s.set_device_counts(True, -1)

"""Throughput benchmark of the CryoSPH step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dx 0.007] [--cpu-steps 20] [--no-cpu]

Workload (the reference dam-break form at scale): single-phase water
column collapse in a 3.0 x 1.0 x 0.8 m box with a 0.15 x 0.4 x 0.15 m block on
the floor at x = 2.3 m; water column 1.0 x 1.0 x 0.5 m at x = 0; rho = 1000,
mu = 1e-3, c0 = 10 sqrt(2 g H), continuity density with delta-SPH 0.1, alpha 0.02,
no-slip walls, thermal / phase change / surface tension off (the reference's
own dam-break settings, validation/spheric_test2/run.py:36-51).  dx = 0.007 m
gives 2,233,076 particles (1,441,726 fluid + 791,350 wall); synthetic lattice
input, no data files.

One "step" is Solver::step (sph_solver.cu:511-533): kick-drift, cell sort,
every neighbour pass, reductions and dt.  `value` counts particle-steps per
second with walls included (as benchmarks/perf_scaling.cpp:103-104 does),
inputs resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import copy
import ctypes as C
import json
import math
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tests"))
# Hardware queues for the ensemble's per-tank streams (HIP's default is 4, so
# at most four tanks' kernels would run side by side); read when HIP starts,
# which is after this line (the first solver created).
os.environ["GPU_MAX_HW_QUEUES"] = "16"

# Algorithmic bytes per particle of the passes the dominant kernel (k_forces)
# fuses for this workload: transport velocity (bulk early exit), momentum with
# wall loads, continuity rate and the dt maxima.
FORCES_BYTES = {"fluid": 21 + 89 + 33 + 44, "wall": 13 + 37 + 5 + 4}
# Whole-step algorithmic bytes of the reference pass structure for this
# workload (single phase, continuity, no thermal / phase change / tension).
STEP_BYTES = {"fluid": 590, "wall": 310}
HBM_PEAK_GBS = 8000.0


C2_DX = 0.007  # the workload the committed counter profiles were taken on
VALU_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
SIMDS = 1024  # 256 CUs x 4 SIMDs


def profile_column(name, kernel, column):
    """Column `column` (0 = the kernel name) of `kernel`'s line in a committed
    counter summary under profiles/ (counters cannot be read inside this run;
    a kernel's byte and instruction counts per launch are fixed for a given
    workload).  None when absent."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) > column and parts[0].startswith(kernel):
            try:
                return float(parts[column].rstrip("%"))
            except ValueError:
                return None
    return None


class marker:
    """A roctx range around a region (shown by rocprofv3 --marker-trace, which
    the ensemble profile uses to find its window); a no-op when the roctx
    library is absent."""

    _lib = None
    _tried = False

    def __init__(self, name):
        self.name = name.encode()

    @classmethod
    def lib(cls):
        if not cls._tried:
            cls._tried = True
            for path in ("/opt/rocm/lib/librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so.1"):
                try:
                    cls._lib = C.CDLL(path)
                    break
                except OSError:
                    continue
        return cls._lib

    def __enter__(self):
        if self.lib() is not None:
            self.lib().roctxRangePushA(C.c_char_p(self.name))
        return self

    def __exit__(self, *exc):
        if self.lib() is not None:
            self.lib().roctxRangePop()
        return False


def c2_particles(cabi, dx):
    water = cabi.hip_fluid_preset("H2O")
    water.rho_liquid = 1000.0
    water.mu_liquid = 1.0e-3
    f32 = lambda v: float(np.float32(v))  # noqa: E731  (Translated keeps float offsets, geometry.hpp:92-103)
    nodes = cabi.geo_nodes([
        (cabi.GEO_BOX, -1, -1, (3.0, 1.0, 0.8)),
        (cabi.GEO_TRANSLATED, 0, -1, (f32(1.5), 0.0, f32(0.4))),
        (cabi.GEO_BOX, -1, -1, (0.15, 0.4, 0.15)),
        (cabi.GEO_TRANSLATED, 2, -1, (f32(2.375), 0.0, f32(0.075))),
        (cabi.GEO_DIFFERENCE, 1, 3, ()),
    ])
    ps = cabi.hip_build_tank_particles(nodes, water, dx, cabi.default_fill(mode=2))
    keep = (ps.phase == cabi.WALL) | ((ps.x[:, 0] < 1.0) & (ps.x[:, 2] < 0.5))
    ps = cabi.ParticleSet(ps.x[keep], ps.v[keep], ps.nrm[keep], ps.m[keep], ps.rho[keep], ps.T[keep], ps.phase[keep],
                          ps.flags[keep], None)
    pad = np.float32(4 * dx)
    lo = np.array([-pad, -0.5 - pad, -pad], np.float32)
    hi = np.array([3.0 + pad, 0.5 + pad, 0.8 + pad], np.float32)
    cfg = cabi.default_config(dx=dx, gravity=(0.0, 0.0, -9.81), surface_tension=0, thermal=0, phase_change=0,
                              density_mode=1, velocity_scale=math.sqrt(2 * 9.81 * 0.5), artificial_viscosity=0.02,
                              delta_sph=0.1, no_slip=1)
    return water, cfg, ps, lo, hi


def c4_tanks(cabi, count=64, seed=0):
    """The batched-ensemble workload: `count` LN2 / LOX cylinder and capsule tanks
    (r = 0.1 m; cylinder h = 0.3 m, capsule straight length 0.2 m), particle
    counts drawn log-uniform in [20k, 400k] (dx solved from the tank volume),
    fill U(0.2, 0.8), lateral acceleration U(0.01, 0.1) m/s^2, Lee coefficient
    U(0.1, 20) 1/s for evaporation and condensation, liquid superheat U(0, 5) K;
    microgravity 1e-4 g, closed tanks, thermal and phase change on."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        name = ("LN2", "LOX")[int(rng.integers(2))]
        shape = ("cylinder", "capsule")[int(rng.integers(2))]
        n_target = float(np.exp(rng.uniform(np.log(20e3), np.log(400e3))))
        fill = float(rng.uniform(0.2, 0.8))
        a_lat = float(rng.uniform(0.01, 0.1))
        lee = float(rng.uniform(0.1, 20.0))
        superheat = float(rng.uniform(0.0, 5.0))
        r = 0.1
        if shape == "cylinder":
            h = 0.3
            vol = lambda d: math.pi * (r + 3 * d) ** 2 * (h + 6 * d)  # noqa: E731  (tank + 3 wall layers)
            nodes = cabi.cylinder(r, h)
        else:
            length = 0.2
            vol = lambda d: math.pi * (r + 3 * d) ** 2 * length + 4.0 / 3.0 * math.pi * (r + 3 * d) ** 3  # noqa: E731
            nodes = cabi.capsule(r, length)
        dx = 0.01
        for _ in range(30):
            dx = (vol(dx) / n_target) ** (1.0 / 3.0)
        fluid = cabi.hip_fluid_preset(name)
        dx = float(np.float32(dx))
        ps = cabi.hip_build_tank_particles(nodes, fluid, dx, cabi.default_fill(fill_fraction=fill,
                                                                              T_liquid=fluid.T_sat + superheat))
        cfg = cabi.default_config(dx=dx, gravity=(0.0, 0.0, -9.81e-4), thermo_mode=1, lee_evaporation=lee,
                                  lee_condensation=lee)
        pad = np.float32(4 * dx)
        out.append(dict(fluid=fluid, cfg=cfg, ps=ps, lo=ps.x.min(0) - pad, hi=ps.x.max(0) + pad, a_lat=a_lat,
                        desc=f"{name} {shape} dx={dx:.4g} N={len(ps)}"))
    return out


def ensemble_measure(cabi, tanks, steps, tune_steps=3):
    """The same tanks advanced `steps` steps by the per-tank loop of Solvers (the
    default launch shape), then autotuned (cryo_solver_autotune: each member
    keeps the fastest launch variant of its neighbour passes, then, where its
    forces pass runs over LDS cell tiles, cryo_solver_autotune_tiles the fastest
    cell-tile shape; the results are the same bits either way), then by the
    loop again and by cryo_ensemble_run_steps
    (bitwise the loop's results; the GPU test checks that), wall-clock per
    ensemble step each way; and the ensemble once more with every member in
    the precision form, the form whose slosh force meets the FP64 bar."""
    solvers = []
    for t in tanks:
        s = cabi.HipSolver(t["cfg"], t["fluid"])
        s.set_particles(t["ps"], t["lo"], t["hi"])
        s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
        solvers.append(s)
    for s in solvers:  # start-up evaluation + one step each, untimed
        s.run_steps(1)
    t0 = time.perf_counter()
    for s in solvers:
        s.run_steps(steps)
    loop = time.perf_counter() - t0
    def tune_member(s):
        v = s.autotune(tune_steps)[0]
        shape = s.autotune_tiles(tune_steps)[0] if (v & 1) == 0 else s.tile_shape()
        return v, shape

    t0 = time.perf_counter()
    tuned = [tune_member(s) for s in solvers]
    picks = [v for v, _ in tuned]
    tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    with marker("per_tank_loop_tuned"):
        for s in solvers:
            s.run_steps(steps)
    loop_tuned = time.perf_counter() - t0
    t0 = time.perf_counter()
    with marker("ensemble_run_steps"):
        cabi.hip_ensemble_run_steps(solvers, steps)
    ens = time.perf_counter() - t0
    for s in solvers:  # the same with captured step graphs (captures included)
        s.set_step_graphs(True)
    t0 = time.perf_counter()
    cabi.hip_ensemble_run_steps(solvers, steps)
    ens_graphs = time.perf_counter() - t0
    for s in solvers:  # counts kept on the device: no host round trip per step
        s.set_step_graphs(False)
        s.set_device_counts(True)
    t0 = time.perf_counter()
    with marker("ensemble_device_counts"):
        cabi.hip_ensemble_run_steps(solvers, steps)
    ens_dc = time.perf_counter() - t0
    for s in solvers:  # and each member's step a replayed graph
        s.set_step_graphs(True)
    t0 = time.perf_counter()
    cabi.hip_ensemble_run_steps(solvers, steps)
    ens_dc_graphs = time.perf_counter() - t0
    for s in solvers:  # and the whole ensemble one captured graph per two steps
        s.set_step_graphs(2)
    t0 = time.perf_counter()
    cabi.hip_ensemble_run_steps(solvers, steps)
    ens_one_graph = time.perf_counter() - t0
    n = sum(s.num_particles for s in solvers)
    del solvers
    # the ensemble at the accuracy the slosh-force bar asks for: every member
    # in the precision form (compensated_sums = 1), autotuned the same way
    prec = []
    for t in tanks:
        cp = copy.copy(t["cfg"])
        cp.compensated_sums = 1
        s = cabi.HipSolver(cp, t["fluid"])
        s.set_particles(t["ps"], t["lo"], t["hi"])
        s.set_frame_motion(acc=(t["a_lat"], 0.0, 0.0))
        prec.append(s)
    for s in prec:
        s.run_steps(1)
        tune_member(s)
    t0 = time.perf_counter()
    cabi.hip_ensemble_run_steps(prec, steps)
    ens_prec = time.perf_counter() - t0
    del prec
    return {"tanks": len(tanks), "particles": n, "steps": steps,
            "per_tank_loop_ms_per_step": 1e3 * loop / steps,
            "autotune_s": tune, "launch_variants_chosen": [picks.count(v) for v in range(4)],
            "tile_shapes_chosen": {"x".join(map(str, sh)): [sh for _, sh in tuned].count(sh)
                                   for sh in sorted({sh for _, sh in tuned})},
            "per_tank_loop_tuned_ms_per_step": 1e3 * loop_tuned / steps,
            "ensemble_ms_per_step": 1e3 * ens / steps,
            "ensemble_step_graphs_ms_per_step": 1e3 * ens_graphs / steps,
            "ensemble_device_counts_ms_per_step": 1e3 * ens_dc / steps,
            "ensemble_device_counts_step_graphs_ms_per_step": 1e3 * ens_dc_graphs / steps,
            "ensemble_one_graph_ms_per_step": 1e3 * ens_one_graph / steps,
            "speedup": loop / ens, "speedup_vs_tuned_loop": loop_tuned / ens,
            "ensemble_particle_steps_per_s": n * steps / ens,
            "precision_form_ensemble_ms_per_step": 1e3 * ens_prec / steps,
            "precision_form_speedup_vs_plain_loop": loop / ens_prec}


def c5_case(cabi, steps, segments=16, seed=0, superheat=0.2):
    """The differentiable-slosh workload: LN2 cylinder tank r = 0.1 m, h = 0.25 m,
    fill 0.5, dx = 3.7 mm, microgravity 1e-4 g, CSF surface tension with a 30
    degree contact angle, conduction and Lee phase change with coefficient 5
    (vented tank, vapour rebalance), the liquid `superheat` K above saturation
    (so that Lee transfer runs; small enough that no particle splits within a
    few hundred steps), a `segments`-piece constant frame-acceleration profile
    drawn N(0, 0.03) m/s^2 per component (seeded), fixed dt = the start-up dt;
    transport velocity off."""
    fluid = cabi.hip_fluid_preset("LN2")
    dx = 0.0037
    cfg = cabi.default_config(dx=dx, gravity=(0.0, 0.0, -9.81e-4), velocity_scale=0.1, thermal=1, phase_change=1,
                              lee_evaporation=5.0, lee_condensation=5.0, thermo_mode=0,
                              surface_tension=1, contact_angle_deg=30.0, transport_velocity=0)
    ps = cabi.hip_build_tank_particles(cabi.cylinder(0.1, 0.25), fluid, dx, cabi.default_fill(fill_fraction=0.5))
    ps.T[ps.phase == 0] += np.float32(superheat)
    pad = np.float32(4 * dx)
    lo, hi = ps.x.min(0) - pad, ps.x.max(0) + pad
    probe = cabi.HipSolver(cfg, fluid)
    probe.set_particles(ps, lo, hi)
    cfg.fixed_dt = float(np.float32(probe.next_dt()))
    del probe
    seg = np.random.default_rng(seed).normal(0.0, 0.03, size=(segments, 3))
    acc = seg[(np.arange(steps) * segments) // steps].astype(np.float32)
    return fluid, cfg, ps, lo, hi, acc


def adjoint_measure(cabi, steps, every):
    """Wall-clock of the forward pass alone and of forward + reverse sweep
    (cryo_solver_adjoint_slosh with and without the gradient) on the
    differentiable-slosh workload, each on a fresh, initialised solver; the
    fastest of three interleaved runs of each."""
    fluid, cfg, ps, lo, hi, acc = c5_case(cabi, steps)
    times = {}
    for name, grad in (("forward", False), ("adjoint", True), ("forward2", False), ("adjoint2", True),
                       ("forward3", False), ("adjoint3", True)):
        s = cabi.HipSolver(cfg, fluid)
        s.set_particles(ps, lo, hi)
        s.initialize()
        s.time  # synchronises
        t0 = time.perf_counter()
        out = s.adjoint_slosh(acc, checkpoint_every=every, grad=grad)
        times[name] = time.perf_counter() - t0
        loss = out[0]
        if grad:
            gnorm, gtheta, glee = float(np.abs(out[1]).max()), out[3], out.lee
        del s
    fwd = min(times["forward"], times["forward2"], times["forward3"])
    adj = min(times["adjoint"], times["adjoint2"], times["adjoint3"])
    # the precision form (the forward whose gradient meets the FP64 bar)
    pcfg = copy.copy(cfg)
    pcfg.compensated_sums = 1
    ptimes = {}
    for name, grad in (("forward", False), ("adjoint", True), ("forward2", False), ("adjoint2", True)):
        s = cabi.HipSolver(pcfg, fluid)
        s.set_particles(ps, lo, hi)
        s.initialize()
        s.time  # synchronises
        t0 = time.perf_counter()
        pout = s.adjoint_slosh(acc, checkpoint_every=every, grad=grad)
        ptimes[name] = time.perf_counter() - t0
        if grad:
            ploss, pgmax = pout[0], float(np.abs(pout[1]).max())
        del s
    pfwd = min(ptimes["forward"], ptimes["forward2"])
    padj = min(ptimes["adjoint"], ptimes["adjoint2"])
    return {"workload": f"LN2 cylinder r=0.1 h=0.25 fill 0.5 dx=0.0037 CSF contact angle 30 deg, Lee 5, "
                        f"liquid 0.2 K superheated, vented, {len(ps)} particles, {steps} steps of "
                        f"dt={cfg.fixed_dt:.3e} s, 16-segment frame-acceleration profile; "
                        f"loss = sum dt |load_xy|^2",
            "scope": "density summation, EOS, Shepard wall BC, momentum, CSF with contact-angle rotation, "
                     "conduction, Lee sources / normalisation / apply, vapour rebalance, venting, kick-drift-kick",
            "particles": len(ps), "steps": steps, "checkpoint_every": every, "loss": loss, "grad_acc_max": gnorm,
            "grad_contact_angle": gtheta, "grad_lee_evaporation": glee[0], "grad_lee_condensation": glee[1],
            "forward_ms_per_step": 1e3 * fwd / steps, "forward_and_adjoint_ms_per_step": 1e3 * adj / steps,
            "cost_ratio": adj / fwd,
            "precision_form": {"loss": ploss, "grad_acc_max": pgmax, "forward_ms_per_step": 1e3 * pfwd / steps,
                               "forward_and_adjoint_ms_per_step": 1e3 * padj / steps, "cost_ratio": padj / pfwd,
                               "cost_ratio_vs_plain_forward": padj / fwd}}


def profile_section(lib, handle, index):
    name = C.create_string_buffer(64)
    ms = C.c_double()
    cnt = C.c_int64()
    rc = lib.cryo_solver_profile_section(handle, C.c_int32(index), name, C.c_size_t(64), C.byref(ms), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(lib.cryo_last_error().decode())
    return name.value.decode(), ms.value, cnt.value


def host_cpu_facts():
    """What the CPU baseline runs on: CPUs this process may use, physical cores
    among them (distinct (socket, core) pairs of `lscpu -p`), the model name."""
    allowed = sorted(os.sched_getaffinity(0))
    cores, model = len(allowed), "unknown"
    try:
        rows = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        pairs = {tuple(r.split(",")[1:3]) for r in rows.splitlines()
                 if r and not r.startswith("#") and int(r.split(",")[0]) in allowed}
        cores = len(pairs) or cores
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    return {"nproc": len(allowed), "physical_cores": cores, "model": model}


def cpu_baseline(dx, steps, warmup=5):
    """The reference CPU/OpenMP build (oracle/_ref; the restatement if that is
    absent) on the host cores, a bounded sample of the same workload: `warmup`
    untimed steps after the start-up evaluation, then `steps` timed steps
    (benchmarks/perf_scaling.cpp:86-104).  One OpenMP thread per physical core
    this process may use, capped by OMP_NUM_THREADS where that is set (the
    host's share of CPUs); threads bound close, one per core.  It runs in a
    child process so that the OpenMP placement applies from the runtime's
    start."""
    facts = host_cpu_facts()
    threads = facts["physical_cores"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-baseline-child", "--dx", str(dx),
           "--cpu-steps", str(steps), "--warmup", str(warmup)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline child failed:\n" + r.stdout + r.stderr)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res.update({"cores": threads, "threads": threads, "host_physical_cores": facts["physical_cores"],
                "host": facts})
    res["sample"] += (f", {threads} OpenMP threads (OMP_PROC_BIND=close, OMP_PLACES=cores) on a host with "
                      f"nproc={facts['nproc']}, {facts['physical_cores']} physical cores ({facts['model']})")
    return res


def cpu_baseline_child(dx, steps, warmup):
    import cabi

    water, cfg, ps, lo, hi = c2_particles(cabi, dx)
    kind = "reference" if cabi.REF_LIB[32].exists() else "port"
    pref = "ref" if kind == "reference" else "orc"
    s = cabi.CpuSolver(pref, cfg, water, 32)
    s.set_particles(ps, lo, hi)
    s.run_steps(warmup)  # start-up evaluation + warm-up steps
    t0 = time.perf_counter()
    s.run_steps(steps)
    sec = time.perf_counter() - t0
    n = s.num_particles
    print(json.dumps({"value": n * steps / sec, "unit": "particle-steps/s", "kind": kind,
                      "sample": f"same workload (N={n}), {steps} timed steps after {warmup} warm-up steps, "
                                f"{sec:.1f} s"}))


def copy_calibration(mib):
    """A plain device-to-device copy of known size: what FETCH_SIZE / WRITE_SIZE
    report for it converts those counters into bytes for the other kernels."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "dispatches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.007)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=10, help="0 skips the 64-tank ensemble leg")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="0 skips the adjoint leg")
    ap.add_argument("--checkpoint-every", type=int, default=20)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the 64-tank ensemble leg and print its JSON object")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (scales the HBM byte counters) and exit")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.cpu_baseline_child:
        cpu_baseline_child(a.dx, a.cpu_steps, a.warmup)
        return
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.ensemble_only:
        import cabi

        print(json.dumps(ensemble_measure(cabi, c4_tanks(cabi), max(a.ensemble_steps, 1))))
        return
    if a.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {a.gpus} is not supported (how the step would shard "
              "across GPUs is described in the design notes)", file=sys.stderr)
        sys.exit(2)

    import cabi

    water, cfg, ps, lo, hi = c2_particles(cabi, a.dx)
    n_fluid = int((ps.phase != cabi.WALL).sum())
    n_wall = int((ps.phase == cabi.WALL).sum())
    s = cabi.HipSolver(cfg, water)
    lib = s.lib
    s.set_particles(ps, lo, hi)
    s.run_steps(a.warmup)  # start-up evaluation, allocations, first sorts
    lib.cryo_solver_profile_reset(s.h)
    lib.cryo_solver_profile_enable(s.h, C.c_uint32(1 << 6))  # forces kernel only
    t0 = time.perf_counter()
    s.run_steps(a.steps)  # ends with a stream synchronisation
    sec = time.perf_counter() - t0
    lib.cryo_solver_profile_enable(s.h, C.c_uint32(0))
    name, ms, cnt = profile_section(lib, s.h, 6)
    n = s.num_particles
    ms_step = 1e3 * sec / a.steps
    forces_ms = ms / max(cnt, 1)
    forces_bytes = n_fluid * FORCES_BYTES["fluid"] + n_wall * FORCES_BYTES["wall"]
    achieved = forces_bytes / (forces_ms * 1e-3) / 1e9
    step_bytes = n_fluid * STEP_BYTES["fluid"] + n_wall * STEP_BYTES["wall"]
    # HBM roofline on algorithmic bytes (live kernel time), and the roof that
    # binds: the kernel's VALU issue share, VALU instructions per launch (from
    # the committed counter profile of this workload) x 4 cycles per wave64
    # instruction over the SIMD-cycles of the live launch time.  Counter-based
    # fields are given only for the workload the profiles were taken on.
    on_profile = abs(a.dx - C2_DX) < 1e-9
    traffic = profile_column("hbm_traffic_c2.txt", "k_forces_tiled", 1) if on_profile else None
    valu_instr = profile_column("valu_l2_c2.txt", "k_forces_tiled", 2) if on_profile else None
    valu_frac = (valu_instr * 4 / SIMDS / (forces_ms * 1e-3 * VALU_CLOCK_HZ)) if valu_instr else None
    tfrac = (traffic / (forces_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None
    if valu_frac is None:
        limiter = None
    elif valu_frac >= 0.8 or (tfrac or 0.0) >= 0.8:
        limiter = "VALU issue" if valu_frac >= (tfrac or 0.0) else "HBM traffic"
    else:
        limiter = "latency: neither the VALU nor the HBM roof is reached"
    roofline = {"bound": "valu" if valu_frac is not None and valu_frac > achieved / HBM_PEAK_GBS else "hbm",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic,
                "traffic_source": "profiles/hbm_traffic_c2.txt (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
                if traffic else None,
                "traffic_frac": tfrac,
                "valu_frac": valu_frac,
                "valu_source": "profiles/valu_l2_c2.txt (SQ_INSTS_VALU per launch) over the live kernel time"
                if valu_frac is not None else None,
                "limiter": limiter,
                "kernel": "k_forces_tiled (transport velocity + momentum + wall loads + continuity + dt "
                          "maxima; LDS-staged cell tiles) + its overflow instance",
                "kernel_ms": forces_ms, "algorithmic_bytes": forces_bytes}
    out = {
        "metric": "particle-steps/s + % HBM roofline at 1M; ensemble speedup; adjoint/forward cost",
        "value": n * a.steps / sec,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (lattice-generated tank, no data files)",
        "config": {"workload": f"water column collapse 3.0x1.0x0.8 m box + 0.15x0.4x0.15 m block, column "
                               f"1.0x1.0x0.5 m, dx={a.dx}, continuity + delta-SPH, no-slip, FP32",
                   "n_particles": n, "n_fluid": n_fluid, "n_wall": n_wall, "dx": a.dx},
        "roofline": roofline,
        "step_roofline": {"algorithmic_bytes": step_bytes,
                          "achieved": step_bytes / (ms_step * 1e-3) / 1e9,
                          "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    del s
    if a.ensemble_steps > 0:
        out["ensemble"] = ensemble_measure(cabi, c4_tanks(cabi), a.ensemble_steps)
    if a.adjoint_steps > 0:
        out["adjoint"] = adjoint_measure(cabi, a.adjoint_steps, a.checkpoint_every)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.dx, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

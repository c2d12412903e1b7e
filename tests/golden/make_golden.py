"""Writes the golden vectors under tests/golden/ with the reference's own CPU
path (oracle/_ref/libcryoref_f32.so, compiled from the reference sources by
oracle/Makefile).  Run in the build container:

    make -C oracle ref && OMP_NUM_THREADS=8 python tests/golden/make_golden.py

Each file holds inputs and the reference's outputs for them:
  hash_sort.npz   random cloud (numpy seed 7, 3000 points, every 5th dead in
                  the second set): sort permutation, n_alive, neighbour counts
                  and id sums at r = 0.08 (test_spatial_hash.cpp:25-93 scenario)
  box_tank.npz    LN2 box tank 0.04^3, dx 4 mm (3,896 particles), jittered:
                  every field after start-up, the state after 20 steps and the
                  diagnostics series
  c1_grid.npz     LN2 cylinder r 0.05 h 0.12, dx 4 mm (26,360 particles),
                  jittered: start-up sort (ids in sorted order), cell tables
                  and per-particle neighbour count / id-sum within 2h
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from cabi import (FIELDS, WALL, CpuSolver, box, cpu_build_tank_particles, cylinder, default_config,  # noqa: E402
                  default_fill, fluid_preset)
import ctypes as C  # noqa: E402


def hash_sort():
    lib = CpuSolver.load("ref", 32)
    rng = np.random.default_rng(7)
    x = rng.random((3000, 3), dtype=np.float32)
    out = {"x": x}
    for tag, dead in (("alive", False), ("dead", True)):
        ph = np.where(np.arange(3000) % 3 == 0, 1, 0).astype(np.uint8)
        if dead:
            ph[::5] = 3
        perm = np.zeros(3000, np.uint32)
        na = C.c_int()
        cnt = np.zeros(3000, np.int32)
        sums = np.zeros(3000, np.uint64)
        lo = np.zeros(3, np.float32)
        hi = np.ones(3, np.float32)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        rc = lib.ref_hash_sort(C.c_int(3000), vp(x), vp(ph), vp(lo), vp(hi), C.c_double(np.float32(0.08)), vp(perm),
                               C.byref(na), C.c_double(np.float32(0.08)), vp(cnt), vp(sums))
        assert rc == 0
        k = na.value
        out.update({f"{tag}_phase": ph, f"{tag}_perm": perm[:k], f"{tag}_n_alive": np.int32(k),
                    f"{tag}_counts": cnt[:k], f"{tag}_id_sums": sums[:k]})
    np.savez_compressed(HERE / "hash_sort.npz", **out)


def jittered(nodes, fluid, dx, fill, seed=0):
    ps = cpu_build_tank_particles("ref", nodes, fluid, dx, fill)
    rng = np.random.default_rng(seed)
    fl = ps.phase != WALL
    ps.x[fl] += rng.uniform(-0.05 * dx, 0.05 * dx, size=(int(fl.sum()), 3)).astype(np.float32)
    lo = ps.x.min(0) - np.float32(4 * dx)
    hi = ps.x.max(0) + np.float32(4 * dx)
    return ps, lo, hi


def box_tank():
    ln2 = fluid_preset("LN2")
    cfg = default_config(dx=0.004, gravity=(0, 0, -0.5), contact_angle_deg=10, velocity_scale=1.0, thermo_mode=1,
                         lee_evaporation=5.0, lee_condensation=5.0)
    ps, lo, hi = jittered(box(0.04, 0.04, 0.04), ln2, 0.004, default_fill(T_liquid=ln2.T_sat + 3.0))
    s = CpuSolver("ref", cfg, ln2, 32)
    s.set_particles(ps, lo, hi)
    s.next_dt()
    out = {"in_" + k: getattr(ps, k) for k in ("x", "v", "nrm", "m", "rho", "T", "phase", "flags", "id")}
    out.update({"lo": lo, "hi": hi})
    for f in FIELDS:
        out["init_" + f] = s.field(f)
    diag = []
    for k in range(20):
        s.step()
        d = s.diagnostics()
        diag.append([d.time, d.mass_liquid, d.mass_vapor, d.enthalpy, d.evaporated_mass, d.p_ullage, d.T_sat,
                     d.max_speed, d.slosh_force.x, d.slosh_force.y, d.slosh_force.z])
    for f in ("x", "v", "m", "T", "rho", "id", "phase", "a", "p"):
        out["step20_" + f] = s.field(f)
    out["diag"] = np.array(diag)
    np.savez_compressed(HERE / "box_tank.npz", **out)


def c1_grid():
    ln2 = fluid_preset("LN2")
    cfg = default_config(dx=0.004, gravity=(0, 0, -9.81e-4), heater_power=20.0, thermo_mode=1)
    ps, lo, hi = jittered(cylinder(0.05, 0.12), ln2, 0.004, default_fill(T_liquid=ln2.T_sat + 2.0))
    s = CpuSolver("ref", cfg, ln2, 32)
    s.set_particles(ps, lo, hi)
    s.next_dt()
    cs, ce = s.grid_tables()
    support = s.params().support
    cnt, sums = s.neighbor_stats(support)
    np.savez_compressed(HERE / "c1_grid.npz", in_x=ps.x, in_phase=ps.phase, lo=lo, hi=hi,
                        sorted_id=s.field("id"), cell_start=cs, cell_end=ce, counts=cnt, id_sums=sums,
                        support=np.float32(support))


if __name__ == "__main__":
    hash_sort()
    box_tank()
    c1_grid()
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size)

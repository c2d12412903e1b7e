"""Lattice generation on the device (cryo_build_tank_particles_device: the
node classes, their compaction, wall normals, fill cut and particle assembly
all evaluated there) against the host generator,
which is bit-exact with the reference's build_tank_particles
(geometry.cpp:242-330): every field of every particle must agree bit for bit,
for analytic tanks, a CSG tree with translations and a difference, mesh tanks
(alone and inside a CSG union), and at the full-size LH2 sphere (1.18 M
particles)."""
import time

import numpy as np
import pytest

import cabi
from test_cabi import GEOMETRIES, icosphere

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "nrm", "m", "rho", "T", "phase", "flags", "id")


def same(a, b):
    assert len(a) == len(b) > 0
    for f in FIELDS:
        assert np.array_equal(getattr(a, f).view(np.uint8), getattr(b, f).view(np.uint8)), f


@pytest.mark.parametrize("name", sorted(GEOMETRIES))
def test_device_lattice_matches_host(name):
    print("test_device_lattice_matches_host", name, flush=True)
    make, dx, fill = GEOMETRIES[name]
    fluid = cabi.hip_fluid_preset("LN2")
    h = cabi.hip_build_tank_particles(make(), fluid, dx, cabi.default_fill(**fill))
    d = cabi.hip_build_tank_particles(make(), fluid, dx, cabi.default_fill(**fill), device=True)
    same(h, d)


# the fill cut and the kept set on their edges: liquid only (the kept fluid
# compacted), nothing liquid, everything liquid, the bubble form
FILLS = {
    "liquid_only": dict(include_vapor=0),
    "no_liquid": dict(fill_fraction=0.0),
    "no_liquid_liquid_only": dict(fill_fraction=0.0, include_vapor=0),
    "all_liquid_liquid_only": dict(fill_fraction=1.0, include_vapor=0),
    "bubble_liquid_only": dict(mode=1, fill_fraction=0.7, bubble_center=(0.01, 0.0, 0.02), include_vapor=0),
    "tilted_settle": dict(fill_fraction=0.45, settle_direction=(0.2, -0.7, -1.0)),
}


@pytest.mark.parametrize("name", sorted(FILLS))
def test_device_lattice_fill_cases_match_host(name):
    print("test_device_lattice_fill_cases_match_host", name, flush=True)
    fluid = cabi.hip_fluid_preset("LN2")
    fill = cabi.default_fill(**FILLS[name])
    h = cabi.hip_build_tank_particles(cabi.capsule(0.04, 0.08), fluid, 0.004, fill)
    d = cabi.hip_build_tank_particles(cabi.capsule(0.04, 0.08), fluid, 0.004, fill, device=True)
    same(h, d)


def test_device_lattice_full_size_sphere():
    print("test_device_lattice_full_size_sphere", flush=True)
    fluid = cabi.hip_fluid_preset("LH2")
    cabi.hip_build_tank_particles(cabi.sphere(0.1), fluid, 0.008, cabi.default_fill(), device=True)  # warm-up
    t0 = time.perf_counter()
    h = cabi.hip_build_tank_particles(cabi.sphere(0.5), fluid, 0.008, cabi.default_fill())
    t1 = time.perf_counter()
    d = cabi.hip_build_tank_particles(cabi.sphere(0.5), fluid, 0.008, cabi.default_fill(), device=True)
    t2 = time.perf_counter()
    # each call builds twice (count, then fill)
    print(f"host generator {t1 - t0:.3f} s, device-evaluated {t2 - t1:.3f} s", flush=True)
    assert len(h) == 1176208
    same(h, d)


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.013, -0.007, 0.021)])
def test_device_mesh_lattice_matches_host(shift):
    """The mesh tank of test_cabi.py's reference check (icosphere, alone and
    translated inside a union with a box, slanted 30 % fill) with 1280
    triangles: nearest-triangle distances and the winding-number sign
    evaluated per lattice node on the device."""
    print("test_device_mesh_lattice_matches_host", shift, flush=True)
    verts, tris = icosphere(0.05, levels=3)
    nodes = [(cabi.GEO_MESH, -1, -1, ())]
    if any(shift):
        nodes += [(cabi.GEO_TRANSLATED, 0, -1, shift), (cabi.GEO_BOX, -1, -1, (0.03, 0.03, 0.14)),
                  (cabi.GEO_UNION, 1, 2, ())]
    arr = cabi.geo_nodes(nodes)
    arr[0].mesh_id = 0
    fluid = cabi.hip_fluid_preset("LN2")
    fill = cabi.default_fill(fill_fraction=0.3, settle_direction=(0.3, 0.2, -1.0))
    t0 = time.perf_counter()
    h = cabi.hip_build_tank_particles(arr, fluid, 0.004, fill, meshes=[(verts, tris)])
    t1 = time.perf_counter()
    d = cabi.hip_build_tank_particles(arr, fluid, 0.004, fill, device=True, meshes=[(verts, tris)])
    t2 = time.perf_counter()
    print(f"{len(h)} particles, {len(tris)} triangles: host {t1 - t0:.3f} s, device-evaluated {t2 - t1:.3f} s",
          flush=True)
    assert int((h.phase == cabi.WALL).sum()) > 0 and int((h.phase == cabi.VAPOR).sum()) > 0
    same(h, d)

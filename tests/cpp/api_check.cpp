// C++ caller of the cryosph:: API (include/cryosph/...), written the way a
// user of the reference's C++ library writes one.  Each check prints a line;
// the exit status is the number of failed checks.
//
//   api_check            every check (needs the GPU)
//   api_check --no-gpu   the checks that need no device
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

#include "cryosph/core/geometry.hpp"
#include "cryosph/core/io.hpp"
#include "cryosph/core/particle_system.hpp"
#include "cryosph/core/spatial_hash.hpp"
#include "cryosph/solvers/sph_solver.hpp"

using namespace cryosph;

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok) ++failures;
}

template <class E, class F>
void expect_throw(F&& f, const std::string& what) {
  try {
    f();
  } catch (const E&) {
    expect(true, what);
    return;
  } catch (...) {
  }
  expect(false, what);
}

bool same_bits(const void* a, const void* b, std::size_t bytes) { return std::memcmp(a, b, bytes) == 0; }

bool same_particles(const HostParticles& a, const HostParticles& b) {
  const std::size_t n = a.size();
  return n == b.size() && same_bits(a.x.data(), b.x.data(), n * sizeof(Vec3)) &&
         same_bits(a.v.data(), b.v.data(), n * sizeof(Vec3)) && same_bits(a.nrm.data(), b.nrm.data(), n * sizeof(Vec3)) &&
         a.m == b.m && a.rho == b.rho && a.T == b.T && a.phase == b.phase && a.flags == b.flags && a.id == b.id;
}

// A box written as a caller's own Geometry subclass: the distance expression of
// an axis-aligned box in the same operation order as the library's BoxTank.
class MyBox final : public Geometry {
 public:
  MyBox(double lx, double ly, double lz) : hx_(0.5 * lx), hy_(0.5 * ly), hz_(0.5 * lz) {}
  double sdf(double x, double y, double z) const override {
    const double qx = std::fabs(x) - hx_, qy = std::fabs(y) - hy_, qz = std::fabs(z) - hz_;
    const double ox = qx < 0 ? 0.0 : qx, oy = qy < 0 ? 0.0 : qy, oz = qz < 0 ? 0.0 : qz;
    const double inner = std::max(std::max(qx, qy), qz);
    return std::sqrt(ox * ox + oy * oy + oz * oz) + std::min(inner, 0.0);
  }
  AABB bounds() const override {
    return AABB{Vec3{float(-hx_), float(-hy_), float(-hz_)}, Vec3{float(hx_), float(hy_), float(hz_)}};
  }
  std::string name() const override { return "my box"; }
  double volume() const override { return 8.0 * hx_ * hy_ * hz_; }

 private:
  double hx_, hy_, hz_;
};

void host_checks() {
  const FluidProperties ln2 = FluidProperties::liquid_nitrogen();
  expect(ln2.name == "LN2" && ln2.rho_liquid > 800 && ln2.rho_liquid / ln2.rho_vapor > 100, "LN2 preset");
  const real p2 = ln2.saturation_pressure(ln2.T_sat + 2);
  expect(p2 > ln2.p_sat && std::fabs(ln2.saturation_temperature(p2) - (ln2.T_sat + 2)) < 1e-3, "Clausius-Clapeyron");

  const CylinderTank cyl(0.05, 0.12);
  expect(std::fabs(cyl.sdf(0, 0, 0) + 0.05) < 1e-12 && std::fabs(cyl.volume() - M_PI * 0.0025 * 0.12) < 1e-12,
         "CylinderTank distance and volume");
  auto outer = std::make_shared<BoxTank>(0.1, 0.1, 0.1);
  auto block = std::make_shared<Translated>(std::make_shared<SphereTank>(0.02), Vec3{0.03f, 0, -0.03f});
  const Difference dif(outer, block);
  expect(dif.sdf(0.03, 0, -0.03) > 0 && dif.sdf(-0.03, 0, 0.03) < 0, "Difference: the obstacle is solid");

  // the C4-style lattice of the reference's own benchmark tank
  const HostParticles c1 = build_tank_particles(cyl, ln2, real(0.004), FillSpec{});
  std::size_t walls = 0;
  for (auto ph : c1.phase) walls += ph == kWall;
  expect(c1.size() == 26360 && walls == 11840, "cylinder r=0.05 h=0.12 dx=4 mm: 26,360 particles, 11,840 walls");

  // a caller's own Geometry subclass (a distance callback) against the analytic box
  FillSpec fill;
  fill.fill_fraction = 0.4;
  fill.settle_direction = Vec3{0.2f, 0, -1};
  const HostParticles mine = build_tank_particles(MyBox(0.05, 0.04, 0.06), ln2, real(0.004), fill);
  const HostParticles lib = build_tank_particles(BoxTank(0.05, 0.04, 0.06), ln2, real(0.004), fill);
  expect(mine.size() > 1000 && same_particles(mine, lib), "own Geometry subclass builds the BoxTank's particles bit for bit");
  const Translated moved(std::make_shared<MyBox>(0.05, 0.04, 0.06), Vec3{0.01f, 0, 0});
  const Translated moved_lib(std::make_shared<BoxTank>(0.05, 0.04, 0.06), Vec3{0.01f, 0, 0});
  expect(same_particles(build_tank_particles(moved, ln2, real(0.004), fill),
                        build_tank_particles(moved_lib, ln2, real(0.004), fill)),
         "own subclass inside Translated");

  // mesh tank from an OBJ file: a unit cube scaled to 6 cm
  const std::string obj = "/tmp/cryosph_api_check_cube.obj";
  {
    std::ofstream f(obj);
    f << "v -0.5 -0.5 -0.5\nv 0.5 -0.5 -0.5\nv 0.5 0.5 -0.5\nv -0.5 0.5 -0.5\n"
         "v -0.5 -0.5 0.5\nv 0.5 -0.5 0.5\nv 0.5 0.5 0.5\nv -0.5 0.5 0.5\n"
         "f 1 4 3 2\nf 5 6 7 8\nf 1 2 6 5\nf 2 3 7 6\nf 3 4 8 7\nf 4 1 5 8\n";
  }
  const MeshTank cube(obj, 0.06);
  expect(cube.num_triangles() == 12 && std::fabs(cube.volume() - 2.16e-4) < 1e-12 &&
             std::fabs(cube.winding_number(0, 0, 0) - 1.0) < 1e-9 && std::fabs(cube.sdf(0, 0, 0) + 0.03) < 1e-12,
         "MeshTank from OBJ: 12 triangles, volume, winding number, distance");

  // file writers against a decode
  HostParticles blk = make_block(Vec3{0, 0, 0}, Vec3{0.03f, 0.02f, 0.01f}, real(0.01), Phase::Liquid, 70.8f, 20.3f);
  blk.push(Vec3{1, 1, 1}, Vec3{3, 4, 0}, 1, 1000, 300, Phase::Wall);
  write_vtk("/tmp/cryosph_api_check.vtk", blk, std::vector<real>(blk.size(), real(1e5)));
  std::ifstream vf("/tmp/cryosph_api_check.vtk", std::ios::binary);
  const std::string vtk((std::istreambuf_iterator<char>(vf)), std::istreambuf_iterator<char>());
  expect(vtk.rfind("# vtk DataFile Version 3.0", 0) == 0 && vtk.find("POINTS 7 float") != std::string::npos &&
             vtk.find("SCALARS pressure float 1") != std::string::npos,
         "write_vtk header");
  const TelemetryFrame fr = make_frame(blk, {}, 0.25, 7);
  expect(fr.n == 7 && fr.speed[6] == 5.0f && fr.pressure[0] == 0.0f, "make_frame: speed column and zero pressure");
  write_json_frame("/tmp/cryosph_api_check.json", fr, false);
  std::ifstream jf("/tmp/cryosph_api_check.json");
  const std::string js((std::istreambuf_iterator<char>(jf)), std::istreambuf_iterator<char>());
  expect(js.find("\"step\":7") != std::string::npos && js.find("\"n\":6") != std::string::npos &&
             js.find("\"phase\":[0,0,0,0,0,0]") != std::string::npos,
         "write_json_frame leaves the wall out");

  ManeuverProfile prof;
  prof.time = {0.0, 1.0};
  prof.acceleration = {Vec3{0, 0, 0}, Vec3{2, 0, 0}};
  prof.angular_velocity = {Vec3{0, 0, 0}, Vec3{0, 0, 1}};
  Vec3 a, w, wd;
  prof.evaluate(0.5, a, w, wd);
  expect(a.x == 1.0f && w.z == 0.5f && wd.z == 1.0f, "ManeuverProfile::evaluate");
  ManeuverProfile bad = prof;
  bad.time = {0.0, 0.0};
  expect_throw<std::invalid_argument>([&] { bad.validate(); }, "ManeuverProfile::validate rejects equal times");
}

void device_checks() {
  expect(mem::device_count() >= 1 && mem::device_name().find("gfx950") != std::string::npos, "gfx950 device");
  const FluidProperties ln2 = FluidProperties::liquid_nitrogen();

  SolverConfig bad;
  bad.contact_angle_deg = 200;
  expect_throw<std::invalid_argument>([&] { Solver s(bad, ln2); }, "contact angle 200 deg: invalid_argument");

  // the header against the C ABI it wraps: the same tank stepped both ways
  SolverConfig cfg;
  cfg.dx = real(0.004);
  cfg.gravity = Vec3{0, 0, real(-0.05)};
  cfg.velocity_scale = real(1.0);
  cfg.thermo_mode = TankThermoMode::Closed;
  cfg.lee_evaporation = cfg.lee_condensation = real(5);
  FillSpec fill;
  fill.T_liquid = ln2.T_sat + real(1.5);
  Solver solver(cfg, ln2);
  solver.build_tank(BoxTank(0.04, 0.04, 0.04), fill);
  solver.relax(0.005);
  solver.run_steps(10);

  const cryo_config cc = cfg.to_c();
  const cryo_fluid cf = ln2.to_c();
  cryo_solver* raw = nullptr;
  cryo_solver_create(&cc, &cf, &raw);
  GeoList g;
  BoxTank(0.04, 0.04, 0.04).emit(g);
  const cryo_fill cfill = fill.to_c();
  cryo_solver_build_tank(raw, int32_t(g.nodes.size()), g.nodes.data(), 0, nullptr, &cfill);
  cryo_solver_relax(raw, 0.005, real(0.98));
  cryo_solver_run_steps(raw, 10);
  const HostParticles hp = solver.particles();
  std::vector<float> x(3 * hp.size()), T(hp.size());
  cryo_solver_download(raw, CRYO_F_X, x.data(), int32_t(hp.size()));
  cryo_solver_download(raw, CRYO_F_T, T.data(), int32_t(hp.size()));
  expect(same_bits(x.data(), hp.x.data(), x.size() * 4) && T == hp.T, "Solver = the C ABI solver, bit for bit");
  cryo_solver_destroy(raw);

  const Diagnostics d = solver.diagnostics();
  expect(d.step == 10 && d.n_total == int(solver.num_particles()) && d.mass_total > 0 && d.evaporated_mass > 0,
         "diagnostics after relax + 10 steps");
  expect(solver.pressure().size() == solver.num_particles() && solver.curvature().size() == solver.num_particles(),
         "pressure and curvature");
  expect(solver.particle_system().capacity() >= 2 * solver.num_particles(), "capacity max(2n, n + 1024)");
  expect(solver.grid().num_cells() > 0 && solver.grid().cell_size() == solver.params().support, "solver grid");
  std::vector<int> counts;
  std::vector<std::uint64_t> sums;
  solver.grid().neighbor_stats(solver.particle_system(), solver.grid().cell_size(), counts, sums, solver.stream());
  expect(counts.size() == solver.num_particles() && counts[0] >= 1, "neighbor_stats through the solver");

  // perf_scaling.cpp's loop: telemetry every 5 steps, zero-copy then async
  for (TelemetryMode mode : {TelemetryMode::ZeroCopy, TelemetryMode::AsyncCopy}) {
    solver.telemetry().configure(solver.particle_system().capacity(), mode);
    TelemetryFrame frame;
    int frames = 0;
    for (int k = 0; k < 20; ++k) {
      solver.step();
      if (k % 5 == 0) {
        if (solver.telemetry().pending()) frames += solver.telemetry().fetch(frame);
        solver.request_telemetry();
      }
    }
    if (solver.telemetry().fetch(frame)) ++frames;
    mem::synchronize(solver.stream());
    const HostParticles now = solver.particles();
    expect(frames == 4 && frame.step == solver.step_count() - 4 && frame.n == int(now.size()),
           std::string("telemetry frames, ") + (mode == TelemetryMode::ZeroCopy ? "ZeroCopy" : "AsyncCopy"));
  }

  // standalone particle system + spatial hash against brute force
  std::mt19937 rng(2024);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  HostParticles cloud;
  for (int i = 0; i < 2500; ++i)
    cloud.push(Vec3{u(rng), u(rng), u(rng)}, Vec3{u(rng), 0, 0}, 1.0f + u(rng), 1000, 20, i % 7 ? Phase::Liquid : Phase::Vapor);
  for (std::size_t i = 0; i < cloud.size(); i += 9) cloud.phase[i] = kDead;
  ParticleSystem ps;
  ps.upload(cloud);
  SpatialHash grid;
  const real radius = real(0.07);
  grid.configure(Vec3{0, 0, 0}, Vec3{1, 1, 1}, radius);
  stream_t st = mem::create_stream();
  const int removed = grid.build(ps, st);
  expect(removed == int((cloud.size() + 8) / 9) && ps.size() == cloud.size() - std::size_t(removed),
         "SpatialHash::build drops dead particles");
  const HostParticles sorted = ps.download();
  bool permuted = true;
  for (std::size_t i = 0; i < sorted.size(); ++i) {
    const std::uint32_t id = sorted.id[i];
    permuted = permuted && sorted.x[i].x == cloud.x[id].x && sorted.m[i] == cloud.m[id] && sorted.phase[i] != kDead;
  }
  expect(permuted, "SpatialHash::build permutes every attribute together");
  grid.neighbor_stats(ps, radius, counts, sums, st);
  bool brute = true;
  for (std::size_t i = 0; i < sorted.size(); i += 11) {
    int bc = 0;
    std::uint64_t bs = 0;
    for (std::size_t j = 0; j < sorted.size(); ++j)
      if (norm2(sorted.x[i] - sorted.x[j]) < radius * radius) ++bc, bs += sorted.id[j];
    brute = brute && counts[i] == bc && sums[i] == bs;
  }
  expect(brute, "neighbor_stats = brute force");
  mem::destroy_stream(st);

  // probes: the liquid pressure is finite inside, NaN far outside
  const std::vector<real> p = solver.probe_pressure({Vec3{0, 0, -0.012f}, Vec3{1, 1, 1}});
  expect(std::isfinite(p[0]) && std::isnan(p[1]), "probe_pressure");
}

}  // namespace

int main(int argc, char** argv) {
  const bool gpu = !(argc > 1 && std::strcmp(argv[1], "--no-gpu") == 0);
  try {
    host_checks();
    if (gpu) device_checks();
  } catch (const std::exception& e) {
    std::printf("FAIL uncaught exception: %s\n", e.what());
    ++failures;
  }
  std::printf("%d failure(s)\n", failures);
  return failures;
}

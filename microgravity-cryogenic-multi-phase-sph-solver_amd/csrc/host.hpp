// Host-side types of the gfx950 CryoSPH library: particle mirror, geometry,
// device buffers and the solver that orchestrates the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "cryosph.h"
#include "cryo_device.hpp"
#include "kernels.hpp"

namespace cryo {

// HIP error -> std::runtime_error (the reference's CRYOSPH_CUDA_CHECK, cuda_utils.cuh:70-80).
#define CRYO_HIP(expr)                                                                                   \
  do {                                                                                                   \
    const hipError_t cryo_err__ = (expr);                                                                \
    if (cryo_err__ != hipSuccess)                                                                        \
      throw std::runtime_error(std::string("HIP error '") + hipGetErrorString(cryo_err__) + "' at " +   \
                               __FILE__ + ":" + std::to_string(__LINE__));                               \
  } while (0)

// Keeps `what` as the calling thread's cryo_last_error() message; returns code.
int record_error(const char* what, int code);

// Body of a C ABI entry point: C++ exceptions become status codes (the
// reference's exception types: invalid_argument, out_of_range, runtime_error).
#define CRYO_GUARD(...)                                                    \
  try {                                                                    \
    __VA_ARGS__;                                                           \
    return CRYO_OK;                                                        \
  } catch (const std::invalid_argument& e) {                               \
    return ::cryo::record_error(e.what(), CRYO_ERR_INVALID_ARGUMENT);      \
  } catch (const std::out_of_range& e) {                                   \
    return ::cryo::record_error(e.what(), CRYO_ERR_OUT_OF_RANGE);          \
  } catch (const std::exception& e) {                                      \
    return ::cryo::record_error(e.what(), CRYO_ERR_RUNTIME);               \
  } catch (...) {                                                          \
    return ::cryo::record_error("unknown error", CRYO_ERR_RUNTIME);        \
  }

// HostParticles (particle_system.hpp:76-90): SoA host mirror.
struct HostParticles {
  std::vector<V3> x, v, nrm;
  std::vector<float> m, rho, T;
  std::vector<uint8_t> phase, flags;
  std::vector<uint32_t> id;
  std::size_t size() const { return x.size(); }
  void push(V3 pos, V3 vel, float mass, float density, float temperature, uint8_t ph, V3 normal = V3{0, 0, 0},
            uint8_t flag = 0) {
    x.push_back(pos);
    v.push_back(vel);
    nrm.push_back(normal);
    m.push_back(mass);
    rho.push_back(density);
    T.push_back(temperature);
    phase.push_back(ph);
    flags.push_back(flag);
    id.push_back(static_cast<uint32_t>(id.size()));
  }
};

// Geometry tree from the flat node list of the C ABI.
class Geometry {
 public:
  Geometry(int n_nodes, const cryo_geo_node* nodes, int n_meshes, const cryo_mesh* meshes, int n_callbacks = 0,
           const cryo_geo_callback* callbacks = nullptr);
  double sdf(double x, double y, double z) const;
  void bounds(V3& lo, V3& hi) const;
  double volume() const;
  V3 outward_normal(double x, double y, double z, double eps) const;
  double mesh_winding(int mesh_id, double x, double y, double z) const;
  // A mesh as flat arrays (vertex k at v[3k..3k+2], triangle t at t[3t..3t+2])
  struct MeshView {
    const double* v;
    const int32_t* t;
    int tris;
  };
  MeshView mesh_view(int mesh_id) const;
  int mesh_count() const { return static_cast<int>(meshes_.size()); }
  std::size_t mesh_vertex_values(int mesh_id) const { return meshes_[mesh_id].v.size(); }
  const std::vector<cryo_geo_node>& node_list() const { return nodes_; }
  bool has_mesh() const;      // mesh or caller-defined nodes (evaluated on the host only)
  bool has_callback() const;
  int depth() const;  // nodes on the longest root-to-leaf path

 private:
  struct MeshData {
    std::vector<double> v;
    std::vector<int32_t> t;
  };
  double sdf_node(int k, double x, double y, double z) const;
  double mesh_sdf(int mesh_id, double x, double y, double z) const;
  void bounds_node(int k, V3& lo, V3& hi) const;
  double volume_node(int k) const;
  std::vector<cryo_geo_node> nodes_;
  std::vector<MeshData> meshes_;
  std::vector<cryo_geo_callback> callbacks_;
};

// on_device: the whole generator runs on the current device (node classes,
// compaction, wall normals, fill cut, particle arrays) and the result is
// downloaded; caller-defined distance functions are refused.
HostParticles build_tank_particles(const Geometry& geo, const cryo_fluid& fluid, float dx, const cryo_fill& fill,
                                   bool on_device = false);

// Owning device buffer (DeviceArray, cuda_utils.cuh:165-244).
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void resize(std::size_t n) {
    if (n == n_) return;
    release();
    if (n) CRYO_HIP(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  // Grows to n elements keeping the first `keep` (stream-ordered copy).
  void grow(std::size_t n, std::size_t keep, hipStream_t s) {
    if (n <= n_) return;
    T* fresh = nullptr;
    CRYO_HIP(hipMalloc(&fresh, n * sizeof(T)));
    keep = keep < n_ ? keep : n_;
    if (keep) CRYO_HIP(hipMemcpyAsync(fresh, p_, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    CRYO_HIP(hipStreamSynchronize(s));
    release();
    p_ = fresh;
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  std::size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

}  // namespace cryo

// Particle file writers of the C ABI: the legacy binary VTK snapshot and the
// JSON viewer frame (io.cpp:52-142 of the reference define both formats, which
// the viewer and ParaView read).  Each file is assembled in memory and written
// with one call, so that a multi-million-particle frame is one sequential write.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "cryosph.h"
#include "host.hpp"

namespace cryo {
namespace {

// Appends big-endian copies of `count` 4-byte words (legacy VTK binary data
// is big-endian; the host is little-endian).
void append_be32(std::string& buf, const void* words, std::size_t count) {
  const std::size_t at = buf.size();
  buf.resize(at + 4 * count);
  const auto* src = static_cast<const uint32_t*>(words);
  char* dst = &buf[at];
  for (std::size_t k = 0; k < count; ++k) {
    uint32_t w;
    std::memcpy(&w, src + k, 4);
    w = __builtin_bswap32(w);
    std::memcpy(dst + 4 * k, &w, 4);
  }
}

void append_text(std::string& buf, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void append_text(std::string& buf, const char* fmt, ...) {
  char tmp[128];
  va_list ap;
  va_start(ap, fmt);
  const int len = std::vsnprintf(tmp, sizeof tmp, fmt, ap);
  va_end(ap);
  if (len < 0) throw std::runtime_error("text formatting failed");
  if (static_cast<std::size_t>(len) < sizeof tmp) {
    buf.append(tmp, static_cast<std::size_t>(len));
    return;
  }
  std::string big(static_cast<std::size_t>(len) + 1, '\0');
  va_start(ap, fmt);
  std::vsnprintf(&big[0], big.size(), fmt, ap);
  va_end(ap);
  buf.append(big.data(), static_cast<std::size_t>(len));
}

void write_file(const std::string& path, const std::string& bytes, const char* who) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error(std::string(who) + ": cannot open '" + path + "'");
  const std::size_t put = bytes.empty() ? 0 : std::fwrite(bytes.data(), 1, bytes.size(), f);
  const bool bad = put != bytes.size() || std::ferror(f);
  if (std::fclose(f) != 0 || bad) throw std::runtime_error(std::string(who) + ": write failed for '" + path + "'");
}

}  // namespace

std::string vtk_image(int32_t n, const float* x, const float* v, const float* T, const float* m, const float* rho,
                      const uint8_t* phase, const float* pressure) {
  const std::size_t N = static_cast<std::size_t>(n);
  std::string b;
  b.reserve(256 + N * (12 + 8 + 4 * 5 + 12));
  b += "# vtk DataFile Version 3.0\nCryoSPH-CUDA particles\nBINARY\nDATASET POLYDATA\n";
  append_text(b, "POINTS %zu float\n", N);
  append_be32(b, x, 3 * N);
  append_text(b, "\nVERTICES %zu %zu\n", N, 2 * N);
  std::vector<int32_t> cells(2 * N);
  for (std::size_t i = 0; i < N; ++i) {
    cells[2 * i] = 1;
    cells[2 * i + 1] = static_cast<int32_t>(i);
  }
  append_be32(b, cells.data(), cells.size());
  append_text(b, "\nPOINT_DATA %zu\n", N);
  const auto scalars = [&](const char* name, const float* vals) {
    append_text(b, "SCALARS %s float 1\nLOOKUP_TABLE default\n", name);
    append_be32(b, vals, N);
    b += '\n';
  };
  std::vector<float> phase_f(N);
  for (std::size_t i = 0; i < N; ++i) phase_f[i] = static_cast<float>(phase[i]);
  scalars("temperature", T);
  scalars("mass", m);
  scalars("density", rho);
  scalars("phase", phase_f.data());
  if (pressure) scalars("pressure", pressure);
  b += "VECTORS velocity float\n";
  append_be32(b, v, 3 * N);
  b += '\n';
  return b;
}

std::string json_frame_image(double time, int64_t step, int32_t n, const float* xyz, const float* T, const float* p,
                             const float* speed, const uint8_t* phase, bool include_walls) {
  std::vector<int32_t> keep;
  keep.reserve(static_cast<std::size_t>(n));
  for (int32_t i = 0; i < n; ++i)
    if (include_walls || phase[i] != CRYO_WALL) keep.push_back(i);
  std::string b;
  b.reserve(64 + keep.size() * 80);
  append_text(b, "{\"time\":%.9g,\"step\":%lld,\"n\":%zu", time, static_cast<long long>(step), keep.size());
  b += ",\"position\":[";
  for (std::size_t k = 0; k < keep.size(); ++k) {
    const float* q = xyz + 3 * static_cast<std::size_t>(keep[k]);
    append_text(b, k ? ",%.6g,%.6g,%.6g" : "%.6g,%.6g,%.6g", double(q[0]), double(q[1]), double(q[2]));
  }
  b += "],\"phase\":[";
  for (std::size_t k = 0; k < keep.size(); ++k) append_text(b, k ? ",%u" : "%u", unsigned(phase[keep[k]]));
  b += ']';
  // non-finite values are written as 0 so that the frame stays valid JSON
  const auto series = [&](const char* name, const float* vals) {
    append_text(b, ",\"%s\":[", name);
    for (std::size_t k = 0; k < keep.size(); ++k) {
      const float val = vals ? vals[keep[k]] : 0.0f;
      append_text(b, k ? ",%.6g" : "%.6g", std::isfinite(val) ? double(val) : 0.0);
    }
    b += ']';
  };
  series("temperature", T);
  series("pressure", p);
  series("speed", speed);
  b += "}\n";
  return b;
}

}  // namespace cryo

extern "C" {

int cryo_write_vtk(const char* path, int32_t n, const float* x, const float* v, const float* T, const float* m,
                   const float* rho, const uint8_t* phase, const float* pressure) {
  CRYO_GUARD({
    if (!path || n < 0 || (n > 0 && (!x || !v || !T || !m || !rho || !phase)))
      throw std::invalid_argument("cryo_write_vtk: null argument");
    cryo::write_file(path, cryo::vtk_image(n, x, v, T, m, rho, phase, pressure), "write_vtk");
  })
}

int cryo_write_json_frame(const char* path, double time, int64_t step, int32_t n, const float* xyz, const float* T,
                          const float* p, const float* speed, const uint8_t* phase, int32_t include_walls) {
  CRYO_GUARD({
    if (!path || n < 0 || (n > 0 && (!xyz || !T || !speed || !phase)))
      throw std::invalid_argument("cryo_write_json_frame: null argument");
    cryo::write_file(path, cryo::json_frame_image(time, step, n, xyz, T, p, speed, phase, include_walls != 0),
                     "write_json_frame");
  })
}

void cryo_frame_speed(int32_t n, const float* v, float* speed) {
  for (int32_t i = 0; i < n; ++i) {
    const float a = v[3 * i], b = v[3 * i + 1], c = v[3 * i + 2];
    speed[i] = std::sqrt(a * a + b * b + c * c);
  }
}

}  // extern "C"

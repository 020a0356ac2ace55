// Output formats of an MD run: the extended-XYZ trajectory writer
// (src/io/trajectory.rs:39-176) and the CSV log writer
// (src/io/log_output.rs:36-110), byte for byte.
//
// Numbers are written as the reference's `{:.Ne}` (Rust LowerExp with N
// fractional digits): the exact binary value rounded once to N+1
// significant decimal digits (ties to even), then `d.ddd..e<exp>` with no
// '+' and no leading zeros in the exponent.  glibc's printf("%.*e") of the
// same value (a float widened to double is exact) rounds the same way, so
// only the exponent and the non-finite spellings are rewritten.
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hmd_internal.h"

namespace hmd {

// Conversion factors atomic unit -> SI per dimension (src/units/mod.rs:85-108,
// constants of src/units/atomic_constants.rs).
static double unit_factor(uint32_t units, uint32_t dim) {
  if (units == HMD_UNITS_ATOMIC || dim == HMD_DIM_DIMENSIONLESS) return 1.0;
  const double bohr = 5.29177210903e-11, me = 9.1093837015e-31, e = 1.602176634e-19;
  const double hartree = 4.359744722207101e-18, tau = 2.4188843265857195e-17;
  switch (dim) {
    case HMD_DIM_LENGTH: return bohr;
    case HMD_DIM_INVERSE_LENGTH: return 1.0 / bohr;
    case HMD_DIM_MASS: return me;
    case HMD_DIM_CHARGE: return e;
    case HMD_DIM_ENERGY: return hartree;
    case HMD_DIM_TIME: return tau;
    case HMD_DIM_INVERSE_TIME: return 1.0 / tau;
    case HMD_DIM_FORCE: return 8.238723498254079e-08;
    case HMD_DIM_PRESSURE: return 29421015696522.1;
    case HMD_DIM_INVERSE_PRESSURE: return 1.0 / 29421015696522.1;
    case HMD_DIM_TEMPERATURE: return 315775.0248040668;
    case HMD_DIM_VELOCITY: return 2187691.2636411153;
    case HMD_DIM_STIFFNESS: return hartree / (bohr * bohr);
    default: throw Error(HMD_ERR_INVALID, "unknown dimension");
  }
}

// Rust `{:.Ne}` of v, appended to out.
static void append_exp(std::string& out, double v, int digits) {
  if (std::isnan(v)) { out += "NaN"; return; }
  if (std::isinf(v)) { out += v < 0 ? "-inf" : "inf"; return; }
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*e", digits, v);
  char* e = std::strchr(buf, 'e');
  *e = '\0';
  out += buf;
  out += 'e';
  const char* x = e + 1;
  bool neg = false;
  if (*x == '+' || *x == '-') { neg = (*x == '-'); ++x; }
  while (*x == '0' && x[1] != '\0') ++x;
  if (neg) out += '-';
  out += x;
}

struct OutFile {
  FILE* f = nullptr;
  std::string buf;
  void open_new(const char* path) {
    // create_new semantics: never overwrite (TrajectoryWriterError::OutputExists)
    f = std::fopen(path, "wx");
    if (!f) {
      if (errno == EEXIST)
        throw Error(HMD_ERR_OUTPUT_EXISTS, std::string("output file already exists: `") + path + "`");
      throw Error(HMD_ERR_IO, std::string(path) + ": " + std::strerror(errno));
    }
  }
  void flush() {
    if (!buf.empty()) {
      if (std::fwrite(buf.data(), 1, buf.size(), f) != buf.size())
        throw Error(HMD_ERR_IO, "failed to write output file");
      buf.clear();
    }
    if (std::fflush(f) != 0) throw Error(HMD_ERR_IO, "failed to flush output file");
  }
  void maybe_flush() { if (buf.size() > (1u << 20)) flush(); }
  ~OutFile() {
    if (f) {
      if (!buf.empty()) std::fwrite(buf.data(), 1, buf.size(), f);
      std::fclose(f);
    }
  }
};

}  // namespace hmd

using namespace hmd;

struct hmd_traj {
  OutFile out;
  uint32_t units = 0;
  bool velocities = false, images = false;
  std::vector<std::string> names;
};

struct hmd_log {
  OutFile out;
  uint32_t units = 0;
  std::vector<uint32_t> dims;
};

extern "C" {

hmd_status hmd_traj_open(const char* path, uint32_t units, int32_t include_velocities,
                         int32_t include_images, const char* const* type_names,
                         uint32_t n_type_names, hmd_traj** out) {
  HMD_GUARD({
    if (!path || !out) throw Error(HMD_ERR_INVALID, "null argument");
    if (units > HMD_UNITS_ATOMIC) throw Error(HMD_ERR_INVALID, "unknown unit system");
    auto* t = new hmd_traj();
    try {
      t->out.open_new(path);
    } catch (...) {
      delete t;
      throw;
    }
    t->units = units;
    t->velocities = include_velocities != 0;
    t->images = include_images != 0;
    for (uint32_t k = 0; k < n_type_names; ++k) t->names.emplace_back(type_names[k]);
    *out = t;
  });
}

// src/io/trajectory.rs:64-168: one frame; lattice {lx, ly, lz, xy, xz, yz}
// in atomic units, positions / velocities as f32 SoA, images as i32 SoA.
hmd_status hmd_traj_write_frame(hmd_traj* t, uint64_t step, double dt, const float lattice[6],
                                uint32_t n, const uint32_t* type_idx, const float* x,
                                const float* y, const float* z, const float* vx, const float* vy,
                                const float* vz, const int32_t* ix, const int32_t* iy,
                                const int32_t* iz) {
  HMD_GUARD({
    if (!t || !lattice || (n && (!type_idx || !x || !y || !z)))
      throw Error(HMD_ERR_INVALID, "null argument");
    if (t->velocities && n && (!vx || !vy || !vz))
      throw Error(HMD_ERR_INVALID, "include_velocities requires velocities");
    if (t->images && n && (!ix || !iy || !iz))
      throw Error(HMD_ERR_INVALID, "include_images requires images");
    const int p = HMD_REAL_FMT_DIGITS;
    // the factors are applied in f32 (the engine's Real), the time in f64
    const float len_f = (float)unit_factor(t->units, HMD_DIM_LENGTH);
    const float vel_f = (float)unit_factor(t->units, HMD_DIM_VELOCITY);
    const double time = (double)step * dt * unit_factor(t->units, HMD_DIM_TIME);
    std::string& b = t->out.buf;
    b += std::to_string(n);
    b += '\n';
    const float lat[9] = {lattice[0] * len_f, 0.0f, 0.0f, lattice[3] * len_f, lattice[1] * len_f,
                          0.0f, lattice[4] * len_f, lattice[5] * len_f, lattice[2] * len_f};
    b += "Lattice=\"";
    for (int k = 0; k < 9; ++k) {
      if (k) b += ' ';
      append_exp(b, (double)lat[k], p);
    }
    b += "\" Properties=";
    b += t->velocities ? (t->images ? "species:S:1:pos:R:3:velo:R:3:image:I:3"
                                    : "species:S:1:pos:R:3:velo:R:3")
                       : (t->images ? "species:S:1:pos:R:3:image:I:3" : "species:S:1:pos:R:3");
    b += " Step=";
    b += std::to_string(step);
    b += " Time=";
    append_exp(b, time, 9);
    b += '\n';
    for (uint32_t i = 0; i < n; ++i) {
      b += type_idx[i] < t->names.size() ? t->names[type_idx[i]] : std::string("?");
      for (const float v : {x[i] * len_f, y[i] * len_f, z[i] * len_f}) {
        b += ' ';
        append_exp(b, (double)v, p);
      }
      if (t->velocities)
        for (const float v : {vx[i] * vel_f, vy[i] * vel_f, vz[i] * vel_f}) {
          b += ' ';
          append_exp(b, (double)v, p);
        }
      if (t->images) {
        char s[48];
        std::snprintf(s, sizeof s, " %d %d %d", ix[i], iy[i], iz[i]);
        b += s;
      }
      b += '\n';
      t->out.maybe_flush();
    }
  });
}

hmd_status hmd_traj_flush(hmd_traj* t) {
  HMD_GUARD({
    if (!t) throw Error(HMD_ERR_INVALID, "null writer");
    t->out.flush();
  });
}

void hmd_traj_close(hmd_traj* t) { delete t; }

// src/io/log_output.rs:39-66: header "step,time,kinetic_energy,temperature"
// plus one column per extra (name, dimension).
hmd_status hmd_log_open(const char* path, uint32_t units, const char* const* extra_names,
                        const uint32_t* extra_dims, uint32_t n_extra, hmd_log** out) {
  HMD_GUARD({
    if (!path || !out || (n_extra && (!extra_names || !extra_dims)))
      throw Error(HMD_ERR_INVALID, "null argument");
    if (units > HMD_UNITS_ATOMIC) throw Error(HMD_ERR_INVALID, "unknown unit system");
    for (uint32_t k = 0; k < n_extra; ++k) unit_factor(HMD_UNITS_SI, extra_dims[k]);   // validate
    auto* l = new hmd_log();
    try {
      l->out.open_new(path);
    } catch (...) {
      delete l;
      throw;
    }
    l->units = units;
    l->out.buf = "step,time,kinetic_energy,temperature";
    for (uint32_t k = 0; k < n_extra; ++k) {
      l->out.buf += ',';
      l->out.buf += extra_names[k];
      l->dims.push_back(extra_dims[k]);
    }
    l->out.buf += '\n';
    *out = l;
  });
}

// src/io/log_output.rs:74-101: one row of atomic-unit scalars converted to
// the output units.
hmd_status hmd_log_write_row(hmd_log* l, uint64_t step, double time, double kinetic_energy,
                             double temperature, const double* extras) {
  HMD_GUARD({
    if (!l || (!l->dims.empty() && !extras)) throw Error(HMD_ERR_INVALID, "null argument");
    const int p = HMD_REAL_FMT_DIGITS;
    std::string& b = l->out.buf;
    b += std::to_string(step);
    b += ',';
    append_exp(b, time * unit_factor(l->units, HMD_DIM_TIME), 9);
    b += ',';
    append_exp(b, kinetic_energy * unit_factor(l->units, HMD_DIM_ENERGY), p);
    b += ',';
    append_exp(b, temperature * unit_factor(l->units, HMD_DIM_TEMPERATURE), p);
    for (size_t k = 0; k < l->dims.size(); ++k) {
      b += ',';
      append_exp(b, extras[k] * unit_factor(l->units, l->dims[k]), p);
    }
    b += '\n';
    l->out.maybe_flush();
  });
}

hmd_status hmd_log_flush(hmd_log* l) {
  HMD_GUARD({
    if (!l) throw Error(HMD_ERR_INVALID, "null writer");
    l->out.flush();
  });
}

void hmd_log_close(hmd_log* l) { delete l; }

// src/timings.rs:513-553 (write_timings_file): a fixed-width table, header
// "{:<28} {:>10} {:>14} {:>13} {:>11} {:>11}", one row per stage with a
// nonzero count as "{:<28} {:>10} {:>14.3} {:>13.1} {:>11.1} {:>11.1}" of
// name, count, total ms, mean us, min us, max us; create-new semantics.
hmd_status hmd_timings_write(const char* path, uint32_t n_stages, const char* const* names,
                             const uint64_t* counts, const uint64_t* total_ns,
                             const uint64_t* min_ns, const uint64_t* max_ns) {
  HMD_GUARD({
    if (!path || (n_stages && (!names || !counts || !total_ns || !min_ns || !max_ns)))
      throw Error(HMD_ERR_INVALID, "null argument");
    OutFile out;
    out.open_new(path);
    char line[512];
    std::snprintf(line, sizeof line, "%-28s %10s %14s %13s %11s %11s\n", "stage", "count",
                  "total_ms", "mean_us", "min_us", "max_us");
    out.buf += line;
    for (uint32_t k = 0; k < n_stages; ++k) {
      if (counts[k] == 0) continue;
      const double total_ms = (double)total_ns[k] / 1000000.0;
      const double mean_us = (double)total_ns[k] / 1000.0 / (double)counts[k];
      std::snprintf(line, sizeof line, "%-28s %10" PRIu64 " %14.3f %13.1f %11.1f %11.1f\n",
                    names[k], counts[k], total_ms, mean_us, (double)min_ns[k] / 1000.0,
                    (double)max_ns[k] / 1000.0);
      out.buf += line;
    }
    out.flush();
  });
}

}  // extern "C"

// sph_capi.cpp - the extern "C" boundary declared in include/sph.h.
//
// Replaces the reference's host-side orchestration: `class_system` methods
// (src/class_system.h:146-651) and the driver loop (src/main.cu:30-561).  Host
// set-up runs on the CPU; the first device call allocates and uploads the
// state.  Methods that mirror a reference method synchronise before returning.
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/sph.h"
#include "sph_device.h"
#include "sph_host.h"

namespace sph {
int device_forces_all(Device *d, int precision, bool store_gradv, std::string &err);
}

struct sph_system {
  sph::HostState hs;
  int precision = SPH_FP64;
  std::unique_ptr<sph::Device> dev;
  std::string err;
  int step = 0;
  bool timing = false;
};

namespace {

using sph::HostState;

int ensure_device(sph_handle h) {
  if (h->dev) return SPH_OK;
  if (!h->hs.initialized) {
    h->err = "the system is not initialized";
    return SPH_ERR_STATE;
  }
  auto dev = sph::make_device(h->precision);
  int rc = dev->upload(h->hs, h->err);
  if (rc) return rc;
  dev->set_timing(h->timing);
  h->dev = std::move(dev);
  return SPH_OK;
}

void magnetic_moment(const sph_params_t &p, int step, double m[3]) {
  double ang = p.omega_magnet * (p.t0 + p.dt * (double)step);
  m[0] = std::cos(ang);
  m[1] = std::sin(ang);
  m[2] = 0.0;
}

struct FieldRef {
  double *d = nullptr;
  int32_t *i = nullptr;
  int64_t n = 0;
  bool device = false;  // changes on the device
};

bool find_field(HostState &s, const char *name, FieldRef &f) {
  const sph_params_t &p = s.p;
  for (int k = 0; k < sph::F_NUM; ++k)
    if (!std::strcmp(name, sph::kPFieldNames[k])) {
      if ((int)s.pf[k].size() < p.N) s.pf[k].resize(p.N, 0.0);
      f.d = s.pf[k].data(); f.n = p.N; f.device = true; return true;
    }
  for (int k = 0; k < sph::C_NUM; ++k)
    if (!std::strcmp(name, sph::kCFieldNames[k])) {
      if ((int)s.cf[k].size() < p.N_colloids) s.cf[k].resize(p.N_colloids, 0.0);
      f.d = s.cf[k].data(); f.n = p.N_colloids; f.device = true; return true;
    }
  if (!std::strcmp(name, "type")) {
    if ((int)s.type.size() < p.N) s.type.resize(p.N, 0);
    f.i = s.type.data(); f.n = p.N; f.device = true; return true;
  }
  if (!std::strcmp(name, "fx_wall") || !std::strcmp(name, "fy_wall") ||
      !std::strcmp(name, "fz_wall")) {
    f.d = s.fwall[name[1] - 'x']; f.n = p.Nwalls; f.device = true; return true;
  }
  struct { const char *n; std::vector<int32_t> *v; } iv[] = {
      {"walls_list", &s.walls_list}, {"walls_start", &s.walls_start},
      {"colloids_list", &s.colloids_list}, {"colloids_start", &s.colloids_start}};
  for (auto &e : iv)
    if (!std::strcmp(name, e.n)) { f.i = e.v->data(); f.n = (int64_t)e.v->size(); return true; }
  struct { const char *n; std::vector<double> *v; } dv[] = {
      {"x_center", &s.x_center}, {"y_center", &s.y_center}, {"z_center", &s.z_center}};
  for (auto &e : dv)
    if (!std::strcmp(name, e.n)) { f.d = e.v->data(); f.n = (int64_t)e.v->size(); return true; }
  return false;
}

int sync_dev(sph_handle h) { return h->dev->synchronize(h->err); }

// Brings the named fields from the device into the host state.
int pull(sph_handle h, std::initializer_list<const char *> names) {
  if (!h->dev) return SPH_OK;
  for (const char *n : names) {
    int rc = h->dev->download_field(h->hs, n, h->err);
    if (rc) return rc;
  }
  return SPH_OK;
}

}  // namespace

extern "C" {

sph_handle sph_create(int precision) {
  if (precision != SPH_FP64 && precision != SPH_FP32) return nullptr;
  sph_handle h = new sph_system();
  h->precision = precision;
  return h;
}

void sph_destroy(sph_handle h) { delete h; }

const char *sph_last_error(sph_handle h) { return h ? h->err.c_str() : "null handle"; }

int sph_precision(sph_handle h) { return h ? h->precision : -1; }

int sph_read_input(sph_handle h, const char *path) {
  if (!h || !path) return SPH_ERR_ARG;
  h->dev.reset();
  h->err.clear();
  return sph::read_input(h->hs, path, h->err);
}

int sph_initialize(sph_handle h, const char *restart_path) {
  if (!h) return SPH_ERR_ARG;
  h->dev.reset();
  h->step = 0;
  return sph::initialize(h->hs, restart_path ? restart_path : "", h->err);
}

int sph_get_params(sph_handle h, sph_params_t *out) {
  if (!h || !out) return SPH_ERR_ARG;
  if (!h->hs.input_read) {
    h->err = "read_input has not succeeded";
    return SPH_ERR_STATE;
  }
  *out = h->hs.p;
  return SPH_OK;
}

int sph_get_field(sph_handle h, const char *name, void *dst, int64_t count) {
  if (!h || !name || (!dst && count)) return SPH_ERR_ARG;
  if (!h->hs.initialized) { h->err = "the system is not initialized"; return SPH_ERR_STATE; }
  FieldRef f;
  if (!find_field(h->hs, name, f)) { h->err = std::string("unknown field ") + name; return SPH_ERR_ARG; }
  if (f.n != count) {
    h->err = std::string("field ") + name + " has " + std::to_string(f.n) + " entries";
    return SPH_ERR_ARG;
  }
  if (h->dev && f.device) {
    int rc = h->dev->download_field(h->hs, name, h->err);
    if (rc) return rc;
    find_field(h->hs, name, f);  // storage may have been reallocated
  }
  if (f.d) std::memcpy(dst, f.d, count * sizeof(double));
  else std::memcpy(dst, f.i, count * sizeof(int32_t));
  return SPH_OK;
}

int sph_set_field(sph_handle h, const char *name, const void *src, int64_t count) {
  if (!h || !name || (!src && count)) return SPH_ERR_ARG;
  if (!h->hs.initialized) { h->err = "the system is not initialized"; return SPH_ERR_STATE; }
  FieldRef f;
  if (!find_field(h->hs, name, f)) { h->err = std::string("unknown field ") + name; return SPH_ERR_ARG; }
  if (f.n != count) {
    h->err = std::string("field ") + name + " has " + std::to_string(f.n) + " entries";
    return SPH_ERR_ARG;
  }
  if (h->dev && !f.device) {
    h->err = std::string("field ") + name + " is fixed once the device state exists";
    return SPH_ERR_STATE;
  }
  if (h->dev && f.device) {
    // Refresh the whole host copy first so untouched entries stay current.
    int rc = h->dev->download_field(h->hs, name, h->err);
    if (rc) return rc;
    find_field(h->hs, name, f);
  }
  if (f.d) std::memcpy(f.d, src, count * sizeof(double));
  else std::memcpy(f.i, src, count * sizeof(int32_t));
  if (h->dev && f.device) return h->dev->upload_field(h->hs, name, h->err);
  return SPH_OK;
}

int sph_get_cell_list(sph_handle h, int32_t *pc, int32_t *pi, int32_t *cs, int32_t *ce) {
  if (!h) return SPH_ERR_ARG;
  if (!h->dev) { h->err = "no neighbours_search has run"; return SPH_ERR_STATE; }
  return h->dev->cell_list(pc, pi, cs, ce, h->err);
}

#define DEVICE_CALL(expr)                          \
  do {                                             \
    if (!h) return SPH_ERR_ARG;                    \
    int rc_ = ensure_device(h);                    \
    if (rc_) return rc_;                           \
    rc_ = (expr);                                  \
    if (rc_) return rc_;                           \
    return sync_dev(h);                            \
  } while (0)

int sph_neighbours_search(sph_handle h) { DEVICE_CALL(h->dev->neighbours_search(h->err)); }
int sph_calculate_density(sph_handle h) { DEVICE_CALL(h->dev->density(h->err)); }
int sph_calculate_vel_gradient(sph_handle h) { DEVICE_CALL(h->dev->vel_gradient(true, h->err)); }
int sph_calculate_pressure(sph_handle h) { DEVICE_CALL(h->dev->pressure(h->err)); }
int sph_calculate_external_forces(sph_handle h) {
  DEVICE_CALL(h->dev->external_forces(h->err));
}
int sph_calculate_force_walls(sph_handle h) { DEVICE_CALL(h->dev->force_walls(h->err)); }
int sph_calculate_force_colloids(sph_handle h) { DEVICE_CALL(h->dev->force_colloids(h->err)); }
int sph_calculate_forces(sph_handle h) {
  DEVICE_CALL(sph::device_forces_all(h->dev.get(), h->precision, true, h->err));
}
int sph_V_Verlet_step1(sph_handle h) { DEVICE_CALL(h->dev->vv1(h->err)); }
int sph_V_Verlet_step2(sph_handle h) { DEVICE_CALL(h->dev->vv2(h->err)); }

int sph_set_step(sph_handle h, int step) {
  if (!h) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  double m[3];
  magnetic_moment(h->hs.p, step, m);
  h->dev->set_magnetic_moment(m);
  h->step = step;
  return SPH_OK;
}

int sph_prepare(sph_handle h) {
  if (!h) return SPH_ERR_ARG;
  int rc = sph_set_step(h, 0);
  if (rc) return rc;
  if ((rc = h->dev->neighbours_search(h->err))) return rc;
  if ((rc = sph::device_forces_all(h->dev.get(), h->precision, true, h->err))) return rc;
  return sync_dev(h);
}

int sph_step(sph_handle h, int nsteps) {
  if (!h || nsteps < 0) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  for (int n = 0; n < nsteps; ++n) {
    const int step = h->step + 1;
    if ((rc = h->dev->vv1(h->err))) return rc;
    if ((rc = h->dev->neighbours_search(h->err))) return rc;
    double m[3];
    magnetic_moment(h->hs.p, step, m);
    h->dev->set_magnetic_moment(m);
    if ((rc = sph::device_forces_all(h->dev.get(), h->precision, false, h->err))) return rc;
    if ((rc = h->dev->vv2(h->err))) return rc;
    h->step = step;
  }
  return SPH_OK;
}

int sph_synchronize(sph_handle h) {
  if (!h) return SPH_ERR_ARG;
  if (!h->dev) return SPH_OK;
  return sync_dev(h);
}

int sph_current_step(sph_handle h) { return h ? h->step : -1; }

int sph_kinetic_energy(sph_handle h, double *ke) {
  if (!h || !ke) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  return h->dev->kinetic_energy(ke, h->err);
}

int sph_observables(sph_handle h, double out[8]) {
  if (!h || !out) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  if ((rc = h->dev->kinetic_energy(&out[0], h->err))) return rc;
  double s[6];
  if ((rc = h->dev->observable_sums(s, h->err))) return rc;
  const sph_params_t &p = h->hs.p;
  const double nan = std::nan("");
  out[1] = s[2] > 0 ? p.nkT * s[0] / s[2] : nan;
  out[2] = s[2] > 0 ? p.nkT * s[1] / s[2] : nan;
  out[3] = out[4] = nan;
  if (p.wall) {
    if ((rc = pull(h, {"fx_wall"}))) return rc;
    const double area = p.dim == 3 ? p.L[0] * p.L[2] : p.L[0];
    out[3] = h->hs.fwall[0][0] / area;
    out[4] = h->hs.fwall[0][1] / area;
  }
  auto half_turn = [](double a) { return a - M_PI * std::ceil(a / M_PI - 0.5); };
  out[5] = s[5] > 0 ? half_turn(0.5 * std::atan2(s[3], s[4])) : nan;
  out[6] = half_turn(p.omega_magnet * (p.t0 + p.dt * (double)h->step));
  out[7] = half_turn(out[6] - out[5]);
  return SPH_OK;
}

int sph_print_output(sph_handle h, int step, const char *out_dir) {
  if (!h || !out_dir) return SPH_ERR_ARG;
  if (!h->hs.initialized) { h->err = "the system is not initialized"; return SPH_ERR_STATE; }
  const sph_params_t &p = h->hs.p;
  std::string dir = out_dir;
  int rc;
  if (h->dev && (rc = sync_dev(h))) return rc;
  if (p.freq_micro > 0 && step % p.freq_micro == 0) {
    if ((rc = pull(h, {"x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "cxx", "cxy",
                       "cyy", "cxz", "cyz", "czz"})))
      return rc;
    if ((rc = sph::write_micro(h->hs, step, dir, h->err))) return rc;
  }
  if (p.wall == 1 && p.freq_walls > 0 && step % p.freq_walls == 0) {
    if ((rc = pull(h, {"fx_wall", "fy_wall", "fz_wall"}))) return rc;
    if ((rc = sph::write_walls(h->hs, step, dir, h->err))) return rc;
  }
  if (p.N_colloids > 0 && p.freq_colloids > 0 && step % p.freq_colloids == 0) {
    for (int k = 0; k < sph::C_NUM; ++k)
      if ((rc = pull(h, {sph::kCFieldNames[k]}))) return rc;
    if ((rc = sph::write_colloids(h->hs, step, dir, h->err))) return rc;
  }
  if (p.freq_macro > 0 && step % p.freq_macro == 0) {
    double ke = 0.0;
    if ((rc = sph_kinetic_energy(h, &ke))) return rc;
    if ((rc = sph::append_macro(h->hs, step, ke, dir, h->err))) return rc;
  }
  return SPH_OK;
}

int sph_run(sph_handle h, const char *out_dir) {
  if (!h || !out_dir) return SPH_ERR_ARG;
  int rc = sph_prepare(h);
  if (rc) return rc;
  if ((rc = sph_print_output(h, 0, out_dir))) return rc;
  const int Nsteps = h->hs.p.Nsteps;
  for (int step = 1; step <= Nsteps; ++step) {
    if ((rc = sph_step(h, 1))) return rc;
    const sph_params_t &p = h->hs.p;
    bool out = (p.freq_micro > 0 && step % p.freq_micro == 0) ||
               (p.freq_macro > 0 && step % p.freq_macro == 0) ||
               (p.wall && p.freq_walls > 0 && step % p.freq_walls == 0) ||
               (p.N_colloids > 0 && p.freq_colloids > 0 && step % p.freq_colloids == 0);
    if (out && (rc = sph_print_output(h, step, out_dir))) return rc;
  }
  return sph_synchronize(h);
}

int sph_write_restart(sph_handle h, const char *path) {
  if (!h || !path) return SPH_ERR_ARG;
  if (!h->hs.initialized) { h->err = "the system is not initialized"; return SPH_ERR_STATE; }
  int rc;
  if (h->dev && (rc = sync_dev(h))) return rc;
  if ((rc = pull(h, {"x", "y", "z", "vx", "vy", "vz", "mass", "cxx", "cxy", "cyy", "cxz", "cyz",
                     "czz"})))
    return rc;
  return sph::write_restart(h->hs, path, h->err);
}

int sph_set_timing(sph_handle h, int enable) {
  if (!h) return SPH_ERR_ARG;
  h->timing = enable != 0;
  if (h->dev) h->dev->set_timing(h->timing);
  return SPH_OK;
}

int sph_get_timing(sph_handle h, double ms[8], int64_t calls[8]) {
  if (!h || !ms || !calls) return SPH_ERR_ARG;
  if (!h->dev) {
    for (int k = 0; k < 8; ++k) { ms[k] = 0.0; calls[k] = 0; }
    return SPH_OK;
  }
  return h->dev->get_timing(ms, calls, h->err);
}

}  // extern "C"

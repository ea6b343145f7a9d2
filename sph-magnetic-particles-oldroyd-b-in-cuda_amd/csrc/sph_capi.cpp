// sph_capi.cpp - the extern "C" boundary declared in include/sph.h.
//
// Replaces the reference's host-side orchestration: `class_system` methods
// (src/class_system.h:146-651) and the driver loop (src/main.cu:30-561).  Host
// set-up runs on the CPU; the first device call allocates and uploads the
// state.  Methods that mirror a reference method synchronise before returning.
// The sph_batch_* calls run many such systems in one device state.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sph.h"
#include "sph_device.h"
#include "sph_host.h"

struct sph_system {
  sph::HostState hs;
  int precision = SPH_FP64;
  std::unique_ptr<sph::Device> dev;
  std::string err;
  int step = 0;
  bool timing = false;
};

struct sph_batch {
  int precision = SPH_FP64;
  std::vector<sph::HostState> sims;  // host copies, ID order
  std::unique_ptr<sph::Device> dev;
  std::string err;
  int step = 0;
  int block = 256, xcd = 1;
  std::vector<int> failed;  // per simulation, sticky: SPH_BATCH_* bits
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
  int rc = dev->upload({&h->hs}, h->err);
  if (rc) return rc;
  dev->set_timing(h->timing);
  h->dev = std::move(dev);
  return SPH_OK;
}

// Steps run as captured hipGraphs unless SPH_NO_GRAPH is set.
bool use_graph() {
  static const bool on = std::getenv("SPH_NO_GRAPH") == nullptr;
  return on;
}

// Rheology observables of one simulation from its device sums (sph.h).
void finish_observables(const sph_params_t &p, int step, double ke, const double s[6],
                        const double fwall_x[2], double out[8]) {
  const double nan = std::nan("");
  out[0] = ke;
  out[1] = s[2] > 0 ? p.nkT * s[0] / s[2] : nan;
  out[2] = s[2] > 0 ? p.nkT * s[1] / s[2] : nan;
  out[3] = out[4] = nan;
  if (p.wall) {
    const double area = p.dim == 3 ? p.L[0] * p.L[2] : p.L[0];
    out[3] = fwall_x[0] / area;
    out[4] = fwall_x[1] / area;
  }
  auto half_turn = [](double a) { return a - M_PI * std::ceil(a / M_PI - 0.5); };
  out[5] = s[5] > 0 ? half_turn(0.5 * std::atan2(s[3], s[4])) : nan;
  out[6] = half_turn(p.omega_magnet * (p.t0 + p.dt * (double)step));
  out[7] = half_turn(out[6] - out[5]);
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
    int rc = h->dev->download_field(0, h->hs, n, h->err);
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
    int rc = h->dev->download_field(0, h->hs, name, h->err);
    if (rc) return rc;
    find_field(h->hs, name, f);  // storage may have been reallocated
  }
  if (count == 0) return SPH_OK;  // (an empty field's storage may be null)
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
    int rc = h->dev->download_field(0, h->hs, name, h->err);
    if (rc) return rc;
    find_field(h->hs, name, f);
  }
  if (count > 0) {  // (an empty field's storage may be null)
    if (f.d) std::memcpy(f.d, src, count * sizeof(double));
    else std::memcpy(f.i, src, count * sizeof(int32_t));
  }
  if (h->dev && f.device) return h->dev->upload_field(0, h->hs, name, h->err);
  return SPH_OK;
}

int sph_get_cell_list(sph_handle h, int32_t *pc, int32_t *pi, int32_t *cs, int32_t *ce) {
  if (!h) return SPH_ERR_ARG;
  if (!h->dev) { h->err = "no neighbours_search has run"; return SPH_ERR_STATE; }
  return h->dev->cell_list(0, pc, pi, cs, ce, h->err);
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
int sph_calculate_forces(sph_handle h) { DEVICE_CALL(h->dev->forces_all(true, h->err)); }
int sph_V_Verlet_step1(sph_handle h) { DEVICE_CALL(h->dev->vv1(h->err)); }
int sph_V_Verlet_step2(sph_handle h) { DEVICE_CALL(h->dev->vv2(h->err)); }

int sph_set_step(sph_handle h, int step) {
  if (!h) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  if ((rc = h->dev->set_step(step, h->err))) return rc;
  h->step = step;
  return SPH_OK;
}

int sph_prepare(sph_handle h) {
  if (!h) return SPH_ERR_ARG;
  int rc = sph_set_step(h, 0);
  if (rc) return rc;
  if ((rc = h->dev->neighbours_search(h->err))) return rc;
  if ((rc = h->dev->forces_all(true, h->err))) return rc;
  return sync_dev(h);
}

int sph_step(sph_handle h, int nsteps) {
  if (!h || nsteps < 0) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  if ((rc = h->dev->run_steps(nsteps, h->step + 1, use_graph(), h->err))) return rc;
  h->step += nsteps;
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
  if ((rc = h->dev->kinetic_energy(ke, h->err))) return rc;
  if (!std::isfinite(*ke)) {
    h->err = "kinetic energy is not finite";
    return SPH_ERR_NAN;
  }
  return SPH_OK;
}

int sph_observables(sph_handle h, double out[8]) {
  if (!h || !out) return SPH_ERR_ARG;
  int rc = ensure_device(h);
  if (rc) return rc;
  double ke = 0.0, s[6];
  if ((rc = sph_kinetic_energy(h, &ke))) return rc;
  if ((rc = h->dev->observable_sums(s, h->err))) return rc;
  if (h->hs.p.wall && (rc = pull(h, {"fx_wall"}))) return rc;
  finish_observables(h->hs.p, h->step, ke, s, h->hs.fwall[0], out);
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

int sph_checkpoint_write(sph_handle h, const char *path) {
  if (!h || !path) return SPH_ERR_ARG;
  if (!h->hs.initialized) { h->err = "the system is not initialized"; return SPH_ERR_STATE; }
  if (h->dev) {
    int rc = sync_dev(h);
    if (rc) return rc;
    // every device-held field, in ID order (walls only exist with walls)
    for (int k = 0; k < sph::F_NUM; ++k)
      if ((rc = h->dev->download_field(0, h->hs, sph::kPFieldNames[k], h->err))) return rc;
    for (int k = 0; k < sph::C_NUM && h->hs.p.N_colloids > 0; ++k)
      if ((rc = h->dev->download_field(0, h->hs, sph::kCFieldNames[k], h->err))) return rc;
    if ((rc = pull(h, {"type"}))) return rc;
    if (h->hs.p.wall && (rc = pull(h, {"fx_wall", "fy_wall", "fz_wall"}))) return rc;
  }
  return sph::write_checkpoint(h->hs, h->precision, h->step, path, h->err);
}

int sph_checkpoint_read(sph_handle h, const char *path) {
  if (!h || !path) return SPH_ERR_ARG;
  int64_t step = 0;
  int rc = sph::read_checkpoint(h->hs, h->precision, &step, path, h->err);
  if (rc) return rc;
  // the next device call uploads the restored state into a new device state
  // (its lists are rebuilt at the first step)
  h->dev.reset();
  h->step = (int)step;
  return SPH_OK;
}

int sph_set_timing(sph_handle h, int enable) {
  if (!h) return SPH_ERR_ARG;
  h->timing = enable != 0;
  if (h->dev) h->dev->set_timing(h->timing);
  return SPH_OK;
}

int sph_get_counters(sph_handle h, int64_t out[4]) {
  if (!h || !out) return SPH_ERR_ARG;
  out[0] = h->step;
  out[1] = out[2] = out[3] = 0;
  if (!h->dev) return SPH_OK;
  int64_t c[3];
  int rc = h->dev->counters(0, c, h->err);
  if (rc) return rc;
  out[1] = c[0];
  out[2] = c[1];
  out[3] = c[2];
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

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
sph_batch_handle sph_batch_create(int precision) {
  if (precision != SPH_FP64 && precision != SPH_FP32) return nullptr;
  sph_batch_handle b = new sph_batch();
  b->precision = precision;
  return b;
}

void sph_batch_destroy(sph_batch_handle b) { delete b; }

const char *sph_batch_last_error(sph_batch_handle b) {
  return b ? b->err.c_str() : "null handle";
}

int sph_batch_add(sph_batch_handle b, sph_handle h, int *index) {
  if (!b || !h) return SPH_ERR_ARG;
  if (b->dev) {
    b->err = "simulations can only be added before sph_batch_prepare";
    return SPH_ERR_STATE;
  }
  if (!h->hs.initialized) {
    b->err = "the system is not initialized";
    return SPH_ERR_STATE;
  }
  int rc;
  if (h->dev && (rc = pull(h, {"x", "y", "z", "vx", "vy", "vz", "mass", "fx", "fy", "fz", "cxx",
                              "cxy", "cyy", "cxz", "cyz", "czz", "type"}))) {
    b->err = h->err;
    return rc;
  }
  if (h->dev)
    for (int k = 0; k < sph::C_NUM; ++k)
      if ((rc = pull(h, {sph::kCFieldNames[k]}))) {
        b->err = h->err;
        return rc;
      }
  if (!b->sims.empty()) {
    const sph_params_t &p0 = b->sims[0].p;
    if (p0.dim != h->hs.p.dim || p0.wall != h->hs.p.wall) {
      b->err = "all simulations of a batch must share dim and wall";
      return SPH_ERR_ARG;
    }
  }
  b->sims.push_back(h->hs);
  if (index) *index = (int)b->sims.size() - 1;
  return SPH_OK;
}

int sph_batch_size(sph_batch_handle b) { return b ? (int)b->sims.size() : -1; }

int sph_batch_prepare(sph_batch_handle b) {
  if (!b) return SPH_ERR_ARG;
  if (b->sims.empty()) {
    b->err = "the batch is empty";
    return SPH_ERR_STATE;
  }
  std::vector<const sph::HostState *> v;
  for (auto &s : b->sims) v.push_back(&s);
  auto dev = sph::make_device(b->precision);
  int rc;
  if ((rc = dev->upload(v, b->err))) return rc;
  if ((rc = dev->set_launch(b->block, b->xcd, b->err))) return rc;
  b->dev = std::move(dev);
  b->step = 0;
  b->failed.assign(b->sims.size(), 0);
  if ((rc = b->dev->set_step(0, b->err))) return rc;
  if ((rc = b->dev->neighbours_search(b->err))) return rc;
  if ((rc = b->dev->forces_all(true, b->err))) return rc;
  return sph_batch_synchronize(b);
}

#define BATCH_READY(b)                               \
  do {                                               \
    if (!(b)) return SPH_ERR_ARG;                    \
    if (!(b)->dev) {                                 \
      (b)->err = "sph_batch_prepare has not run";    \
      return SPH_ERR_STATE;                          \
    }                                                \
  } while (0)

int sph_batch_step(sph_batch_handle b, int nsteps) {
  BATCH_READY(b);
  if (nsteps < 0) return SPH_ERR_ARG;
  int rc = b->dev->run_steps(nsteps, b->step + 1, use_graph(), b->err);
  if (rc) return rc;
  b->step += nsteps;
  return SPH_OK;
}

// Waits for the queued steps and collects the per-simulation flags; a failed
// simulation does not fail the call (the others are unaffected).
int sph_batch_synchronize(sph_batch_handle b) {
  BATCH_READY(b);
  std::vector<int> f;
  int rc = b->dev->poll_flags(f, b->err);
  if (rc) return rc;
  for (size_t s = 0; s < f.size(); ++s)
    if (f[s]) b->failed[s] |= SPH_BATCH_OUT_OF_GRID;
  return SPH_OK;
}

int sph_batch_failed(sph_batch_handle b, int32_t *flags) {
  BATCH_READY(b);
  if (!flags) return SPH_ERR_ARG;
  int rc = sph_batch_synchronize(b);
  if (rc) return rc;
  for (size_t s = 0; s < b->failed.size(); ++s) flags[s] = b->failed[s];
  return SPH_OK;
}

int sph_batch_current_step(sph_batch_handle b) { return b ? b->step : -1; }

int sph_batch_get_field(sph_batch_handle b, int sim, const char *name, void *dst,
                        int64_t count) {
  if (!b || !name || (!dst && count)) return SPH_ERR_ARG;
  if (sim < 0 || sim >= (int)b->sims.size()) {
    b->err = "simulation index out of range";
    return SPH_ERR_ARG;
  }
  sph::HostState &hs = b->sims[sim];
  FieldRef f;
  if (!find_field(hs, name, f)) {
    b->err = std::string("unknown field ") + name;
    return SPH_ERR_ARG;
  }
  if (f.n != count) {
    b->err = std::string("field ") + name + " has " + std::to_string(f.n) + " entries";
    return SPH_ERR_ARG;
  }
  if (b->dev && f.device) {
    int rc = b->dev->download_field(sim, hs, name, b->err);
    if (rc) return rc;
    find_field(hs, name, f);
  }
  if (count == 0) return SPH_OK;  // (an empty field's storage may be null)
  if (f.d) std::memcpy(dst, f.d, count * sizeof(double));
  else std::memcpy(dst, f.i, count * sizeof(int32_t));
  return SPH_OK;
}

int sph_batch_kinetic_energy(sph_batch_handle b, double *ke) {
  BATCH_READY(b);
  if (!ke) return SPH_ERR_ARG;
  int rc = b->dev->kinetic_energy(ke, b->err);
  if (rc) return rc;
  for (size_t s = 0; s < b->sims.size(); ++s)
    if (!std::isfinite(ke[s])) b->failed[s] |= SPH_BATCH_NAN;
  return SPH_OK;
}

int sph_batch_observables(sph_batch_handle b, double *out) {
  BATCH_READY(b);
  if (!out) return SPH_ERR_ARG;
  const int S = (int)b->sims.size();
  std::vector<double> ke(S), sums(6 * S), fw(6 * S);
  int rc;
  if ((rc = sph_batch_kinetic_energy(b, ke.data()))) return rc;
  if ((rc = b->dev->observable_sums(sums.data(), b->err))) return rc;
  if ((rc = b->dev->wall_forces(fw.data(), b->err))) return rc;
  for (int s = 0; s < S; ++s) {
    sph::HostState &hs = b->sims[s];
    if (hs.p.wall)
      for (int a = 0; a < 3; ++a)
        for (int w = 0; w < 2; ++w) hs.fwall[a][w] = fw[6 * s + 2 * a + w];
    finish_observables(hs.p, b->step, ke[s], &sums[6 * s], hs.fwall[0], out + 8 * s);
  }
  return SPH_OK;
}

int sph_batch_set_launch(sph_batch_handle b, int block, int xcd) {
  if (!b) return SPH_ERR_ARG;
  if (block != 64 && block != 128 && block != 256) {
    b->err = "block must be 64, 128 or 256 threads";
    return SPH_ERR_ARG;
  }
  b->block = block;
  b->xcd = xcd ? 1 : 0;
  return b->dev ? b->dev->set_launch(block, b->xcd, b->err) : SPH_OK;
}

int sph_batch_get_launch(sph_batch_handle b, int *block, int *xcd) {
  if (!b || !block || !xcd) return SPH_ERR_ARG;
  *block = b->block;
  *xcd = b->xcd;
  return SPH_OK;
}

int sph_batch_autotune(sph_batch_handle b, int steps, int *block, int *xcd,
                       double *ms_per_step) {
  BATCH_READY(b);
  if (steps < 1) return SPH_ERR_ARG;
  static const int shapes[6][2] = {{256, 1}, {128, 1}, {64, 1}, {256, 0}, {128, 0}, {64, 0}};
  double best = -1.0;
  int bb = b->block, bx = b->xcd, rc;
  for (auto &sh : shapes) {
    if ((rc = sph_batch_set_launch(b, sh[0], sh[1]))) return rc;
    // one untimed step captures the graphs of both buffer parities
    if ((rc = sph_batch_step(b, 2)) || (rc = sph_batch_synchronize(b))) return rc;
    auto t0 = std::chrono::steady_clock::now();
    if ((rc = sph_batch_step(b, steps)) || (rc = sph_batch_synchronize(b))) return rc;
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count() / steps;
    if (best < 0 || ms < best) {
      best = ms;
      bb = sh[0];
      bx = sh[1];
    }
  }
  if ((rc = sph_batch_set_launch(b, bb, bx))) return rc;
  if (block) *block = bb;
  if (xcd) *xcd = bx;
  if (ms_per_step) *ms_per_step = best;
  return SPH_OK;
}

}  // extern "C"

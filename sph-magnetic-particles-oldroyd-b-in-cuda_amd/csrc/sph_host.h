// sph_host.h - host-side state of one simulation (particle-ID order).
//
// Mirrors the data members of the reference's `class_system`
// (src/class_system.h:12-143): SoA host arrays of length Nmax in particle-ID
// order, wall and colloid lists, per-colloid rigid-body state and the
// derived constants (held in sph_params_t).  The device copy lives in
// sph_device.hip; this part compiles and runs without a GPU.
#pragma once

#include <string>
#include <vector>
#include <cstdint>

#include "../../include/sph.h"

namespace sph {

constexpr int kMaxColloids = 1000;  // src/config.h:22 (MAX_COLLOIDS)

// Names of the per-particle real fields, in the order of HostState::pf.
enum PField {
  F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_MASS, F_DENS, F_PRESS, F_FX, F_FY, F_FZ,
  F_CXX, F_CXY, F_CYY, F_CXZ, F_CYZ, F_CZZ,
  F_DCDT_XX, F_DCDT_XY, F_DCDT_YY, F_DCDT_XZ, F_DCDT_YZ, F_DCDT_ZZ,
  F_GXX, F_GXY, F_GXZ, F_GYX, F_GYY, F_GYZ, F_GZX, F_GZY, F_GZZ,
  F_NUM
};
extern const char *const kPFieldNames[F_NUM];

// Names of the per-colloid real fields, in the order of HostState::cf.
enum CField {
  C_X, C_Y, C_Z, C_VX, C_VY, C_VZ, C_OX, C_OY, C_OZ, C_THETA,
  C_Q0, C_QX, C_QY, C_QZ, C_FX, C_FY, C_FZ, C_TX, C_TY, C_TZ,
  C_NUM
};
extern const char *const kCFieldNames[C_NUM];

struct HostState {
  sph_params_t p{};
  bool input_read = false;
  bool initialized = false;
  std::string input_dir;  // directory of the input file (restart default)

  // Colloid centres as read from the input (coll_x lines).
  std::vector<double> coll_x_data;  // 3 * N_colloids

  std::vector<double> pf[F_NUM];    // per particle, length Nmax
  std::vector<int32_t> type;        // per particle
  std::vector<double> cf[C_NUM];    // per colloid
  double fwall[3][2] = {{0, 0}, {0, 0}, {0, 0}};

  std::vector<int32_t> walls_list, walls_start;
  std::vector<int32_t> colloids_list, colloids_start;
  std::vector<double> x_center, y_center, z_center;
};

// Parses the reference input grammar.  Returns an SPH_* status and a message.
int read_input(HostState &s, const std::string &path, std::string &err);
// constructor + initialize + new/old sim particle set-up.
int initialize(HostState &s, const std::string &restart_path, std::string &err);

// Output writers (ID order).  All return SPH_OK or SPH_ERR_IO.
int write_micro(const HostState &s, int step, const std::string &dir, std::string &err);
int write_walls(const HostState &s, int step, const std::string &dir, std::string &err);
int write_colloids(const HostState &s, int step, const std::string &dir, std::string &err);
int append_macro(const HostState &s, int step, double ke, const std::string &dir,
                 std::string &err);
int write_restart(const HostState &s, const std::string &path, std::string &err);
// Binary checkpoint of the whole host state (parameters, every particle,
// colloid and wall array and list, at full precision) and the step counter.
// read_checkpoint replaces `s` only when the file reads completely; it keeps
// s.input_dir.  SPH_ERR_IO on a missing, foreign or damaged file, SPH_ERR_ARG
// when the file holds the other precision form.
int write_checkpoint(const HostState &s, int precision, int64_t step, const std::string &path,
                     std::string &err);
int read_checkpoint(HostState &s, int precision, int64_t *step, const std::string &path,
                    std::string &err);

}  // namespace sph

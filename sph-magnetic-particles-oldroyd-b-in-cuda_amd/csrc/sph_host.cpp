// sph_host.cpp - input parsing, system set-up and output writers (host only).
//
// Behaviour follows the reference's host class:
//   read_input      src/system_read_input.cpp:18-512
//   constructor     src/system_constructor.cpp:15-108
//   initialize      src/system_initialize.cpp:16-359
//   new-sim lattice src/system_initialize_particles_new_sim.cpp:16-253
//   restart reader  src/system_initialize_particles_old_sim.cpp:17-112
//   writers         src/system_print_particles.cpp:57-91, src/system_print_walls.cpp:32-44,
//                   src/system_print_colloids.cpp:81-111, src/system_print_macro_vars.cu:37-47
// Deliberate differences (each is a defect of the reference):
//   * a missing input file is an error (reference returns success, read_input:58-62);
//   * the 3D colloid-overlap check uses z for z (reference uses y, read_input:485);
//   * colloid boundary particles get mass = 0, v = 0, c = 0 (reference leaves them
//     uninitialised, new_sim:115 vs 155-244);
//   * every appended colloid-shell particle is checked against Nmax (reference checks
//     only the 3D north pole, new_sim:190);
//   * the 3D shell poles are wrapped into the box like the ring particles (reference
//     leaves them unwrapped, new_sim:194-196, 241-243, so a colloid within R of z = 0
//     or z = L puts a pole outside the cell grid and trips its device assert).
#include "sph_host.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <fstream>

namespace sph {

const char *const kPFieldNames[F_NUM] = {
    "x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "fx", "fy", "fz",
    "cxx", "cxy", "cyy", "cxz", "cyz", "czz",
    "dcdt_xx", "dcdt_xy", "dcdt_yy", "dcdt_xz", "dcdt_yz", "dcdt_zz",
    "gradvxx", "gradvxy", "gradvxz", "gradvyx", "gradvyy", "gradvyz",
    "gradvzx", "gradvzy", "gradvzz"};

const char *const kCFieldNames[C_NUM] = {
    "coll_x", "coll_y", "coll_z", "coll_vx", "coll_vy", "coll_vz",
    "coll_omegax", "coll_omegay", "coll_omegaz", "coll_theta",
    "coll_quat0", "coll_quatx", "coll_quaty", "coll_quatz",
    "fx_colloid", "fy_colloid", "fz_colloid", "tx_colloid", "ty_colloid", "tz_colloid"};

namespace {

std::string dir_of(const std::string &path) {
  size_t k = path.find_last_of('/');
  if (k == std::string::npos) return ".";
  if (k == 0) return "/";
  return path.substr(0, k);
}

std::string join(const std::string &dir, const std::string &name) {
  if (dir.empty()) return name;
  if (dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

}  // namespace

// ---------------------------------------------------------------------------
// read_input
// ---------------------------------------------------------------------------
int read_input(HostState &s, const std::string &path, std::string &err) {
  sph_params_t &p = s.p;
  p = sph_params_t{};
  // Defaults of the driver's variables (src/main.cu:48-84); the 2D forms of N, L
  // and ext_force leave the third entry at these values.
  int N[3] = {20, 20, 20};
  double L[3] = {1.0, 1.0, 1.0};
  double ext_force[3] = {0.0, 0.0, 0.0};
  int dim = 2, Nsteps = 1000, ext_force_type = 0;
  double dt = 0.01, t0 = 0.0, overlap = 4.0, rho = 1.0, c = 1.0, P0 = 1.0, eta = 1.0,
         zeta = 0.0, np_ = 1.0, kT = 1.0, tau = 1.0;
  int freq_micro = 1000, freq_macro = 100, freq_walls = 1000, freq_colloids = 1000;
  int wall = 0, N_coll = 0, coll_move = 0, new_sim = 0;
  double Vwall_top = 0.0, Vwall_bottom = 0.0, coll_R = 1.0, coll_rho = 1.0;
  double cuton = 0.001, F0_rep = 1.0, tau_rep = 0.01, r0_mag = 0.1, F0_mag = 1.0,
         omega_mag = 0.01;
  std::vector<double> coll_data;  // 3 per colloid
  int coll_x_count = 0;

  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) {
    err = "read_input: cannot open input file '" + path + "'";
    return SPH_ERR_IO;
  }

  enum { K_N, K_L, K_DIM, K_DT, K_T0, K_NSTEPS, K_OVERLAP, K_RHO, K_C, K_P0, K_ETA,
         K_ZETA, K_NP, K_KT, K_TAU, K_EXTT, K_EXTF, K_FMICRO, K_FMACRO, K_FWALLS,
         K_FCOLL, K_WALL, K_VTOP, K_VBOT, K_NCOLL, K_CR, K_CRHO, K_CMOVE, K_CUTON,
         K_F0REP, K_TAUREP, K_RMAG, K_F0MAG, K_OMAG, K_NEWSIM, K_NUM };
  static const char *const names[K_NUM] = {
      "N", "L", "dim", "dt", "t0", "Nsteps", "overlap", "rho", "c", "P0", "eta",
      "zeta", "np", "kT", "tau", "ext_force_type", "ext_force", "freq_micro",
      "freq_macro", "freq_walls", "freq_colloids", "wall", "Vwall_top",
      "Vwall_bottom", "N_colloids", "coll_R", "coll_rho", "coll_move",
      "coll_repulsion_cuton", "F0_repulsion", "tau_repulsion", "cutoff_magnetic",
      "F0_magnetic", "omega_magnetic", "new_sim"};
  bool seen[K_NUM] = {false};
  int dim_N = 0, dim_L = 0, dim_ext = 0;

  char line[256];  // MAX_LINE of the reference
  while (std::fgets(line, sizeof(line), f)) {
    char *start = line;
    while (*start == ' ' || *start == '\t') start++;
    char *end = start + std::strlen(start) - 1;
    while (end > start && (*end == ' ' || *end == '\t' || *end == '\n')) *end-- = '\0';
    char var[256];
    double v1 = 0, v2 = 0, v3 = 0;
    int count = std::sscanf(start, "%255s %lf %lf %lf", var, &v1, &v2, &v3);
    if (count < 1) continue;
    auto is = [&](const char *k) { return std::strcmp(var, k) == 0; };
    if (is("N") && count == 4) {
      N[0] = (int)v1; N[1] = (int)v2; N[2] = (int)v3; dim_N = 3; seen[K_N] = true;
    } else if (is("N") && count == 3) {
      N[0] = (int)v1; N[1] = (int)v2; dim_N = 2; seen[K_N] = true;
    } else if (is("L") && count == 4) {
      L[0] = v1; L[1] = v2; L[2] = v3; dim_L = 3; seen[K_L] = true;
    } else if (is("L") && count == 3) {
      L[0] = v1; L[1] = v2; dim_L = 2; seen[K_L] = true;
    } else if (is("ext_force") && count == 4) {
      ext_force[0] = v1; ext_force[1] = v2; ext_force[2] = v3; dim_ext = 3;
      seen[K_EXTF] = true;
    } else if (is("ext_force") && count == 3) {
      ext_force[0] = v1; ext_force[1] = v2; dim_ext = 2; seen[K_EXTF] = true;
    } else if (is("coll_x")) {
      if (count >= 3 && N_coll > 0) {
        if (coll_x_count >= N_coll) {
          std::fclose(f);
          err = "system read input error: different coll_x entries than N_colloids";
          return SPH_ERR_PARSE;
        }
        // Values are the space-separated tokens after the key (strtok on ' ').
        double vals[3] = {0, 0, 0};
        int nv = 0;
        for (char *tok = std::strtok(start, " "); tok; tok = std::strtok(nullptr, " ")) {
          if (std::strcmp(tok, "coll_x") != 0) {
            if (nv < 3) vals[nv] = std::atof(tok);
            nv++;
          }
        }
        if (nv != dim) {
          std::fclose(f);
          err = "system_read_input error: the number of dimensions of coll_x is not "
                "compatible with the number of dimensions of the system";
          return SPH_ERR_PARSE;
        }
        coll_data.push_back(vals[0]);
        coll_data.push_back(vals[1]);
        coll_data.push_back(vals[2]);
        coll_x_count++;
      }
    } else if (count >= 2) {
      int k = -1;
      for (int q = 0; q < K_NUM; ++q)
        if (q != K_N && q != K_L && q != K_EXTF && is(names[q])) { k = q; break; }
      switch (k) {
        case K_DIM: dim = (int)v1; break;
        case K_DT: dt = v1; break;
        case K_T0: t0 = v1; break;
        case K_NSTEPS: Nsteps = (int)v1; break;
        case K_OVERLAP: overlap = v1; break;
        case K_RHO: rho = v1; break;
        case K_C: c = v1; break;
        case K_P0: P0 = v1; break;
        case K_ETA: eta = v1; break;
        case K_ZETA: zeta = v1; break;
        case K_NP: np_ = v1; break;
        case K_KT: kT = v1; break;
        case K_TAU: tau = v1; break;
        case K_EXTT: ext_force_type = (int)v1; break;
        case K_FMICRO: freq_micro = (int)v1; break;
        case K_FMACRO: freq_macro = (int)v1; break;
        case K_FWALLS: freq_walls = (int)v1; break;
        case K_FCOLL: freq_colloids = (int)v1; break;
        case K_WALL: wall = (int)v1; break;
        case K_VTOP: Vwall_top = v1; break;
        case K_VBOT: Vwall_bottom = v1; break;
        case K_NCOLL: N_coll = (int)v1; break;
        case K_CR: coll_R = v1; break;
        case K_CRHO: coll_rho = v1; break;
        case K_CMOVE: coll_move = (int)v1; break;
        case K_CUTON: cuton = v1; break;
        case K_F0REP: F0_rep = v1; break;
        case K_TAUREP: tau_rep = v1; break;
        case K_RMAG: r0_mag = v1; break;
        case K_F0MAG: F0_mag = v1; break;
        case K_OMAG: omega_mag = v1; break;
        case K_NEWSIM: new_sim = (int)v1; break;
        default: break;
      }
      if (k >= 0) seen[k] = true;
    }
  }
  std::fclose(f);

  for (int q = 0; q < K_NUM; ++q) {
    if (!seen[q]) {
      err = std::string("system_read_input error: ") + names[q] + " was not read";
      return SPH_ERR_PARSE;
    }
  }
  if (coll_x_count != N_coll) {
    err = "system read input error: different number of coll_x variables than N_colloids";
    return SPH_ERR_PARSE;
  }
  if (dim != dim_N) {
    err = "system_read_input error: the number of dimensions of the system is not "
          "compatible with the number of dimensions of N";
    return SPH_ERR_PARSE;
  }
  if (dim != dim_L) {
    err = "system_read_input error: the number of dimensions of the system is not "
          "compatible with the number of dimensions of L";
    return SPH_ERR_PARSE;
  }
  if (ext_force_type != 0 && dim != dim_ext) {
    err = "system_read_input error: the number of dimensions of the system is not "
          "compatible with the number of dimensions of ext_force";
    return SPH_ERR_PARSE;
  }
  if (ext_force_type < 0 || ext_force_type > 3) {
    err = "system_read_input error: the option ext_force_type = " +
          std::to_string(ext_force_type) + " is not available.";
    return SPH_ERR_PARSE;
  }
  if (N_coll > kMaxColloids) {
    err = "System read input error: N_colloids is too high";
    return SPH_ERR_PARSE;
  }
  if (dim != 2 && dim != 3) {
    err = "system_read_input error: dim must be 2 or 3";
    return SPH_ERR_PARSE;
  }
  // Colloids inside the box and not overlapping (checked for N_colloids >= 2 only,
  // as in the reference).
  if (N_coll >= 2) {
    for (int i = 0; i < N_coll; ++i) {
      const double *ci = &coll_data[3 * i];
      if (ci[0] < 0 || ci[0] > L[0] || ci[1] < 0 || ci[1] > L[1] ||
          (dim == 3 && (ci[2] < 0 || ci[2] > L[2]))) {
        err = "System read input error: Colloid is out of the simulation box";
        return SPH_ERR_GEOMETRY;
      }
    }
    for (int i = 0; i < N_coll; ++i)
      for (int j = i + 1; j < N_coll; ++j) {
        double d[3];
        for (int a = 0; a < dim; ++a) d[a] = coll_data[3 * i + a] - coll_data[3 * j + a];
        for (int a = 0; a < dim; ++a) {
          if (a == 1 && wall != 0) continue;
          if (d[a] > 0.5 * L[a]) d[a] -= L[a];
          if (d[a] < -0.5 * L[a]) d[a] += L[a];
        }
        double r2 = d[0] * d[0] + d[1] * d[1] + (dim == 3 ? d[2] * d[2] : 0.0);
        if (std::sqrt(r2) < 2 * coll_R) {
          err = "System read input error: Overlapping colloids";
          return SPH_ERR_GEOMETRY;
        }
      }
  }

  p.dim = dim;
  for (int a = 0; a < 3; ++a) {
    p.Nxyz[a] = N[a];
    p.L[a] = L[a];
    p.ext_force[a] = ext_force[a];
  }
  p.dt = dt; p.t0 = t0; p.Nsteps = Nsteps; p.overlap = overlap; p.rho = rho; p.c = c;
  p.P0 = P0; p.eta = eta; p.zeta = zeta; p.np = np_; p.kT = kT; p.tau = tau;
  p.ext_force_type = ext_force_type;
  p.freq_micro = freq_micro; p.freq_macro = freq_macro; p.freq_walls = freq_walls;
  p.freq_colloids = freq_colloids;
  p.wall = wall; p.V_top = Vwall_top; p.V_bottom = Vwall_bottom;
  p.N_colloids = N_coll; p.coll_R = coll_R; p.coll_rho = coll_rho; p.coll_move = coll_move;
  p.coll_rep_cuton = cuton; p.F0_rep = F0_rep; p.tau_rep = tau_rep; p.r0_magnet = r0_mag;
  p.F0_magnet = F0_mag; p.omega_magnet = omega_mag; p.new_sim = new_sim;
  s.coll_x_data = coll_data;
  s.input_dir = dir_of(path);
  s.input_read = true;
  s.initialized = false;
  return SPH_OK;
}

// ---------------------------------------------------------------------------
// initialize
// ---------------------------------------------------------------------------
namespace {

void resize_particles(HostState &s, int n) {
  for (int k = 0; k < F_NUM; ++k) s.pf[k].assign(n, 0.0);
  s.type.assign(n, 0);
}

// Lattice of fluid (and wall) particles plus colloid shells
// (src/system_initialize_particles_new_sim.cpp:16-253).
int particles_new_sim(HostState &s, const double dx[3], std::string &err) {
  sph_params_t &p = s.p;
  const int dim = p.dim;
  double *x = s.pf[F_X].data(), *y = s.pf[F_Y].data(), *z = s.pf[F_Z].data();
  const double *cx = s.cf[C_X].data(), *cy = s.cf[C_Y].data(), *cz = s.cf[C_Z].data();
  const double R = p.coll_R, L0 = p.L[0], L1 = p.L[1], L2 = p.L[2];
  const int Nc = p.N_colloids;
  int counter = 0;

  auto inside_colloid = [&](double xi, double yi, double zi) {
    for (int l = 0; l < Nc; ++l) {
      double xij = xi - cx[l], yij = yi - cy[l];
      if (xij > 0.5 * L0) xij -= L0;
      if (xij < -0.5 * L0) xij += L0;
      if (p.wall == 0) {
        if (yij > 0.5 * L1) yij -= L1;
        if (yij < -0.5 * L1) yij += L1;
      }
      double r2 = xij * xij + yij * yij;
      if (dim == 3) {
        double zij = zi - cz[l];
        if (zij > 0.5 * L2) zij -= L2;
        if (zij < -0.5 * L2) zij += L2;
        r2 = xij * xij + yij * yij + zij * zij;
      }
      if (r2 <= R * R) return true;
    }
    return false;
  };

  if (dim == 3) {
    for (int i = 0; i < p.Nxyz[0]; ++i)
      for (int j = 0; j < p.Nxyz[1]; ++j)
        for (int k = 0; k < p.Nxyz[2]; ++k) {
          double xi = dx[0] / 2.0 + i * dx[0];
          double yi = dx[1] / 2.0 + j * dx[1];
          if (p.wall) yi = yi - p.wall_width;
          double zi = dx[2] / 2.0 + k * dx[2];
          x[counter] = xi; y[counter] = yi; z[counter] = zi;
          if (!inside_colloid(xi, yi, zi)) counter++;
        }
  } else {
    for (int i = 0; i < p.Nxyz[0]; ++i)
      for (int j = 0; j < p.Nxyz[1]; ++j) {
        double xi = dx[0] / 2.0 + i * dx[0];
        double yi = dx[1] / 2.0 + j * dx[1];
        if (p.wall) yi = yi - p.wall_width;
        x[counter] = xi; y[counter] = yi; z[counter] = 0.0;
        if (!inside_colloid(xi, yi, 0.0)) counter++;
      }
  }
  int N = counter;
  for (int i = 0; i < N; ++i) {
    if (p.wall == 0) s.type[i] = 0;
    else if (y[i] < p.y_bottom) s.type[i] = 1;
    else if (y[i] > p.y_top) s.type[i] = 2;
    else s.type[i] = 0;
  }
  double mass;
  if (dim == 2)
    mass = p.rho * p.L[0] * (p.L[1] + 2.0 * p.wall_width) / (p.Nxyz[0] * p.Nxyz[1]);
  else
    mass = p.rho * p.L[0] * (p.L[1] + 2.0 * p.wall_width) * p.L[2] /
           (p.Nxyz[0] * p.Nxyz[1] * p.Nxyz[2]);
  for (int i = 0; i < N; ++i) {
    s.pf[F_VX][i] = s.pf[F_VY][i] = s.pf[F_VZ][i] = 0.0;
    s.pf[F_MASS][i] = mass;
    s.pf[F_CXX][i] = 1.0; s.pf[F_CXY][i] = 0.0; s.pf[F_CYY][i] = 1.0;
    s.pf[F_CXZ][i] = 0.0; s.pf[F_CYZ][i] = 0.0; s.pf[F_CZZ][i] = 1.0;
  }

  auto append = [&](double xi, double yi, double zi, int t) -> bool {
    if (N + 1 > p.Nmax) {
      err = "system initialize particles new sim error: N > Nmax";
      return false;
    }
    x[N] = xi; y[N] = yi; z[N] = zi; s.type[N] = t;
    // mass, v, c and F of shell particles stay 0 (zero-filled arrays).
    N++;
    return true;
  };
  auto wrap_xy = [&](double &xi, double &yi) {
    if (xi < 0) xi = xi + L0;
    if (p.wall == 0 && yi < 0) yi = yi + L1;
    if (xi > L0) xi = xi - L0;
    if (p.wall == 0 && yi > L1) yi = yi - L1;
  };

  if (dim == 2) {
    int Nlayers = (int)std::round(R / dx[0]);
    for (int i = 0; i < Nc; ++i) {
      double r = 0.5 * dx[0];
      for (int j = 0; j < Nlayers; ++j) {
        if (r < R && r > R - p.wall_width) {
          double d_phi = dx[0] / r;
          int Npart_phi = (int)std::round(2.0 * M_PI / d_phi);
          d_phi = 2.0 * M_PI / Npart_phi;
          double phi = 0.0;
          for (int k = 0; k < Npart_phi; ++k) {
            double xi = cx[i] + r * std::cos(phi);
            double yi = cy[i] + r * std::sin(phi);
            wrap_xy(xi, yi);
            if (!append(xi, yi, 0.0, 3 + i)) return SPH_ERR_GEOMETRY;
            phi = phi + d_phi;
          }
        }
        r = r + dx[0];
      }
    }
  } else {
    int Nlayers = (int)std::round(R / dx[0]);
    for (int i = 0; i < Nc; ++i) {
      double r = 0.5 * dx[0];
      for (int j = 0; j < Nlayers; ++j) {
        if (r < R && r > R - p.wall_width) {
          auto wrap_z = [&](double zi) {
            if (zi < 0) zi = zi + L2;
            if (zi > L2) zi = zi - L2;
            return zi;
          };
          // North pole.
          if (!append(cx[i], cy[i], wrap_z(cz[i] + r), 3 + i)) return SPH_ERR_GEOMETRY;
          double d_theta = dx[0] / r;
          int Npart_theta = (int)std::round(M_PI / d_theta);
          d_theta = M_PI / Npart_theta;
          double theta = 0.0;
          for (int k = 0; k < Npart_theta; ++k) {
            double d_phi = dx[0] / (r * std::sin(theta));
            int Npart_phi = (int)std::round(2.0 * M_PI / d_phi);
            d_phi = 2.0 * M_PI / Npart_phi;
            double phi = 0.0;
            for (int l = 0; l < Npart_phi; ++l) {
              double xi = cx[i] + r * std::sin(theta) * std::cos(phi);
              double yi = cy[i] + r * std::sin(theta) * std::sin(phi);
              double zi = cz[i] + r * std::cos(theta);
              wrap_xy(xi, yi);
              if (zi < 0) zi = zi + L2;
              if (zi > L2) zi = zi - L2;
              if (!append(xi, yi, zi, 3 + i)) return SPH_ERR_GEOMETRY;
              phi = phi + d_phi;
            }
            theta = theta + d_theta;
          }
          // South pole.
          if (!append(cx[i], cy[i], wrap_z(cz[i] - r), 3 + i)) return SPH_ERR_GEOMETRY;
        }
        r = r + dx[0];
      }
    }
  }
  p.N = N;
  return SPH_OK;
}

// Restart file: columns x y [z] vx vy [vz] mass type cxx cxy cyy [cxz cyz czz],
// parsed as float (src/system_initialize_particles_old_sim.cpp:31-63).
int particles_old_sim(HostState &s, const std::string &path, std::string &err) {
  std::ifstream file(path);
  if (!file.is_open()) {
    err = "initialize: cannot open restart file '" + path + "'";
    return SPH_ERR_IO;
  }
  const int dim = s.p.dim;
  const int ncol = dim == 2 ? 9 : 14;
  std::vector<float> rows;
  float v[14];
  while (true) {
    bool ok = true;
    for (int k = 0; k < ncol; ++k)
      if (!(file >> v[k])) { ok = false; break; }
    if (!ok) break;
    rows.insert(rows.end(), v, v + ncol);
  }
  int N = (int)(rows.size() / ncol);
  s.p.N = N;
  s.p.Nmax = N;
  resize_particles(s, N);
  for (int i = 0; i < N; ++i) {
    const float *r = &rows[(size_t)i * ncol];
    if (dim == 2) {
      s.pf[F_X][i] = r[0]; s.pf[F_Y][i] = r[1]; s.pf[F_VX][i] = r[2]; s.pf[F_VY][i] = r[3];
      s.pf[F_MASS][i] = r[4]; s.type[i] = (int)r[5];
      s.pf[F_CXX][i] = r[6]; s.pf[F_CXY][i] = r[7]; s.pf[F_CYY][i] = r[8];
    } else {
      s.pf[F_X][i] = r[0]; s.pf[F_Y][i] = r[1]; s.pf[F_Z][i] = r[2];
      s.pf[F_VX][i] = r[3]; s.pf[F_VY][i] = r[4]; s.pf[F_VZ][i] = r[5];
      s.pf[F_MASS][i] = r[6]; s.type[i] = (int)r[7];
      s.pf[F_CXX][i] = r[8]; s.pf[F_CXY][i] = r[9]; s.pf[F_CYY][i] = r[10];
      s.pf[F_CXZ][i] = r[11]; s.pf[F_CYZ][i] = r[12]; s.pf[F_CZZ][i] = r[13];
    }
  }
  (void)err;
  return SPH_OK;
}

}  // namespace

int initialize(HostState &s, const std::string &restart_path, std::string &err) {
  if (!s.input_read) {
    err = "initialize: read_input has not succeeded";
    return SPH_ERR_STATE;
  }
  sph_params_t &p = s.p;
  const int dim = p.dim;
  const int N_coll = p.N_colloids;

  // ---- constructor (src/system_constructor.cpp:15-108) ----
  p.Nwalls = p.wall ? 2 : 0;
  if (p.new_sim == 0) {
    int Ny = p.wall == 1 ? p.Nxyz[1] + 2 * (int)std::ceil(p.overlap) : p.Nxyz[1];
    int N = dim == 2 ? p.Nxyz[0] * Ny : p.Nxyz[0] * Ny * p.Nxyz[2];
    if (N_coll > 0) N = (int)(N * 1.01);
    p.N = N;
    p.Nmax = N;
    resize_particles(s, N);
  }
  for (int k = 0; k < C_NUM; ++k) s.cf[k].assign(N_coll, 0.0);

  // ---- derived constants (src/system_initialize.cpp:47-143) ----
  p.nkT = p.np * p.kT;
  p.rcut = p.L[0] / p.Nxyz[0] * p.overlap;
  p.rcutsq = p.rcut * p.rcut;
  p.rcut_inv = 1.0 / p.rcut;
  double dx[3] = {0, 0, 0};
  for (int i = 0; i < dim; ++i) dx[i] = p.L[i] / p.Nxyz[i];
  p.ky = 2.0 * M_PI / p.L[1];
  const double eps = 1.0e-10;
  if (std::fabs(p.overlap - std::floor(p.overlap)) < eps)
    p.wall_width = std::floor(p.overlap) * dx[1];
  else
    p.wall_width = (std::floor(p.overlap) + 1) * dx[1];
  p.y_bottom = 0.0;
  p.y_top = 0.0;
  if (p.wall == 1) {
    p.Nxyz[1] = p.Nxyz[1] + 2 * (int)std::ceil(p.overlap);
    p.y_bottom = 0.0;
    p.y_top = p.L[1];
  }
  if (dim == 2)
    p.cw = 63.0 / (478.0 * M_PI * p.rcut * p.rcut);
  else
    p.cw = 27.0 / (120.0 * M_PI * p.rcut * p.rcut * p.rcut);
  p.c_gradw = -15.0 * p.cw / p.rcut;
  p.csq = p.c * p.c;
  p.rho0 = p.P0 * p.rho;
  p.a = 5.0 * p.eta / 3.0 - p.zeta;
  p.b = (dim == 3 ? 5.0 : 4.0) * (p.zeta + p.eta / 3.0);
  p.beta_max = 1.5;

  p.Ncells[0] = (int)(p.L[0] / p.rcut);
  p.cell_size[0] = p.L[0] / p.Ncells[0];
  if (p.wall == 1) {
    p.Ncells[1] = (int)((p.L[1] + 2 * p.wall_width) / p.rcut);
    p.cell_size[1] = (p.L[1] + 2 * p.wall_width) / p.Ncells[1];
  } else {
    p.Ncells[1] = (int)(p.L[1] / p.rcut);
    p.cell_size[1] = p.L[1] / p.Ncells[1];
  }
  if (dim == 3) {
    p.Ncells[2] = (int)(p.L[2] / p.rcut);
    p.cell_size[2] = p.L[2] / p.Ncells[2];
  } else {
    p.Ncells[2] = 1;
    p.cell_size[2] = 1;
  }
  p.Ntotal_cells = p.Ncells[0] * p.Ncells[1] * p.Ncells[2];
  if (p.Ncells[0] < 3 || p.Ncells[1] < 3 || (dim == 3 && p.Ncells[2] < 3)) {
    err = "System initialize error: neighbour cells are too big";
    return SPH_ERR_GEOMETRY;
  }

  // ---- colloids (src/system_initialize.cpp:154-262) ----
  p.coll_rep_cutoff = 5.0 * p.coll_R / p.tau_rep;
  p.r0_magnet_sq = p.r0_magnet * p.r0_magnet;
  if (dim == 2) {
    p.coll_mass = M_PI * p.coll_R * p.coll_R * p.coll_rho;
    p.coll_I = 0.5 * p.coll_mass * p.coll_R * p.coll_R;
  } else {
    p.coll_mass = 4.0 / 3.0 * M_PI * p.coll_R * p.coll_R * p.coll_R * p.coll_rho;
    p.coll_I = 2.0 / 5.0 * p.coll_mass * p.coll_R * p.coll_R;
  }
  for (int i = 0; i < N_coll; ++i) {
    s.cf[C_X][i] = s.coll_x_data[3 * i + 0];
    s.cf[C_Y][i] = s.coll_x_data[3 * i + 1];
    s.cf[C_Z][i] = dim == 3 ? s.coll_x_data[3 * i + 2] : 0.0;
    s.cf[C_Q0][i] = 1.0;
    if (p.wall) {
      if (s.cf[C_Y][i] - p.coll_R < p.y_bottom) {
        err = "System initialize error: colloid " + std::to_string(i) +
              " is overlapping with the bottom wall";
        return SPH_ERR_GEOMETRY;
      }
      if (s.cf[C_Y][i] + p.coll_R > p.y_top) {
        err = "System initialize error: colloid " + std::to_string(i) +
              " is overlapping with the top wall";
        return SPH_ERR_GEOMETRY;
      }
    }
  }
  p.Ncells_colloids[0] = p.Ncells_colloids[1] = p.Ncells_colloids[2] = 0;
  p.Ntotal_cells_colloids = 0;
  if (N_coll > 0) {
    p.rcuton_coll_rep = 2.0 * p.coll_R + p.coll_rep_cuton;
    p.rcuton_coll_rep_sq = p.rcuton_coll_rep * p.rcuton_coll_rep;
    p.rcutoff_coll_rep = 2.0 * p.coll_R + p.coll_rep_cutoff;
    p.rcutoff_coll_rep_sq = p.rcutoff_coll_rep * p.rcutoff_coll_rep;
    p.rcutoff_coll = p.rcutoff_coll_rep > p.r0_magnet ? p.rcutoff_coll_rep : p.r0_magnet;
    p.rcutoff_coll_sq = p.rcutoff_coll * p.rcutoff_coll;
    p.Ncells_colloids[0] = (int)(p.L[0] / p.rcutoff_coll);
    p.cell_colloids_size[0] = p.L[0] / p.Ncells_colloids[0];
    p.Ncells_colloids[1] = (int)(p.L[1] / p.rcutoff_coll);
    p.cell_colloids_size[1] = p.L[1] / p.Ncells_colloids[1];
    if (dim == 3) {
      p.Ncells_colloids[2] = (int)(p.L[2] / p.rcutoff_coll);
      p.cell_colloids_size[2] = p.L[2] / p.Ncells_colloids[2];
    } else {
      p.Ncells_colloids[2] = 1;
      p.cell_colloids_size[2] = 1;
    }
    p.Ntotal_cells_colloids =
        p.Ncells_colloids[0] * p.Ncells_colloids[1] * p.Ncells_colloids[2];
    if (p.Ncells_colloids[0] < 3 || p.Ncells_colloids[1] < 3 ||
        (dim == 3 && p.Ncells_colloids[2] < 3)) {
      err = "System initialize error: neighbour colloids cells are too big";
      return SPH_ERR_GEOMETRY;
    }
  }

  // ---- particles ----
  int rc;
  if (p.new_sim == 0) {
    rc = particles_new_sim(s, dx, err);
  } else {
    std::string path = restart_path.empty() ? join(s.input_dir, "input_particles.in")
                                            : restart_path;
    rc = particles_old_sim(s, path, err);
  }
  if (rc != SPH_OK) return rc;
  const int N = p.N;

  // ---- wall list (src/system_initialize.cpp:276-300) ----
  s.walls_list.clear();
  s.walls_start.assign(p.Nwalls, 0);
  if (p.wall) {
    for (int w = 0; w < p.Nwalls; ++w) {
      s.walls_start[w] = (int)s.walls_list.size();
      for (int j = 0; j < N; ++j)
        if (s.type[j] == w + 1) s.walls_list.push_back(j);
    }
  }
  p.Nlist_walls = (int)s.walls_list.size();

  // ---- colloid lists and body-frame offsets (src/system_initialize.cpp:302-355) ----
  s.colloids_list.clear();
  s.colloids_start.assign(N_coll, 0);
  for (int i = 0; i < N_coll; ++i) {
    s.colloids_start[i] = (int)s.colloids_list.size();
    for (int j = 0; j < N; ++j)
      if (s.type[j] == i + 3) s.colloids_list.push_back(j);
  }
  p.Nlist_colloids = (int)s.colloids_list.size();
  const int nl = p.Nlist_colloids;
  s.x_center.assign(nl, 0.0);
  s.y_center.assign(nl, 0.0);
  s.z_center.assign(nl, 0.0);
  for (int i = 0; i < nl; ++i) {
    int part = s.colloids_list[i];
    int col = s.type[part] - 3;
    double xc = s.pf[F_X][part] - s.cf[C_X][col];
    double yc = s.pf[F_Y][part] - s.cf[C_Y][col];
    double zc = dim == 3 ? s.pf[F_Z][part] - s.cf[C_Z][col] : 0.0;
    if (xc > 0.5 * p.L[0]) xc -= p.L[0];
    if (xc < -0.5 * p.L[0]) xc += p.L[0];
    if (p.wall == 0) {
      if (yc > 0.5 * p.L[1]) yc -= p.L[1];
      if (yc < -0.5 * p.L[1]) yc += p.L[1];
    }
    if (dim == 3) {
      if (zc > 0.5 * p.L[2]) zc -= p.L[2];
      if (zc < -0.5 * p.L[2]) zc += p.L[2];
    }
    s.x_center[i] = xc;
    s.y_center[i] = yc;
    s.z_center[i] = zc;
  }
  s.initialized = true;
  return SPH_OK;
}

// ---------------------------------------------------------------------------
// writers
// ---------------------------------------------------------------------------
int write_micro(const HostState &s, int step, const std::string &dir, std::string &err) {
  std::string fn = join(dir, "micro-" + std::to_string(step) + ".dat");
  FILE *f = std::fopen(fn.c_str(), "w");
  if (!f) { err = "print particles: cannot open " + fn; return SPH_ERR_IO; }
  const auto &P = s.pf;
  for (int i = 0; i < s.p.N; ++i) {
    if (s.p.dim == 2)
      std::fprintf(f, "%d %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %d %.8lf %.8lf %.8lf \n",
                   i, P[F_X][i], P[F_Y][i], P[F_VX][i], P[F_VY][i], P[F_MASS][i],
                   P[F_DENS][i], P[F_PRESS][i], s.type[i], P[F_CXX][i], P[F_CXY][i],
                   P[F_CYY][i]);
    else
      std::fprintf(f,
                   "%d %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %d %.8lf %.8lf "
                   "%.8lf %.8lf %.8lf %.8lf\n",
                   i, P[F_X][i], P[F_Y][i], P[F_Z][i], P[F_VX][i], P[F_VY][i], P[F_VZ][i],
                   P[F_MASS][i], P[F_DENS][i], P[F_PRESS][i], s.type[i], P[F_CXX][i],
                   P[F_CXY][i], P[F_CYY][i], P[F_CXZ][i], P[F_CYZ][i], P[F_CZZ][i]);
  }
  std::fclose(f);
  return SPH_OK;
}

int write_walls(const HostState &s, int step, const std::string &dir, std::string &err) {
  std::string fn = join(dir, "walls-" + std::to_string(step) + ".dat");
  FILE *f = std::fopen(fn.c_str(), "w");
  if (!f) { err = "print walls: cannot open " + fn; return SPH_ERR_IO; }
  for (int w = 0; w < s.p.Nwalls; ++w) {
    if (s.p.dim == 2)
      std::fprintf(f, "%d %.8lf %.8lf\n", w, s.fwall[0][w], s.fwall[1][w]);
    else
      std::fprintf(f, "%d %.8lf %.8lf %.8lf\n", w, s.fwall[0][w], s.fwall[1][w],
                   s.fwall[2][w]);
  }
  std::fclose(f);
  return SPH_OK;
}

int write_colloids(const HostState &s, int step, const std::string &dir, std::string &err) {
  std::string fn = join(dir, "colloids-" + std::to_string(step) + ".dat");
  FILE *f = std::fopen(fn.c_str(), "w");
  if (!f) { err = "print colloids: cannot open " + fn; return SPH_ERR_IO; }
  const auto &C = s.cf;
  for (int i = 0; i < s.p.N_colloids; ++i) {
    if (s.p.dim == 2)
      std::fprintf(f, "%d %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf\n", i, C[C_X][i],
                   C[C_Y][i], C[C_VX][i], C[C_VY][i], C[C_OZ][i], C[C_FX][i], C[C_FY][i],
                   C[C_TZ][i]);
    else
      std::fprintf(f,
                   "%d %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf %.8lf "
                   "%.8lf %.8lf %.8lf %.8lf\n",
                   i, C[C_X][i], C[C_Y][i], C[C_Z][i], C[C_VX][i], C[C_VY][i], C[C_VZ][i],
                   C[C_OX][i], C[C_OY][i], C[C_OZ][i], C[C_FX][i], C[C_FY][i], C[C_FZ][i],
                   C[C_TX][i], C[C_TY][i], C[C_TZ][i]);
  }
  std::fclose(f);
  return SPH_OK;
}

int append_macro(const HostState &s, int step, double ke, const std::string &dir,
                 std::string &err) {
  (void)s;
  std::string fn = join(dir, "macro.dat");
  FILE *f = std::fopen(fn.c_str(), step == 0 ? "w" : "a");
  if (!f) { err = "print macro vars: cannot open " + fn; return SPH_ERR_IO; }
  std::fprintf(f, "%d %f\n", step, ke);
  std::fclose(f);
  return SPH_OK;
}

int write_restart(const HostState &s, const std::string &path, std::string &err) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) { err = "write restart: cannot open " + path; return SPH_ERR_IO; }
  // The reader parses floats, so values are rounded to float first and printed
  // with 9 significant digits, which round-trips every float exactly.
  auto P = [&](int k, int i) { return (double)(float)s.pf[k][i]; };
  for (int i = 0; i < s.p.N; ++i) {
    if (s.p.dim == 2)
      std::fprintf(f, "%.9g %.9g %.9g %.9g %.9g %d %.9g %.9g %.9g\n", P(F_X, i), P(F_Y, i),
                   P(F_VX, i), P(F_VY, i), P(F_MASS, i), s.type[i], P(F_CXX, i),
                   P(F_CXY, i), P(F_CYY, i));
    else
      std::fprintf(f,
                   "%.9g %.9g %.9g %.9g %.9g %.9g %.9g %d %.9g %.9g %.9g %.9g %.9g %.9g\n",
                   P(F_X, i), P(F_Y, i), P(F_Z, i), P(F_VX, i), P(F_VY, i), P(F_VZ, i),
                   P(F_MASS, i), s.type[i], P(F_CXX, i), P(F_CXY, i), P(F_CYY, i),
                   P(F_CXZ, i), P(F_CYZ, i), P(F_CZZ, i));
  }
  std::fclose(f);
  return SPH_OK;
}

// ---------------------------------------------------------------------------
// Binary checkpoint: the whole host state at full precision.  Layout (native
// little-endian): the 8-byte magic, int32 precision, int32 sizeof(sph_params_t),
// the parameter block, int64 step, then arrays as records (char name[16],
// int32 kind 0 = double / 1 = int32, int64 count, the values), closed by a
// record named "end" with no values.
namespace {

constexpr char kCkptMagic[8] = {'S', 'P', 'H', 'C', 'K', 'P', 'T', '\x01'};

struct CkptArray {
  std::string name;
  std::vector<double> *d = nullptr;
  std::vector<int32_t> *i = nullptr;
  int64_t count;  // entries written and expected: fixed by the parameters
  int64_t alloc;  // storage the state keeps (Nmax for the per-particle arrays)
};

// Every array of the state with its length under the parameters p (the
// lengths initialize gives them).
std::vector<CkptArray> ckpt_arrays(HostState &s, const sph_params_t &p) {
  std::vector<CkptArray> a;
  const int64_t nc = p.N_colloids, nl = p.Nlist_colloids;
  for (int k = 0; k < F_NUM; ++k) a.push_back({kPFieldNames[k], &s.pf[k], nullptr, p.N, p.Nmax});
  a.push_back({"type", nullptr, &s.type, p.N, p.Nmax});
  for (int k = 0; k < C_NUM; ++k) a.push_back({kCFieldNames[k], &s.cf[k], nullptr, nc, nc});
  a.push_back({"coll_x_data", &s.coll_x_data, nullptr, 3 * nc, 3 * nc});
  a.push_back({"walls_list", nullptr, &s.walls_list, p.Nlist_walls, p.Nlist_walls});
  a.push_back({"walls_start", nullptr, &s.walls_start, p.Nwalls, p.Nwalls});
  a.push_back({"colloids_list", nullptr, &s.colloids_list, nl, nl});
  a.push_back({"colloids_start", nullptr, &s.colloids_start, nc, nc});
  a.push_back({"x_center", &s.x_center, nullptr, nl, nl});
  a.push_back({"y_center", &s.y_center, nullptr, nl, nl});
  a.push_back({"z_center", &s.z_center, nullptr, nl, nl});
  return a;
}

bool put(FILE *f, const void *p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
bool get(FILE *f, void *p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

bool put_record(FILE *f, const std::string &name, int32_t kind, int64_t count, const void *v) {
  char nm[16] = {0};
  std::strncpy(nm, name.c_str(), sizeof(nm) - 1);
  const size_t bytes = (size_t)count * (kind == 0 ? sizeof(double) : sizeof(int32_t));
  return put(f, nm, sizeof(nm)) && put(f, &kind, sizeof(kind)) && put(f, &count, sizeof(count)) &&
         put(f, v, bytes);
}

}  // namespace

int write_checkpoint(const HostState &cs, int precision, int64_t step, const std::string &path,
                     std::string &err) {
  HostState &s = const_cast<HostState &>(cs);  // (arrays are only read)
  if (!s.initialized) { err = "checkpoint: the system is not initialized"; return SPH_ERR_STATE; }
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { err = "checkpoint: cannot open " + path; return SPH_ERR_IO; }
  const int32_t prec = precision, psize = (int32_t)sizeof(sph_params_t);
  bool ok = put(f, kCkptMagic, sizeof(kCkptMagic)) && put(f, &prec, sizeof(prec)) &&
            put(f, &psize, sizeof(psize)) && put(f, &s.p, sizeof(s.p)) &&
            put(f, &step, sizeof(step));
  double fw[6];
  for (int a = 0; a < 3; ++a)
    for (int w = 0; w < 2; ++w) fw[2 * a + w] = s.fwall[a][w];
  ok = ok && put_record(f, "fwall", 0, 6, fw);
  for (const CkptArray &a : ckpt_arrays(s, s.p)) {
    const size_t have = a.d ? a.d->size() : a.i->size();
    if ((int64_t)have < a.count) {
      std::fclose(f);
      err = "checkpoint: array " + a.name + " is shorter than the state";
      return SPH_ERR_STATE;
    }
    ok = ok && put_record(f, a.name, a.d ? 0 : 1, a.count,
                          a.d ? (const void *)a.d->data() : (const void *)a.i->data());
  }
  ok = ok && put_record(f, "end", 0, 0, nullptr);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) { err = "checkpoint: write failed on " + path; return SPH_ERR_IO; }
  return SPH_OK;
}

int read_checkpoint(HostState &out, int precision, int64_t *step, const std::string &path,
                    std::string &err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) { err = "checkpoint: cannot open " + path; return SPH_ERR_IO; }
  auto fail = [&](const std::string &m) {
    std::fclose(f);
    err = "checkpoint " + path + ": " + m;
    return SPH_ERR_IO;
  };
  char magic[8];
  int32_t prec = -1, psize = 0;
  HostState s;
  int64_t st = 0;
  if (!get(f, magic, sizeof(magic)) || std::memcmp(magic, kCkptMagic, sizeof(magic)))
    return fail("not a checkpoint of this library");
  if (!get(f, &prec, sizeof(prec)) || !get(f, &psize, sizeof(psize)) ||
      psize != (int32_t)sizeof(sph_params_t) || !get(f, &s.p, sizeof(s.p)) ||
      !get(f, &st, sizeof(st)))
    return fail("truncated header");
  if (prec != precision) {
    std::fclose(f);
    err = "checkpoint " + path + " holds the other precision form";
    return SPH_ERR_ARG;
  }
  const sph_params_t &p = s.p;
  if (p.N < 0 || p.N > p.Nmax || p.N_colloids < 0 || p.Nlist_colloids < 0 ||
      p.Nlist_walls < 0 || p.Nwalls < 0 || p.Nwalls > 2 || (p.dim != 2 && p.dim != 3))
    return fail("inconsistent parameters");
  std::vector<CkptArray> arrays = ckpt_arrays(s, p);
  for (CkptArray &a : arrays) {
    if (a.d) a.d->assign(a.alloc, 0.0);
    else a.i->assign(a.alloc, 0);
  }
  std::vector<bool> seen(arrays.size(), false);
  bool fwall = false;
  for (;;) {
    char nm[16];
    int32_t kind;
    int64_t count;
    if (!get(f, nm, sizeof(nm)) || !get(f, &kind, sizeof(kind)) || !get(f, &count, sizeof(count)))
      return fail("truncated record");
    nm[15] = 0;
    const std::string name = nm;
    if (name == "end") break;
    if ((kind != 0 && kind != 1) || count < 0 || count > (int64_t)1 << 32)
      return fail("bad record " + name);
    if (name == "fwall") {
      double fw[6];
      if (kind != 0 || count != 6 || !get(f, fw, sizeof(fw))) return fail("bad record fwall");
      for (int a = 0; a < 3; ++a)
        for (int w = 0; w < 2; ++w) s.fwall[a][w] = fw[2 * a + w];
      fwall = true;
      continue;
    }
    size_t k = 0;
    while (k < arrays.size() && arrays[k].name != name) ++k;
    if (k == arrays.size()) return fail("unknown record " + name);
    CkptArray &a = arrays[k];
    if (seen[k] || (a.d != nullptr) != (kind == 0) || count != a.count)
      return fail("record " + name + " does not fit the parameters");
    const bool got = a.d ? get(f, a.d->data(), (size_t)count * sizeof(double))
                         : get(f, a.i->data(), (size_t)count * sizeof(int32_t));
    if (!got) return fail("truncated " + name);
    seen[k] = true;
  }
  std::fclose(f);
  for (size_t k = 0; k < arrays.size(); ++k)
    if (!seen[k]) { err = "checkpoint " + path + ": no record " + arrays[k].name; return SPH_ERR_IO; }
  if (!fwall) { err = "checkpoint " + path + ": no record fwall"; return SPH_ERR_IO; }
  s.input_read = s.initialized = true;
  s.input_dir = out.input_dir;
  out = std::move(s);
  *step = st;
  return SPH_OK;
}

}  // namespace sph

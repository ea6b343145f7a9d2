// sph_oracle.cpp - FP64 OpenMP oracle of the SPH Oldroyd-B step.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Each function names
// the reference kernel whose arithmetic it restates; loops run over particle
// IDs, neighbours come through particle_index after a stable (cell, ID) sort,
// exactly the reference's data flow.  Defects of the reference are handled as
// follows (and identically in the HIP path):
//   * empty cells get start > end every step (reference keeps stale ranges,
//     src/kernel_cell_ranges.cu:26-31 - no config here has an empty cell);
//   * 3D colloid Morris terms use a 3-vector (reference writes ri_coll[2] of a
//     2-array, src/kernel_vel_gradients.cu:463-466, 567-570);
//   * the 2D colloid torque accumulator starts from 0 (uninitialised in the
//     reference, src/kernel_collect_force_colloids.cu:41,80);
//   * a 3D colloid with |omega| == 0 is not rotated (reference divides by 0,
//     src/kernel_move_colloids_VV_part1.cu:121-125);
//   * the 3D magnetic force keeps the reference's non-unit e_ij = R r_ij/|r_ij|
//     (src/kernel_magnetic_force_colloids.cu:181-184).
#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum {
  X, Y, Z, VX, VY, VZ, MASS, DENS, PRESS, FX, FY, FZ, CXX, CXY, CYY, CXZ, CYZ, CZZ,
  DXX, DXY, DYY, DXZ, DYZ, DZZ, GXX, GXY, GXZ, GYX, GYY, GYZ, GZX, GZY, GZZ, NPF
};
const char *const kPF[NPF] = {
    "x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "fx", "fy", "fz",
    "cxx", "cxy", "cyy", "cxz", "cyz", "czz",
    "dcdt_xx", "dcdt_xy", "dcdt_yy", "dcdt_xz", "dcdt_yz", "dcdt_zz",
    "gradvxx", "gradvxy", "gradvxz", "gradvyx", "gradvyy", "gradvyz",
    "gradvzx", "gradvzy", "gradvzz"};
enum {
  CX, CY, CZ, CVX, CVY, CVZ, COX, COY, COZ, CTH, CQ0, CQX, CQY, CQZ,
  CFX, CFY, CFZ, CTX, CTY, CTZ, NCF
};
const char *const kCF[NCF] = {
    "coll_x", "coll_y", "coll_z", "coll_vx", "coll_vy", "coll_vz",
    "coll_omegax", "coll_omegay", "coll_omegaz", "coll_theta",
    "coll_quat0", "coll_quatx", "coll_quaty", "coll_quatz",
    "fx_colloid", "fy_colloid", "fz_colloid", "tx_colloid", "ty_colloid", "tz_colloid"};

}  // namespace

struct orc_system {
  sph_params_t p;
  std::vector<double> f[NPF];
  std::vector<int32_t> type;
  std::vector<double> cf[NCF];
  double fwall[3][2];
  std::vector<int32_t> walls_list, walls_start, colloids_list, colloids_start;
  std::vector<double> xc, yc, zc;
  std::vector<int32_t> particle_cell, particle_index, cell_start, cell_end;
  std::vector<int32_t> coll_cell, coll_index, coll_cell_start, coll_cell_end;
  double mag[3];
  int step;
  int slip;  // a planted mistake (orc_set_slip), 0 = none
};

namespace {

// Minimum image along one axis, the reference's "if (d > L/2) d -= L" form.
inline double mimg(double d, double L, double halfL) {
  if (d > halfL) d -= L;
  if (d < -halfL) d += L;
  return d;
}

// kernel_W (src/kernel_W.cu:13-59): quintic spline.
double W(const sph_params_t &p, double r) {
  double q3 = r * p.rcut_inv * 3.0;
  double a = 3.0 - q3, a2 = a * a, a5 = a2 * a2 * a;
  if (q3 < 1.0) {
    double b = 2.0 - q3, b2 = b * b, b5 = 6.0 * (b2 * b2 * b);
    double c = 1.0 - q3, c2 = c * c, c5 = 15.0 * (c2 * c2 * c);
    return p.cw * ((a5 - b5) + c5);
  } else if (q3 < 2.0) {
    double b = 2.0 - q3, b2 = b * b, b5 = 6.0 * (b2 * b2 * b);
    return p.cw * (a5 - b5);
  }
  return p.cw * a5;
}

// kernel_gradW (src/kernel_gradW.cu:13-53): dW/dr (<= 0).
double gradW(const sph_params_t &p, double r) {
  double q3 = 3.0 * (r * p.rcut_inv);
  double a = 3.0 - q3, a2 = a * a, a4 = a2 * a2;
  if (q3 < 1.0) {
    double b = 2.0 - q3, b2 = b * b, b4 = 6.0 * (b2 * b2);
    double c = 1.0 - q3, c2 = c * c, c4 = 15.0 * (c2 * c2);
    return p.c_gradw * ((a4 - b4) + c4);
  } else if (q3 < 2.0) {
    double b = 2.0 - q3, b2 = b * b, b4 = 6.0 * (b2 * b2);
    return p.c_gradw * (a4 - b4);
  }
  return p.c_gradw * a4;
}

// Iterates the 9/27 neighbour cells of particle i in the reference's order
// (dx outer, dy, dz inner; src/kernel_densities.cu:55-163) and calls fn(j)
// for every particle index j stored in them.
template <class Fn>
void for_each_candidate(const orc_system &s, int i, Fn fn) {
  const sph_params_t &p = s.p;
  const double *x = s.f[X].data(), *y = s.f[Y].data(), *z = s.f[Z].data();
  const int N0 = p.Ncells[0], N1 = p.Ncells[1], N2 = p.Ncells[2];
  int cxi = (int)std::floor(x[i] / p.cell_size[0]);
  int cyi = p.wall == 0 ? (int)std::floor(y[i] / p.cell_size[1])
                        : (int)std::floor((y[i] + p.wall_width) / p.cell_size[1]);
  int czi = p.dim == 3 ? (int)std::floor(z[i] / p.cell_size[2]) : 0;
  const int dzlo = p.dim == 3 ? -1 : 0, dzhi = p.dim == 3 ? 1 : 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = dzlo; dz <= dzhi; ++dz) {
        int cx = cxi + dx, cy = cyi + dy, cz = czi + dz;
        if (cx < 0) cx += N0;
        if (cx >= N0) cx -= N0;
        if (p.wall == 0) {
          if (cy < 0) cy += N1;
          if (cy >= N1) cy -= N1;
        }
        if (p.dim == 3) {
          if (cz < 0) cz += N2;
          if (cz >= N2) cz -= N2;
        }
        if (cy < 0 || cy >= N1) continue;
        int cell = p.dim == 3 ? cz * N0 * N1 + cy * N0 + cx : cy * N0 + cx;
        for (int k = s.cell_start[cell]; k <= s.cell_end[cell]; ++k) fn(s.particle_index[k]);
      }
}

// r_ij = x_i - x_j with the minimum image in x, z and (without walls) y.
inline void rij_of(const orc_system &s, int i, int j, double r[3]) {
  const sph_params_t &p = s.p;
  r[0] = mimg(s.f[X][i] - s.f[X][j], p.L[0], 0.5 * p.L[0]);
  r[1] = s.f[Y][i] - s.f[Y][j];
  if (p.wall == 0) r[1] = mimg(r[1], p.L[1], 0.5 * p.L[1]);
  r[2] = p.dim == 3 ? mimg(s.f[Z][i] - s.f[Z][j], p.L[2], 0.5 * p.L[2]) : 0.0;
}

inline double rsq(const sph_params_t &p, const double r[3]) {
  return p.dim == 3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
}

// Relative velocity with Morris boundary conditions for a pair (i, j) whose
// types differ or are both fluid (src/kernel_vel_gradients.cu:157-332 and
// 415-630; identical branches in src/kernel_forces.cu:196-419, 533-822).
void morris_vij(const orc_system &s, int i, int j, const double rij[3], double v[3]) {
  const sph_params_t &p = s.p;
  const int ti = s.type[i], tj = s.type[j];
  const double *vx = s.f[VX].data(), *vy = s.f[VY].data(), *vz = s.f[VZ].data();
  const double *y = s.f[Y].data();
  const bool d3 = p.dim == 3;
  const double vi[3] = {vx[i], vy[i], d3 ? vz[i] : 0.0};
  const double vj[3] = {vx[j], vy[j], d3 ? vz[j] : 0.0};
  double beta;
  if (ti == tj) {
    for (int a = 0; a < 3; ++a) v[a] = vi[a] - vj[a];
  } else if (ti == 0) {
    if (tj == 1 || tj == 2) {
      double di, dj, Vw;
      if (tj == 1) { di = y[i] - p.y_bottom; dj = p.y_bottom - y[j]; Vw = p.V_bottom; }
      else { di = p.y_top - y[i]; dj = y[j] - p.y_top; Vw = p.V_top; }
      beta = 1.0 + dj / di;
      if (beta > p.beta_max) beta = p.beta_max;
      if (di > 0) { v[0] = beta * (vi[0] - Vw); v[1] = beta * vi[1]; v[2] = beta * vi[2]; }
      else { v[0] = vi[0] - Vw; v[1] = vi[1]; v[2] = vi[2]; }
    } else {
      int c = tj - 3;
      double rc[3];
      rc[0] = mimg(s.f[X][i] - s.cf[CX][c], p.L[0], 0.5 * p.L[0]);
      rc[1] = y[i] - s.cf[CY][c];
      if (p.wall == 0) rc[1] = mimg(rc[1], p.L[1], 0.5 * p.L[1]);
      rc[2] = d3 ? mimg(s.f[Z][i] - s.cf[CZ][c], p.L[2], 0.5 * p.L[2]) : 0.0;
      double dc = std::sqrt(rsq(p, rc));
      double dis = dc - p.coll_R;
      const double Vc[3] = {s.cf[CVX][c], s.cf[CVY][c], d3 ? s.cf[CVZ][c] : 0.0};
      if (dis < 0) {
        for (int a = 0; a < 3; ++a) v[a] = vi[a] - Vc[a];
      } else {
        double xs = rc[0] / dc * p.coll_R, ys = rc[1] / dc * p.coll_R;
        double zs = d3 ? rc[2] / dc * p.coll_R : 0.0;
        double ox = d3 ? s.cf[COX][c] : 0.0, oy = d3 ? s.cf[COY][c] : 0.0, oz = s.cf[COZ][c];
        double vr[3] = {oy * zs - oz * ys, oz * xs - ox * zs, ox * ys - oy * xs};
        if (!d3) { vr[0] = -oz * ys; vr[1] = oz * xs; vr[2] = 0.0; }
        double sv[3] = {rc[0] / dc, rc[1] / dc, d3 ? rc[2] / dc : 0.0};
        double djs = (d3 ? rij[0] * sv[0] + rij[1] * sv[1] + rij[2] * sv[2]
                         : rij[0] * sv[0] + rij[1] * sv[1]) - dis;
        beta = 1.0 + djs / dis;
        if (beta > p.beta_max) beta = p.beta_max;
        for (int a = 0; a < 3; ++a) v[a] = beta * (vi[a] - Vc[a] - vr[a]);
      }
    }
  } else {
    if (ti == 1 || ti == 2) {
      double di, dj, Vw;
      if (ti == 1) { di = p.y_bottom - y[i]; dj = y[j] - p.y_bottom; Vw = p.V_bottom; }
      else { di = y[i] - p.y_top; dj = p.y_top - y[j]; Vw = p.V_top; }
      beta = 1.0 + di / dj;
      if (beta > p.beta_max) beta = p.beta_max;
      if (dj > 0) { v[0] = beta * (Vw - vj[0]); v[1] = -beta * vj[1]; v[2] = -beta * vj[2]; }
      else { v[0] = Vw - vj[0]; v[1] = -vj[1]; v[2] = -vj[2]; }
    } else {
      int c = ti - 3;
      double rc[3];
      rc[0] = mimg(s.f[X][j] - s.cf[CX][c], p.L[0], 0.5 * p.L[0]);
      rc[1] = y[j] - s.cf[CY][c];
      if (p.wall == 0) rc[1] = mimg(rc[1], p.L[1], 0.5 * p.L[1]);
      rc[2] = d3 ? mimg(s.f[Z][j] - s.cf[CZ][c], p.L[2], 0.5 * p.L[2]) : 0.0;
      double dc = std::sqrt(rsq(p, rc));
      double djs = dc - p.coll_R;
      const double Vc[3] = {s.cf[CVX][c], s.cf[CVY][c], d3 ? s.cf[CVZ][c] : 0.0};
      if (djs < 0) {
        for (int a = 0; a < 3; ++a) v[a] = Vc[a] - vj[a];
      } else {
        double xs = rc[0] / dc * p.coll_R, ys = rc[1] / dc * p.coll_R;
        double zs = d3 ? rc[2] / dc * p.coll_R : 0.0;
        double ox = d3 ? s.cf[COX][c] : 0.0, oy = d3 ? s.cf[COY][c] : 0.0, oz = s.cf[COZ][c];
        double vr[3] = {oy * zs - oz * ys, oz * xs - ox * zs, ox * ys - oy * xs};
        if (!d3) { vr[0] = -oz * ys; vr[1] = oz * xs; vr[2] = 0.0; }
        double sv[3] = {rc[0] / dc, rc[1] / dc, d3 ? rc[2] / dc : 0.0};
        double dis = (d3 ? -rij[0] * sv[0] - rij[1] * sv[1] - rij[2] * sv[2]
                         : -rij[0] * sv[0] - rij[1] * sv[1]) - djs;
        beta = 1.0 + dis / djs;
        if (beta > p.beta_max) beta = p.beta_max;
        for (int a = 0; a < 3; ++a) v[a] = beta * (Vc[a] + vr[a] - vj[a]);
      }
    }
  }
  if (!d3) v[2] = 0.0;
}

int cell_of(const sph_params_t &p, double x, double y, double z) {
  // kernel_cell_list (src/kernel_cell_list.cu:21-61)
  int cx = (int)std::floor(x / p.cell_size[0]);
  if (cx == p.Ncells[0]) cx = cx - 1;
  int cy;
  if (p.wall == 0) {
    cy = (int)std::floor(y / p.cell_size[1]);
    if (cy == p.Ncells[1]) cy = cy - 1;
  } else {
    cy = (int)std::floor((y + p.wall_width) / p.cell_size[1]);
  }
  if (p.dim == 2) return cy * p.Ncells[0] + cx;
  int cz = (int)std::floor(z / p.cell_size[2]);
  if (cz == p.Ncells[2]) cz = cz - 1;
  return cz * p.Ncells[0] * p.Ncells[1] + cy * p.Ncells[0] + cx;
}

void colloid_pair_forces(orc_system &s, bool magnetic);

}  // namespace

extern "C" {

orc_system *orc_create(const sph_params_t *p) {
  orc_system *s = new orc_system();
  s->p = *p;
  const int N = p->N;
  for (int k = 0; k < NPF; ++k) s->f[k].assign(N, 0.0);
  s->type.assign(N, 0);
  for (int k = 0; k < NCF; ++k) s->cf[k].assign(p->N_colloids, 0.0);
  std::memset(s->fwall, 0, sizeof(s->fwall));
  s->walls_list.assign(p->Nlist_walls, 0);
  s->walls_start.assign(p->Nwalls, 0);
  s->colloids_list.assign(p->Nlist_colloids, 0);
  s->colloids_start.assign(p->N_colloids, 0);
  s->xc.assign(p->Nlist_colloids, 0.0);
  s->yc.assign(p->Nlist_colloids, 0.0);
  s->zc.assign(p->Nlist_colloids, 0.0);
  s->particle_cell.assign(N, 0);
  s->particle_index.assign(N, 0);
  s->cell_start.assign(p->Ntotal_cells, 0);
  s->cell_end.assign(p->Ntotal_cells, -1);
  s->coll_cell.assign(p->N_colloids, 0);
  s->coll_index.assign(p->N_colloids, 0);
  s->coll_cell_start.assign(std::max(p->Ntotal_cells_colloids, 0), -1);
  s->coll_cell_end.assign(std::max(p->Ntotal_cells_colloids, 0), -2);
  s->mag[0] = s->mag[1] = s->mag[2] = 0.0;
  s->step = 0;
  return s;
}

void orc_destroy(orc_system *s) { delete s; }

static int field_ptr(orc_system *s, const char *name, double **dp, int32_t **ip, int64_t *n) {
  *dp = nullptr; *ip = nullptr;
  for (int k = 0; k < NPF; ++k)
    if (!std::strcmp(name, kPF[k])) { *dp = s->f[k].data(); *n = (int64_t)s->f[k].size(); return 0; }
  for (int k = 0; k < NCF; ++k)
    if (!std::strcmp(name, kCF[k])) { *dp = s->cf[k].data(); *n = (int64_t)s->cf[k].size(); return 0; }
  struct { const char *n; std::vector<int32_t> *v; } iv[] = {
      {"type", &s->type}, {"walls_list", &s->walls_list}, {"walls_start", &s->walls_start},
      {"colloids_list", &s->colloids_list}, {"colloids_start", &s->colloids_start}};
  for (auto &e : iv)
    if (!std::strcmp(name, e.n)) { *ip = e.v->data(); *n = (int64_t)e.v->size(); return 0; }
  struct { const char *n; std::vector<double> *v; } dv[] = {
      {"x_center", &s->xc}, {"y_center", &s->yc}, {"z_center", &s->zc}};
  for (auto &e : dv)
    if (!std::strcmp(name, e.n)) { *dp = e.v->data(); *n = (int64_t)e.v->size(); return 0; }
  if (!std::strcmp(name, "fx_wall")) { *dp = &s->fwall[0][0]; *n = s->p.Nwalls; return 0; }
  if (!std::strcmp(name, "fy_wall")) { *dp = &s->fwall[1][0]; *n = s->p.Nwalls; return 0; }
  if (!std::strcmp(name, "fz_wall")) { *dp = &s->fwall[2][0]; *n = s->p.Nwalls; return 0; }
  return SPH_ERR_ARG;
}

int orc_set_field(orc_system *s, const char *name, const void *src, int64_t count) {
  double *dp; int32_t *ip; int64_t n;
  if (field_ptr(s, name, &dp, &ip, &n) || n != count) return SPH_ERR_ARG;
  if (n == 0) return SPH_OK;  // (an empty field's storage may be null)
  if (dp) std::memcpy(dp, src, n * sizeof(double));
  else std::memcpy(ip, src, n * sizeof(int32_t));
  return SPH_OK;
}

int orc_get_field(orc_system *s, const char *name, void *dst, int64_t count) {
  double *dp; int32_t *ip; int64_t n;
  if (field_ptr(s, name, &dp, &ip, &n) || n != count) return SPH_ERR_ARG;
  if (n == 0) return SPH_OK;
  if (dp) std::memcpy(dst, dp, n * sizeof(double));
  else std::memcpy(dst, ip, n * sizeof(int32_t));
  return SPH_OK;
}

int orc_get_cell_list(orc_system *s, int32_t *pc, int32_t *pi, int32_t *cs, int32_t *ce) {
  const int N = s->p.N, NC = s->p.Ntotal_cells;
  if (pc && N) std::memcpy(pc, s->particle_cell.data(), N * sizeof(int32_t));
  if (pi && N) std::memcpy(pi, s->particle_index.data(), N * sizeof(int32_t));
  if (cs && NC) std::memcpy(cs, s->cell_start.data(), NC * sizeof(int32_t));
  if (ce && NC) std::memcpy(ce, s->cell_end.data(), NC * sizeof(int32_t));
  return SPH_OK;
}

int orc_get_colloid_cell_list(orc_system *s, int32_t *cc, int32_t *ci, int32_t *cs,
                              int32_t *ce) {
  const int Nc = s->p.N_colloids, NCC = std::max(s->p.Ntotal_cells_colloids, 0);
  if (cc && Nc) std::memcpy(cc, s->coll_cell.data(), Nc * sizeof(int32_t));
  if (ci && Nc) std::memcpy(ci, s->coll_index.data(), Nc * sizeof(int32_t));
  if (cs && NCC) std::memcpy(cs, s->coll_cell_start.data(), NCC * sizeof(int32_t));
  if (ce && NCC) std::memcpy(ce, s->coll_cell_end.data(), NCC * sizeof(int32_t));
  return SPH_OK;
}

// neighbours_search (src/system_neighbours_search.cu:19-110): hash, stable sort
// by (cell, ID) - Thrust's int-key radix sort is stable and starts from
// particle_index[i] = i -, ranges; then the colloid grid when N_colloids >= 2.
int orc_neighbours_search(orc_system *s) {
  const sph_params_t &p = s->p;
  const int N = p.N;
  std::vector<int32_t> cell(N);
  int bad = 0;
#pragma omp parallel for reduction(| : bad)
  for (int i = 0; i < N; ++i) {
    cell[i] = cell_of(p, s->f[X][i], s->f[Y][i], s->f[Z][i]);
    if (cell[i] < 0 || cell[i] > p.Ntotal_cells - 1) bad = 1;
  }
  if (bad) return SPH_ERR_OUT_OF_GRID;
  std::vector<int32_t> idx(N);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cell[a] < cell[b]; });
  for (int k = 0; k < N; ++k) {
    s->particle_index[k] = idx[k];
    s->particle_cell[k] = cell[idx[k]];
  }
  std::fill(s->cell_start.begin(), s->cell_start.end(), 0);
  std::fill(s->cell_end.begin(), s->cell_end.end(), -1);
  for (int k = 0; k < N; ++k) {  // kernel_cell_ranges (src/kernel_cell_ranges.cu:13-33)
    int c = s->particle_cell[k];
    if (k == 0 || s->particle_cell[k - 1] != c) s->cell_start[c] = k;
    if (k == N - 1 || s->particle_cell[k + 1] != c) s->cell_end[c] = k;
  }
  if (p.N_colloids >= 2) {
    const int Nc = p.N_colloids;
    std::vector<int32_t> cc(Nc);
    for (int i = 0; i < Nc; ++i) {  // kernel_cell_colloids_list.cu:21-37
      int cx = (int)std::floor(s->cf[CX][i] / p.cell_colloids_size[0]);
      int cy = (int)std::floor(s->cf[CY][i] / p.cell_colloids_size[1]);
      int cz = p.dim == 3 ? (int)std::floor(s->cf[CZ][i] / p.cell_colloids_size[2]) : 0;
      cc[i] = p.dim == 3 ? cz * p.Ncells_colloids[0] * p.Ncells_colloids[1] +
                               cy * p.Ncells_colloids[0] + cx
                         : cy * p.Ncells_colloids[0] + cx;
      if (cc[i] < 0 || cc[i] >= p.Ntotal_cells_colloids) return SPH_ERR_OUT_OF_GRID;
    }
    std::vector<int32_t> ci(Nc);
    std::iota(ci.begin(), ci.end(), 0);
    std::stable_sort(ci.begin(), ci.end(), [&](int a, int b) { return cc[a] < cc[b]; });
    for (int k = 0; k < Nc; ++k) { s->coll_index[k] = ci[k]; s->coll_cell[k] = cc[ci[k]]; }
    std::fill(s->coll_cell_start.begin(), s->coll_cell_start.end(), -1);
    std::fill(s->coll_cell_end.begin(), s->coll_cell_end.end(), -2);
    for (int k = 0; k < Nc; ++k) {
      int c = s->coll_cell[k];
      if (k == 0 || s->coll_cell[k - 1] != c) s->coll_cell_start[c] = k;
      if (k == Nc - 1 || s->coll_cell[k + 1] != c) s->coll_cell_end[c] = k;
    }
  }
  return SPH_OK;
}

// kernel_densities (src/kernel_densities.cu:16-165)
int orc_calculate_density(orc_system *s) {
  const sph_params_t &p = s->p;
  const int N = p.N;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    double d = 0.0;
    for_each_candidate(*s, i, [&](int j) {
      double r[3];
      rij_of(*s, i, j, r);
      double r2 = rsq(p, r);
      if (r2 < p.rcutsq) d = d + W(p, std::sqrt(r2));
    });
    s->f[DENS][i] = d;
  }
  return SPH_OK;
}

// kernel_vel_gradients (src/kernel_vel_gradients.cu:15-662); component
// gradv_ab = dv_b/dx_a accumulates -(1/d_i) gradW e_a v_ij,b.
int orc_calculate_vel_gradient(orc_system *s) {
  const sph_params_t &p = s->p;
  const int N = p.N;
  const int nd = p.dim;
  static const int G[3][3] = {{GXX, GXY, GXZ}, {GYX, GYY, GYZ}, {GZX, GZY, GZZ}};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const int ti = s->type[i];
    const double dinv = 1.0 / s->f[DENS][i];
    double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for_each_candidate(*s, i, [&](int j) {
      if (i == j) return;
      if (ti == s->type[j] && ti != 0) return;
      double r[3];
      rij_of(*s, i, j, r);
      double r2 = rsq(p, r);
      if (!(r2 < p.rcutsq)) return;
      double rr = std::sqrt(r2), rinv = 1.0 / rr;
      double e[3] = {r[0] * rinv, r[1] * rinv, r[2] * rinv};
      double gw = gradW(p, rr);
      double v[3];
      morris_vij(*s, i, j, r, v);
      for (int a = 0; a < nd; ++a)
        for (int b = 0; b < nd; ++b) g[a][b] += -dinv * gw * e[a] * v[b];
    });
    for (int a = 0; a < nd; ++a)
      for (int b = 0; b < nd; ++b) s->f[G[a][b]][i] = g[a][b];
  }
  return SPH_OK;
}

// kernel_pressures (src/kernel_pressures.cu:13-23)
int orc_calculate_pressure(orc_system *s) {
  const sph_params_t &p = s->p;
  const bool no_mass = s->slip == ORC_SLIP_EOS_NO_MASS;
#pragma omp parallel for
  for (int i = 0; i < p.N; ++i)
    s->f[PRESS][i] = p.csq * ((no_mass ? 1.0 : s->f[MASS][i]) * s->f[DENS][i] - p.rho0);
  return SPH_OK;
}

int orc_set_slip(orc_system *s, int slip) {
  s->slip = slip;
  return SPH_OK;
}

}  // extern "C"

namespace {

// kernel_forces (src/kernel_forces.cu:15-865) - pair forces, gathered per i.
void pair_forces(orc_system *s) {
  const sph_params_t &p = s->p;
  const int N = p.N;
  const bool d3 = p.dim == 3;
  const double nkT = p.nkT;
  const double *c[6] = {s->f[CXX].data(), s->f[CXY].data(), s->f[CYY].data(),
                        s->f[CXZ].data(), s->f[CYZ].data(), s->f[CZZ].data()};
  const double *press = s->f[PRESS].data(), *dens = s->f[DENS].data();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const int ti = s->type[i];
    // Pi = p I - nkT c over d^2, in the order xx xy yy xz yz zz.
    auto pi_over_d2 = [&](int k, double pr, double dinv2, double out[6]) {
      out[0] = (pr - nkT * c[0][k]) * dinv2;
      out[1] = (-nkT * c[1][k]) * dinv2;
      out[2] = (pr - nkT * c[2][k]) * dinv2;
      out[3] = d3 ? (-nkT * c[3][k]) * dinv2 : 0.0;
      out[4] = d3 ? (-nkT * c[4][k]) * dinv2 : 0.0;
      out[5] = d3 ? (pr - nkT * c[5][k]) * dinv2 : 0.0;
    };
    const double di0 = dens[i], pi0 = press[i];
    const double di0inv = 1.0 / di0, di0inv2 = di0inv * di0inv;
    double Pi0[6];
    pi_over_d2(i, pi0, di0inv2, Pi0);
    double f[3] = {0.0, 0.0, 0.0};
    for_each_candidate(*s, i, [&](int j) {
      if (i == j) return;
      const int tj = s->type[j];
      if (ti != 0 && tj != 0) return;
      double r[3];
      rij_of(*s, i, j, r);
      double r2 = rsq(p, r);
      if (!(r2 < p.rcutsq)) return;
      double Pi[6], Pj[6], diinv, djinv;
      if (ti == tj) {  // fluid - fluid
        diinv = di0inv;
        djinv = 1.0 / dens[j];
        for (int k = 0; k < 6; ++k) Pi[k] = Pi0[k];
        pi_over_d2(j, press[j], djinv * djinv, Pj);
      } else if (ti == 0) {  // fluid i - wall/colloid j: j takes i's d and p
        diinv = di0inv;
        djinv = di0inv;
        for (int k = 0; k < 6; ++k) Pi[k] = Pi0[k];
        pi_over_d2(j, pi0, djinv * djinv, Pj);
      } else {  // wall/colloid i - fluid j: both sides use j's d, p and c
        djinv = 1.0 / dens[j];
        diinv = djinv;
        pi_over_d2(j, press[j], djinv * djinv, Pj);
        for (int k = 0; k < 6; ++k) Pi[k] = Pj[k];
      }
      double v[3];
      morris_vij(*s, i, j, r, v);
      double rr = std::sqrt(r2), rinv = 1.0 / rr;
      double gw = gradW(p, rr);
      double e[3] = {r[0] * rinv, r[1] * rinv, d3 ? r[2] * rinv : 0.0};
      double ev = d3 ? e[0] * v[0] + e[1] * v[1] + e[2] * v[2] : e[0] * v[0] + e[1] * v[1];
      double axx = -gw * (Pi[0] + Pj[0]), axy = -gw * (Pi[1] + Pj[1]);
      double ayy = -gw * (Pi[2] + Pj[2]);
      if (d3) {
        double axz = -gw * (Pi[3] + Pj[3]), ayz = -gw * (Pi[4] + Pj[4]);
        double azz = -gw * (Pi[5] + Pj[5]);
        f[0] += axx * e[0] + axy * e[1] + axz * e[2];
        f[1] += axy * e[0] + ayy * e[1] + ayz * e[2];
        f[2] += axz * e[0] + ayz * e[1] + azz * e[2];
      } else {
        f[0] += axx * e[0] + axy * e[1];
        f[1] += axy * e[0] + ayy * e[1];
      }
      double aux = gw * diinv * djinv * rinv;
      double a1 = p.a * aux, a2 = p.b * aux * ev;
      f[0] += a1 * v[0] + a2 * e[0];
      f[1] += a1 * v[1] + a2 * e[1];
      if (d3) f[2] += a1 * v[2] + a2 * e[2];
    });
    s->f[FX][i] = f[0];
    s->f[FY][i] = f[1];
    s->f[FZ][i] = d3 ? f[2] : 0.0;
  }
}

// kernel_ext_force (src/kernel_ext_force.cu:13-73)
void ext_forces(orc_system *s) {
  const sph_params_t &p = s->p;
  if (p.ext_force_type == 0) return;
#pragma omp parallel for
  for (int i = 0; i < p.N; ++i) {
    if (s->type[i] != 0) continue;
    double m = s->f[MASS][i];
    if (p.ext_force_type == 1) {
      s->f[FX][i] += m * p.ext_force[0];
      s->f[FY][i] += m * p.ext_force[1];
      if (p.dim == 3) s->f[FZ][i] += m * p.ext_force[2];
    } else if (p.ext_force_type == 2) {
      s->f[FX][i] += m * p.ext_force[0] * std::sin(p.ky * s->f[Y][i]);
    } else if (p.ext_force_type == 3) {
      if (s->f[Y][i] < 0.5 * p.L[1]) s->f[FX][i] += m * p.ext_force[0];
      else s->f[FX][i] += -(m * p.ext_force[0]);
    }
  }
}

// kernel_dcdt (src/kernel_dcdt.cu:15-104): dc/dt = c.gradv + (c.gradv)^T + (I - c)/tau
void dcdt(orc_system *s) {
  const sph_params_t &p = s->p;
  const bool d3 = p.dim == 3;
  const double tinv = 1.0 / p.tau;
#pragma omp parallel for
  for (int i = 0; i < p.N; ++i) {
    if (s->type[i] != 0) continue;
    double cxx = s->f[CXX][i], cxy = s->f[CXY][i], cyy = s->f[CYY][i];
    double gxx = s->f[GXX][i], gxy = s->f[GXY][i], gyx = s->f[GYX][i], gyy = s->f[GYY][i];
    if (!d3) {
      double cgxx = cxx * gxx + cxy * gyx, cgxy = cxx * gxy + cxy * gyy;
      double cgyx = cxy * gxx + cyy * gyx, cgyy = cxy * gxy + cyy * gyy;
      s->f[DXX][i] = 2.0 * cgxx + tinv * (1.0 - cxx);
      s->f[DXY][i] = cgxy + cgyx - tinv * cxy;
      s->f[DYY][i] = 2.0 * cgyy + tinv * (1.0 - cyy);
    } else {
      double cxz = s->f[CXZ][i], cyz = s->f[CYZ][i], czz = s->f[CZZ][i];
      double gxz = s->f[GXZ][i], gyz = s->f[GYZ][i], gzx = s->f[GZX][i];
      double gzy = s->f[GZY][i], gzz = s->f[GZZ][i];
      double cgxx = cxx * gxx + cxy * gyx + cxz * gzx;
      double cgxy = cxx * gxy + cxy * gyy + cxz * gzy;
      double cgxz = cxx * gxz + cxy * gyz + cxz * gzz;
      double cgyx = cxy * gxx + cyy * gyx + cyz * gzx;
      double cgyy = cxy * gxy + cyy * gyy + cyz * gzy;
      double cgyz = cxy * gxz + cyy * gyz + cyz * gzz;
      double cgzx = cxz * gxx + cyz * gyx + czz * gzx;
      double cgzy = cxz * gxy + cyz * gyy + czz * gzy;
      double cgzz = cxz * gxz + cyz * gyz + czz * gzz;
      s->f[DXX][i] = 2.0 * cgxx + tinv * (1.0 - cxx);
      s->f[DXY][i] = cgxy + cgyx - tinv * cxy;
      s->f[DYY][i] = 2.0 * cgyy + tinv * (1.0 - cyy);
      s->f[DXZ][i] = cgxz + cgzx - tinv * cxz;
      s->f[DYZ][i] = cgyz + cgzy - tinv * cyz;
      s->f[DZZ][i] = 2.0 * cgzz + tinv * (1.0 - czz);
    }
  }
}

// kernel_collect_force_walls (src/kernel_collect_force_walls.cu:15-50)
void walls(orc_system *s) {
  const sph_params_t &p = s->p;
  for (int w = 0; w < p.Nwalls; ++w) {
    int end = w < p.Nwalls - 1 ? s->walls_start[w + 1] - 1 : p.Nlist_walls - 1;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int k = s->walls_start[w]; k <= end; ++k) {
      int part = s->walls_list[k];
      fx = fx + s->f[FX][part];
      fy = fy + s->f[FY][part];
      if (p.dim == 3) fz = fz + s->f[FZ][part];
    }
    s->fwall[0][w] = fx;
    s->fwall[1][w] = fy;
    s->fwall[2][w] = p.dim == 3 ? fz : 0.0;
  }
}

// kernel_colloid_forces_to_zero + kernel_collect_force_colloids
// (src/kernel_collect_force_colloids.cu:15-99).
void collect_colloids(orc_system *s) {
  const sph_params_t &p = s->p;
  const bool d3 = p.dim == 3;
  for (int i = 0; i < p.N_colloids; ++i) {
    int end = i < p.N_colloids - 1 ? s->colloids_start[i + 1] - 1 : p.Nlist_colloids - 1;
    double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0;
    for (int k = s->colloids_start[i]; k <= end; ++k) {
      int part = s->colloids_list[k];
      double Fx = s->f[FX][part], Fy = s->f[FY][part], Fz = d3 ? s->f[FZ][part] : 0.0;
      fx = fx + Fx;
      fy = fy + Fy;
      if (d3) fz = fz + Fz;
      double rx = mimg(s->f[X][part] - s->cf[CX][i], p.L[0], 0.5 * p.L[0]);
      double ry = s->f[Y][part] - s->cf[CY][i];
      if (p.wall == 0) ry = mimg(ry, p.L[1], 0.5 * p.L[1]);
      double rz = d3 ? mimg(s->f[Z][part] - s->cf[CZ][i], p.L[2], 0.5 * p.L[2]) : 0.0;
      tz = tz + rx * Fy - ry * Fx;
      if (d3) {
        tx = tx + ry * Fz - rz * Fy;
        ty = ty + rz * Fx - rx * Fz;
      }
    }
    s->cf[CFX][i] = fx; s->cf[CFY][i] = fy; s->cf[CFZ][i] = fz;
    s->cf[CTX][i] = tx; s->cf[CTY][i] = ty; s->cf[CTZ][i] = tz;
  }
}

// kernel_repulsion_force_colloids (src/kernel_repulsion_force_colloids.cu:21-207)
// and kernel_magnetic_force_colloids (src/kernel_magnetic_force_colloids.cu:15-203),
// both over the colloid cell list.
void colloid_pair_forces(orc_system &s, bool magnetic) {
  const sph_params_t &p = s.p;
  const bool d3 = p.dim == 3;
  const int Nc = p.N_colloids;
  const int N0 = p.Ncells_colloids[0], N1 = p.Ncells_colloids[1], N2 = p.Ncells_colloids[2];
  std::vector<double> add[3];
  for (int a = 0; a < 3; ++a) add[a].assign(Nc, 0.0);
  for (int i = 0; i < Nc; ++i) {
    double xi = s.cf[CX][i], yi = s.cf[CY][i], zi = s.cf[CZ][i];
    int cxi = (int)std::floor(xi / p.cell_colloids_size[0]);
    int cyi = (int)std::floor(yi / p.cell_colloids_size[1]);
    int czi = d3 ? (int)std::floor(zi / p.cell_colloids_size[2]) : 0;
    double F[3] = {s.cf[CFX][i], s.cf[CFY][i], s.cf[CFZ][i]};
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = (d3 ? -1 : 0); dz <= (d3 ? 1 : 0); ++dz) {
          int cx = cxi + dx, cy = cyi + dy, cz = czi + dz;
          if (cx < 0) cx += N0;
          if (cx >= N0) cx -= N0;
          if (p.wall == 0) {
            if (cy < 0) cy += N1;
            if (cy >= N1) cy -= N1;
          }
          if (d3) {
            if (cz < 0) cz += N2;
            if (cz >= N2) cz -= N2;
          }
          if (cy < 0 || cy >= N1) continue;
          int cell = d3 ? cz * N0 * N1 + cy * N0 + cx : cy * N0 + cx;
          for (int k = s.coll_cell_start[cell]; k <= s.coll_cell_end[cell]; ++k) {
            int j = s.coll_index[k];
            if (i == j) continue;
            double r[3];
            r[0] = mimg(xi - s.cf[CX][j], p.L[0], 0.5 * p.L[0]);
            r[1] = yi - s.cf[CY][j];
            if (p.wall == 0) r[1] = mimg(r[1], p.L[1], 0.5 * p.L[1]);
            r[2] = d3 ? mimg(zi - s.cf[CZ][j], p.L[2], 0.5 * p.L[2]) : 0.0;
            double r2 = rsq(p, r);
            if (!magnetic) {
              if (r2 < p.rcuton_coll_rep_sq) r2 = p.rcuton_coll_rep_sq;
              if (r2 <= p.rcutoff_coll_rep_sq) {
                double rr = std::sqrt(r2);
                double ts = -p.tau_rep * (rr - 2.0 * p.coll_R) / p.coll_R;
                double ex = std::exp(ts);
                double Fm = (p.F0_rep * p.tau_rep) * ex / (1.0 - ex);
                for (int a = 0; a < p.dim; ++a) F[a] = F[a] + Fm * (r[a] / rr);
              }
            } else if (r2 <= p.r0_magnet_sq) {
              double rr = d3 ? std::sqrt(r2) / p.coll_R : std::sqrt(r2);
              double e[3] = {r[0] / rr, r[1] / rr, d3 ? r[2] / rr : 0.0};
              double me = d3 ? s.mag[0] * e[0] + s.mag[1] * e[1] + s.mag[2] * e[2]
                             : s.mag[0] * e[0] + s.mag[1] * e[1];
              double me2 = me * me;
              double F0r4 = p.F0_magnet / (r2 * r2);
              for (int a = 0; a < p.dim; ++a)
                F[a] = F[a] + F0r4 * (2.0 * me * s.mag[a] - (5.0 * me2 - 1.0) * e[a]);
            }
          }
        }
    for (int a = 0; a < 3; ++a) add[a][i] = F[a];
  }
  for (int i = 0; i < Nc; ++i) {
    s.cf[CFX][i] = add[0][i];
    s.cf[CFY][i] = add[1][i];
    if (d3) s.cf[CFZ][i] = add[2][i];
  }
}

void quat_mul(const double c[4], const double d[4], double r[4]) {
  // kernel_quaternion_product (src/kernel_quaternion_product.cu:15-32)
  r[0] = c[0] * d[0] - c[1] * d[1] - c[2] * d[2] - c[3] * d[3];
  r[1] = c[0] * d[1] + c[1] * d[0] + c[2] * d[3] - c[3] * d[2];
  r[2] = c[0] * d[2] - c[1] * d[3] + c[2] * d[0] + c[3] * d[1];
  r[3] = c[0] * d[3] + c[1] * d[2] - c[2] * d[1] + c[3] * d[0];
}

}  // namespace

extern "C" {

int orc_calculate_pair_forces(orc_system *s) { pair_forces(s); return SPH_OK; }
int orc_calculate_external_forces(orc_system *s) { ext_forces(s); return SPH_OK; }
int orc_calculate_dcdt(orc_system *s) { dcdt(s); return SPH_OK; }
int orc_calculate_force_walls(orc_system *s) {
  if (s->p.wall) walls(s);
  return SPH_OK;
}
// src/system_calculate_force_colloids.cu:40-94
int orc_calculate_force_colloids(orc_system *s) {
  if (s->p.N_colloids > 0) {
    collect_colloids(s);
    if (s->p.N_colloids >= 2) {
      colloid_pair_forces(*s, false);
      colloid_pair_forces(*s, true);
    }
  }
  return SPH_OK;
}

int orc_calculate_forces(orc_system *s) {
  // src/system_calculate_forces.cu:86-170
  orc_calculate_density(s);
  orc_calculate_vel_gradient(s);
  orc_calculate_pressure(s);
  pair_forces(s);
  ext_forces(s);
  dcdt(s);
  orc_calculate_force_walls(s);
  return orc_calculate_force_colloids(s);
}

// kernel_move_sys_VV_part1 (src/kernel_move_sys_VV_part1.cu:13-114) and
// kernel_move_colloids_VV_part1 (src/kernel_move_colloids_VV_part1.cu:15-204).
int orc_V_Verlet_step1(orc_system *s) {
  const sph_params_t &p = s->p;
  const bool d3 = p.dim == 3;
  static const int C6[6] = {CXX, CXY, CYY, CXZ, CYZ, CZZ};
  static const int D6[6] = {DXX, DXY, DYY, DXZ, DYZ, DZZ};
#pragma omp parallel for
  for (int i = 0; i < p.N; ++i) {
    int t = s->type[i];
    if (t > 2) continue;
    double xi = s->f[X][i], yi = s->f[Y][i], zi = d3 ? s->f[Z][i] : 0.0;
    if (t == 0) {
      double hdt = 0.5 * p.dt;
      double hdtm = hdt / s->f[MASS][i];
      s->f[VX][i] = s->f[VX][i] + hdtm * s->f[FX][i];
      s->f[VY][i] = s->f[VY][i] + hdtm * s->f[FY][i];
      int nc = d3 ? 6 : 3;
      for (int k = 0; k < nc; ++k) s->f[C6[k]][i] = s->f[C6[k]][i] + hdt * s->f[D6[k]][i];
      if (d3) s->f[VZ][i] = s->f[VZ][i] + hdtm * s->f[FZ][i];
    } else {
      s->f[VX][i] = t == 1 ? p.V_bottom : p.V_top;
      s->f[VY][i] = 0.0;
      if (d3) s->f[VZ][i] = 0.0;
    }
    xi = xi + s->f[VX][i] * p.dt;
    yi = yi + s->f[VY][i] * p.dt;
    if (d3) zi = zi + s->f[VZ][i] * p.dt;
    if (xi < 0) xi = xi + p.L[0];
    if (p.wall == 0 && yi < 0) yi = yi + p.L[1];
    if (d3 && zi < 0) zi = zi + p.L[2];
    if (xi > p.L[0]) xi = xi - p.L[0];
    if (p.wall == 0 && yi > p.L[1]) yi = yi - p.L[1];
    if (d3 && zi > p.L[2]) zi = zi - p.L[2];
    s->f[X][i] = xi;
    s->f[Y][i] = yi;
    if (d3) s->f[Z][i] = zi;
  }
  if (p.N_colloids > 0 && p.coll_move == 0) {
    for (int i = 0; i < p.N_colloids; ++i) {
      double hdtm = 0.5 * p.dt / p.coll_mass, hdtI = 0.5 * p.dt / p.coll_I;
      auto &C = s->cf;
      C[CVX][i] += hdtm * C[CFX][i];
      C[CVY][i] += hdtm * C[CFY][i];
      if (d3) C[CVZ][i] += hdtm * C[CFZ][i];
      C[COZ][i] += hdtI * C[CTZ][i];
      double omega;
      if (!d3) omega = C[COZ][i];
      else {
        C[COX][i] += hdtI * C[CTX][i];
        C[COY][i] += hdtI * C[CTY][i];
        omega = std::sqrt(C[COX][i] * C[COX][i] + C[COY][i] * C[COY][i] +
                          C[COZ][i] * C[COZ][i]);
      }
      double cx = C[CX][i] + C[CVX][i] * p.dt;
      double cy = C[CY][i] + C[CVY][i] * p.dt;
      double cz = d3 ? C[CZ][i] + C[CVZ][i] * p.dt : 0.0;
      if (cx < 0) cx = cx + p.L[0];
      if (cx > p.L[0]) cx = cx - p.L[0];
      if (p.wall == 0) {
        if (cy < 0) cy = cy + p.L[1];
        if (cy > p.L[1]) cy = cy - p.L[1];
      }
      if (d3) {
        if (cz < 0) cz = cz + p.L[2];
        if (cz > p.L[2]) cz = cz - p.L[2];
      }
      C[CX][i] = cx; C[CY][i] = cy;
      if (d3) C[CZ][i] = cz;
      double ca = 1.0, sa = 0.0, q[4] = {1, 0, 0, 0};
      if (!d3) {
        double th = C[CTH][i] + C[COZ][i] * p.dt;
        sa = std::sin(th);
        ca = std::cos(th);
        C[CTH][i] = th;
      } else {
        double hth = 0.5 * omega * p.dt;
        double dq[4] = {1.0, 0.0, 0.0, 0.0};
        if (omega != 0.0) {
          double ax[3] = {C[COX][i] / omega, C[COY][i] / omega, C[COZ][i] / omega};
          double sh = std::sin(hth);
          dq[0] = std::cos(hth); dq[1] = sh * ax[0]; dq[2] = sh * ax[1]; dq[3] = sh * ax[2];
        }
        double qi[4] = {C[CQ0][i], C[CQX][i], C[CQY][i], C[CQZ][i]};
        quat_mul(dq, qi, q);
        C[CQ0][i] = q[0]; C[CQX][i] = q[1]; C[CQY][i] = q[2]; C[CQZ][i] = q[3];
      }
      int j0 = s->colloids_start[i];
      int j1 = i < p.N_colloids - 1 ? s->colloids_start[i + 1] - 1 : p.Nlist_colloids - 1;
      for (int j = j0; j <= j1; ++j) {
        int part = s->colloids_list[j];
        double px, py, pz = 0.0;
        if (!d3) {
          px = s->xc[j] * ca - s->yc[j] * sa;
          py = s->xc[j] * sa + s->yc[j] * ca;
        } else {
          double qp[4] = {0, s->xc[j], s->yc[j], s->zc[j]}, qinv[4] = {q[0], -q[1], -q[2], -q[3]};
          double t1[4], t2[4];
          quat_mul(q, qp, t1);
          quat_mul(t1, qinv, t2);
          px = t2[1]; py = t2[2]; pz = t2[3];
        }
        double X_ = px + C[CX][i], Y_ = py + C[CY][i], Z_ = d3 ? pz + C[CZ][i] : 0.0;
        if (X_ < 0) X_ = X_ + p.L[0];
        if (X_ > p.L[0]) X_ = X_ - p.L[0];
        if (p.wall == 0) {
          if (Y_ < 0) Y_ = Y_ + p.L[1];
          if (Y_ > p.L[1]) Y_ = Y_ - p.L[1];
        }
        if (d3) {
          if (Z_ < 0) Z_ = Z_ + p.L[2];
          if (Z_ > p.L[2]) Z_ = Z_ - p.L[2];
        }
        s->f[X][part] = X_;
        s->f[Y][part] = Y_;
        if (d3) s->f[Z][part] = Z_;
      }
    }
  }
  return SPH_OK;
}

// kernel_move_sys_VV_part2 (src/kernel_move_sys_VV_part2.cu:13-58) and
// kernel_move_colloids_VV_part2 (src/kernel_move_colloids_VV_part2.cu:14-47).
int orc_V_Verlet_step2(orc_system *s) {
  const sph_params_t &p = s->p;
  const bool d3 = p.dim == 3;
  static const int C6[6] = {CXX, CXY, CYY, CXZ, CYZ, CZZ};
  static const int D6[6] = {DXX, DXY, DYY, DXZ, DYZ, DZZ};
#pragma omp parallel for
  for (int i = 0; i < p.N; ++i) {
    if (s->type[i] != 0) continue;
    double hdt = 0.5 * p.dt;
    double hdtm = hdt / s->f[MASS][i];
    s->f[VX][i] = s->f[VX][i] + hdtm * s->f[FX][i];
    s->f[VY][i] = s->f[VY][i] + hdtm * s->f[FY][i];
    int nc = d3 ? 6 : 3;
    for (int k = 0; k < nc; ++k) s->f[C6[k]][i] = s->f[C6[k]][i] + hdt * s->f[D6[k]][i];
    if (d3) s->f[VZ][i] = s->f[VZ][i] + hdtm * s->f[FZ][i];
  }
  if (p.coll_move == 0 && p.N_colloids > 0) {
    auto &C = s->cf;
    for (int i = 0; i < p.N_colloids; ++i) {
      double hdtm = 0.5 * p.dt / p.coll_mass, hdtI = 0.5 * p.dt / p.coll_I;
      C[CVX][i] += hdtm * C[CFX][i];
      C[CVY][i] += hdtm * C[CFY][i];
      if (d3) C[CVZ][i] += hdtm * C[CFZ][i];
      C[COZ][i] += hdtI * C[CTZ][i];
      if (d3) {
        C[COX][i] += hdtI * C[CTX][i];
        C[COY][i] += hdtI * C[CTY][i];
      }
    }
  }
  return SPH_OK;
}

// Magnetic moment of time step `step` (src/main.cu:395-400).
int orc_set_step(orc_system *s, int step) {
  double ang = s->p.omega_magnet * (s->p.t0 + s->p.dt * (double)step);
  s->mag[0] = std::cos(ang);
  s->mag[1] = std::sin(ang);
  s->mag[2] = 0.0;
  s->step = step;
  return SPH_OK;
}

int orc_prepare(orc_system *s) {
  double ang = s->p.omega_magnet * s->p.t0;  // src/main.cu:290-293
  s->mag[0] = std::cos(ang);
  s->mag[1] = std::sin(ang);
  s->mag[2] = 0.0;
  s->step = 0;
  int rc = orc_neighbours_search(s);
  if (rc) return rc;
  return orc_calculate_forces(s);
}

int orc_step(orc_system *s, int nsteps) {
  for (int n = 0; n < nsteps; ++n) {
    int step = s->step + 1;
    orc_V_Verlet_step1(s);
    int rc = orc_neighbours_search(s);
    if (rc) return rc;
    orc_set_step(s, step);
    orc_calculate_forces(s);
    orc_V_Verlet_step2(s);
  }
  return SPH_OK;
}

// kernel_calculate_macro_vars (src/kernel_calculate_macro_vars.cu:14-40)
double orc_kinetic_energy(orc_system *s) {
  const sph_params_t &p = s->p;
  double ke = 0.0;
#pragma omp parallel for reduction(+ : ke)
  for (int i = 0; i < p.N; ++i) {
    double v2 = s->f[VX][i] * s->f[VX][i] + s->f[VY][i] * s->f[VY][i];
    if (p.dim == 3) v2 = v2 + s->f[VZ][i] * s->f[VZ][i];
    ke += 0.5 * s->f[MASS][i] * v2;
  }
  return ke;
}

int orc_observables(orc_system *s, double out[8]) {
  const sph_params_t &p = s->p;
  const double nan = std::nan("");
  out[0] = orc_kinetic_energy(s);
  double n1 = 0.0, sxy = 0.0, nf = 0.0;
  for (int i = 0; i < p.N; ++i) {
    if (s->type[i] != 0) continue;
    n1 += s->f[CXX][i] - s->f[CYY][i];
    sxy += s->f[CXY][i];
    nf += 1.0;
  }
  out[1] = nf > 0 ? p.nkT * n1 / nf : nan;
  out[2] = nf > 0 ? p.nkT * sxy / nf : nan;
  const double area = p.dim == 3 ? p.L[0] * p.L[2] : p.L[0];
  out[3] = p.wall ? s->fwall[0][0] / area : nan;
  out[4] = p.wall ? s->fwall[0][1] / area : nan;
  double ss = 0.0, cc = 0.0, np_ = 0.0;
  for (int i = 0; i < p.N_colloids; ++i)
    for (int j = i + 1; j < p.N_colloids; ++j) {
      double dx = mimg(s->cf[CX][j] - s->cf[CX][i], p.L[0], 0.5 * p.L[0]);
      double dy = s->cf[CY][j] - s->cf[CY][i];
      if (p.wall == 0) dy = mimg(dy, p.L[1], 0.5 * p.L[1]);
      double dz = p.dim == 3 ? mimg(s->cf[CZ][j] - s->cf[CZ][i], p.L[2], 0.5 * p.L[2]) : 0.0;
      if (dx * dx + dy * dy + dz * dz > p.r0_magnet_sq) continue;
      double phi = std::atan2(dy, dx);
      ss += std::sin(2.0 * phi);
      cc += std::cos(2.0 * phi);
      np_ += 1.0;
    }
  auto half_turn = [](double a) { return a - M_PI * std::ceil(a / M_PI - 0.5); };
  out[5] = np_ > 0 ? half_turn(0.5 * std::atan2(ss, cc)) : nan;
  out[6] = half_turn(p.omega_magnet * (p.t0 + p.dt * (double)s->step));
  out[7] = half_turn(out[6] - out[5]);
  return SPH_OK;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double orc_kernel_W(const sph_params_t *p, double r) { return W(*p, r); }
double orc_kernel_gradW(const sph_params_t *p, double r) { return gradW(*p, r); }

int orc_density_brute_force(orc_system *s, double *out) {
  const sph_params_t &p = s->p;
  const int N = p.N;
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < N; ++i) {
    double d = 0.0;
    for (int j = 0; j < N; ++j) {
      double r[3];
      rij_of(*s, i, j, r);
      double r2 = rsq(p, r);
      if (r2 < p.rcutsq) d += W(p, std::sqrt(r2));
    }
    out[i] = d;
  }
  return SPH_OK;
}

}  // extern "C"

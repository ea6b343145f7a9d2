// sph_device.hip - gfx950 kernels and the device-side owner of a batch of
// simulations (a single simulation is a batch of one).
//
// Layout: particles live in cell-sorted order as structure-of-arrays.  Positions
// are double in both precision forms; every other per-particle field is `R`
// (double for SPH_FP64, float for SPH_FP32).  Each neighbours_search hashes
// the particles, counts per cell, scans the counts into cell_start, and moves
// every particle to cell_start[cell] + (number of particles of its cell with a
// smaller ID) - the stable (cell, ID) order of the reference's Thrust sort - while
// gathering its state into the second buffer set.  Neighbour cells x-1..x+1 of
// one (y, z) row are then one contiguous range ("strip") of the arrays, so the
// pair passes read neighbours coalesced instead of through particle_index.
//
// Kernels and the reference kernels they replace:
//   k_hash, k_scan_*, k_scatter, k_rank_gather   kernel_cell_list.cu:15-63,
//       thrust::sort_by_key (system_neighbours_search.cu:56-58), kernel_cell_ranges.cu:13-33
//   k_density            kernel_densities.cu:16-193 + kernel_pressures.cu:13-23
//   k_gradv              kernel_vel_gradients.cu:15-662 + kernel_dcdt.cu:15-104
//   k_forces             kernel_forces.cu:15-865 + kernel_ext_force.cu:13-73
//   k_walls              kernel_collect_force_walls.cu:15-50 (wavefront reduction)
//   k_colloids           kernel_collect_force_colloids.cu:15-99,
//                        kernel_repulsion_force_colloids.cu:21-207,
//                        kernel_magnetic_force_colloids.cu:15-203
//   k_vv1, k_vv2         kernel_move_sys_VV_part1.cu:13-114, kernel_move_sys_VV_part2.cu:13-58
//   k_coll_vv1, k_coll_place, k_coll_vv2
//                        kernel_move_colloids_VV_part1.cu:15-204, kernel_move_colloids_VV_part2.cu:14-47
//   k_ke_partial, k_ke_final  kernel_calculate_macro_vars.cu:14-40 (deterministic FP64)
//   k_set_step           magnetic moment of src/main.cu:290-293, 395-400
//
// Batches: the simulations' arrays are concatenated (particles, cells,
// colloids, wall and colloid lists) and every launch covers the whole batch.
// Particle IDs, cell indices and list entries are global (offset by the
// simulation's first particle / cell), so the hash, scan and reorder need no
// per-simulation logic and each simulation stays one contiguous range of the
// sorted arrays.  A kernel finds its simulation's parameters through SimTable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "sph_cells.h"
#include "sph_device.h"

// Multiply-add contraction only inside one source expression (a * b + c in a
// single statement becomes one fma), never across statements: the arithmetic
// of a pair is then fixed by the source alone, so the Verlet walk, the strip
// walk and the overflow tail - separate inlined copies of the same pair code -
// round identically, and a batch reproduces its simulations run alone.
#pragma clang fp contract(on)

namespace sph {
namespace {

constexpr int kBlock = 256;
constexpr int kRedBlock = 1024;
constexpr int kKeBlocks = 1024;
constexpr int kTailBlocks = 64;

#define HIPCHK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
      return SPH_ERR_DEVICE;                                                  \
    }                                                                         \
  } while (0)

// ---------------------------------------------------------------------------
// device views
// ---------------------------------------------------------------------------
template <class R>
struct Particles {  // per-particle state that moves with the reorder
  double *x, *y, *z;
  R *v[3];
  R *c[6];  // xx xy yy xz yz zz
  R *mass;
  int *type, *id;
  // Packed copies of the positions, in the current sorted order, kept current
  // wherever positions change (reorder, VV1, colloid placement, uploads).
  // FP64 form: (x, y, z, 0) in double, same order: one 32-byte access per
  // gathered neighbour instead of three scattered 8-byte ones (null in FP32).
  double4 *pd;
  // FP32 form: fixed-point positions (see Fix), same order (null in FP64).
  uint4 *pq;
};

template <class R>
struct Derived {  // per-particle fields recomputed every step
  R *dens, *press;
  R *dinv;   // 1 / dens
  R *f[3];
  R *dcdt[6];
  R *gv[9];  // xx xy xz yx yy yz zx zy zz ; gv_ab = dv_b/dx_a
};

struct Bodies {  // colloids (double) and wall/colloid lists (particle IDs)
  double *s[C_NUM];
  const int *clist, *cstart, *cowner;
  const double *xc, *yc, *zc;
  const int *wlist;
  double *fwall;  // [nsim][3][2]
};

// Where each simulation of the batch sits.  Per-particle launches use blocks
// that never straddle two simulations: tile b covers particles blk_first[b] ..
// blk_first[b] + blockDim.x - 1 of simulation blk_sim[b] (the last tile of a
// simulation is partial).
struct SimTable {
  const sph_params_t *P;  // [nsim]
  const int *blk_sim;     // [nblk]
  const int *blk_first;   // [nblk]
  const int *pbeg;        // [nsim + 1] first particle of each simulation
  const int *cell_off;    // [nsim] first cell
  const int *coll_off;    // [nsim + 1] first colloid
  const int *coll_sim;    // [total colloids] owning simulation
  const int *wrange;      // [2 nsim + 1] wall-list ranges: wall w of sim s is [2s+w, 2s+w+1)
  // [nsim] nonzero: the simulation has failed (a particle left the grid) and
  // is held: from then on no kernel moves its particles or colloids, reorders
  // or rebuilds it, or evaluates its pairs, so a failed member of a batch costs
  // nothing and its state stays as it was when it failed.  Set by the hash.
  int *halt;
  int nsim, nblk, ntotal, xcd;
};

// Tile of this workgroup: XCD-aware (see xcd_tile) or plain.
__device__ __forceinline__ int xcd_tile();
__device__ __forceinline__ int tile_of(const SimTable &T) {
  return T.xcd ? xcd_tile() : (int)blockIdx.x;
}

// Simulation of global particle i (binary search over the particle offsets).
__device__ __forceinline__ int sim_of(const SimTable &T, int i) {
  int lo = 0, hi = T.nsim - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.pbeg[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Verlet lists: every particle within rcut + skin at the last rebuild, in the
// order of a scan over the cells it can reach.  Until any particle of a simulation has
// moved more than skin / 2 from its position at the rebuild (ref), the list
// still holds every pair closer than rcut, so the per-step passes test only the
// list instead of 27 cells.  rebuild[s] asks for a rebuild of simulation s in
// the current step; the reorder, list build and strip scans of the other
// simulations are skipped (their sorted order stays put).
struct VerletCtl {
  int *rebuild;        // [nsim] 1: rebuild this step (set by the displacement checks)
  int *overflow;       // [nsim] a list overflowed at the last build
  float4 *ref;         // [N] positions at the last rebuild
  float half_skin_frac;  // (skin / 2) / rcut, slightly reduced for rounding
  int on;              // displacement checks active
  // [1] some simulation of the batch rebuilds this step: raised with rebuild[s];
  // the step's hash, scan and scatter exit at once while it is 0, and the
  // step's copy-back clears it for the next step (after which an overflowing
  // build raises it again).
  int *any;
  unsigned long long *rebuilds;  // [nsim] list builds so far (instrumentation)
};

struct VList {
  // (cap / 4) * N groups, column-major like NList; an entry is the neighbour's
  // index within its simulation (sorted position - pbeg)
  int4 *idx;
  int *count;  // N: entries found at the build (may exceed cap)
  // Per step, the density pass marks which entries are inside rcut: bit e % 32
  // of word e / 32 (column-major, words * N), so the pair pass visits exactly
  // the neighbours without a second index list.
  unsigned *mask;
  // FP32 form, staged density walk (null when off): each entry's window slot,
  // eight 16-bit slots per group ((cap / 8) * N groups, column-major); per
  // window tile its window (kWinInts ints; see window_build); per simulation
  // the index of its first window
  uint4 *slot;
  int *win;
  const int *wbase;
  int stage_cap;    // positions per window (<= the LDS window of the form)
  int cap;
  float skin_frac;  // skin / rcut
};

// The bodies of simulation s: colloid arrays shifted to its first colloid, so
// a particle of type 3 + k addresses colloid k of its own simulation.
__device__ __forceinline__ Bodies bodies_of(const Bodies &B, const SimTable &T, int s) {
  Bodies r = B;
  const int off = T.coll_off[s];
#pragma unroll
  for (int k = 0; k < C_NUM; ++k) r.s[k] = B.s[k] + off;
  return r;
}

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mimg(double d, double L, double h) {
  if (d > h) d -= L;
  if (d < -h) d += L;
  return d;
}

// The periodic box of the pair loops: lengths and their reciprocals.
struct Box {
  double L[3], iL[3];
};
__device__ __forceinline__ Box box_of(const sph_params_t &P) {
  Box b;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b.L[a] = P.L[a];
    b.iL[a] = 1.0 / P.L[a];
  }
  return b;
}

// Minimum image as d - L rint(d / L) (three DP operations, no compares).  For
// every separation that can be within reach (|d| < L/2 up to an ulp) it equals
// mimg(d, L, L/2) bit for bit: the shift is 0 or +-1, so fma(-L, shift, d)
// rounds d -+ L exactly once, as the subtraction does.
__device__ __forceinline__ double mimg_r(double d, double L, double iL) {
  return fma(-L, rint(d * iL), d);
}

// Fixed-point positions of the FP32 form: coordinate a is stored as the 32-bit
// integer rint((x_a - o_a) 2^32 / span_a).  A periodic axis spans the box
// (o = 0), so the 32-bit difference of two coordinates is their minimum-image
// separation; the wall-normal axis spans twice the walled channel, centred on
// it, so every difference inside the channel is exact.  One 16-byte load
// carries a particle's position, and a separation costs an integer
// subtraction and a rounding to float (relative error ~1e-7 - the rounding of
// the double difference - plus the 2^-32 span quantisation).
struct Fix {
  double s[3], o[3];  // 2^32 / span, origin
  double ud[3];       // span / 2^32
  float u[3];         // span / 2^32
};
__device__ __forceinline__ Fix fix_of(const sph_params_t &P) {
  Fix f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double span = a < P.dim ? P.L[a] : 0.0, o = 0.0;
    if (a == 1 && P.wall) {
      const double h = P.L[1] + 2.0 * P.wall_width;
      span = 2.0 * h;
      o = 0.5 * P.L[1] - h;
    }
    f.s[a] = span > 0.0 ? 4294967296.0 / span : 0.0;
    f.o[a] = o;
    f.ud[a] = span / 4294967296.0;
    f.u[a] = (float)f.ud[a];
  }
  return f;
}
__device__ __forceinline__ unsigned fix_enc(double x, double s, double o) {
  return (unsigned)(long long)rint((x - o) * s);
}
// Loads of the neighbour lists' index groups and hit-mask words: each is read
// once per pass, so they are marked non-temporal, which keeps the L2 for the
// positions and pair records the entries point at (re-read by every particle
// that lists the same neighbour).
__device__ __forceinline__ unsigned ld_stream(const unsigned *p) {
  return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ uint4 ld_stream(const uint4 *p) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const v4u v = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ld_stream(const int4 *p) {
  typedef int v4i __attribute__((ext_vector_type(4)));
  const v4i v = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(p));
  return make_int4(v.x, v.y, v.z, v.w);
}

// (.w = 1 marks a stored position: a buffer load past the records reads 0)
__device__ __forceinline__ uint4 fix_pos(const Fix &f, double x, double y, double z) {
  return make_uint4(fix_enc(x, f.s[0], f.o[0]), fix_enc(y, f.s[1], f.o[1]),
                    fix_enc(z, f.s[2], f.o[2]), 1u);
}
template <int DIM>
__device__ __forceinline__ void fix_sep(const float u[3], uint4 a, uint4 b, float r[3]) {
  r[0] = (float)(int)(a.x - b.x) * u[0];
  r[1] = (float)(int)(a.y - b.y) * u[1];
  r[2] = DIM == 3 ? (float)(int)(a.z - b.z) * u[2] : 0.0f;
}
template <int DIM>
__device__ __forceinline__ float fix_r2(const float u[3], uint4 a, uint4 b) {
  float r[3];
  fix_sep<DIM>(u, a, b, r);
  return DIM == 3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
}

// Writes the packed position copy of particle k: double4 (FP64 form) or fixed
// point (FP32 form).
template <class R>
__device__ __forceinline__ void store_pos(const Particles<R> &S, const sph_params_t &P, int k,
                                          double x, double y, double z) {
  if constexpr (sizeof(R) == 4) S.pq[k] = fix_pos(fix_of(P), x, y, z);
  else S.pd[k] = make_double4(x, y, z, 0.0);
}

// Distance r and 1/r of a pair from r^2 (pair_root), and r alone (pair_sqrt).
// FP64 form: the hardware reciprocal square root (relative error up to 5.2e-8)
// refined by two Newton steps, each of which squares the error: 1/r within
// 1.4e-16 relative (measured over 2^20 inputs in 1e-14 .. 1e-3 against long
// double) - about the correctly rounded sqrt and division of the reference, at
// a third of their instructions.  FP32 form: the hardware
// v_rsq_f32 / v_sqrt_f32 (1 ulp) unless SPH_FP32_IEEE_MATH is defined.
__device__ __forceinline__ double rsqrt_nr(double a) {
  double y = __builtin_amdgcn_rsq(a);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double e = fma(-a * y, y, 1.0);  // 1 - a y^2
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ __forceinline__ void pair_root(double r2, double &rr, double &ri) {
  ri = rsqrt_nr(r2);
  rr = r2 * ri;
}
__device__ __forceinline__ double pair_sqrt(double r2) { return r2 * rsqrt_nr(r2); }
#ifndef SPH_FP32_IEEE_MATH
__device__ __forceinline__ void pair_root(float r2, float &rr, float &ri) {
  ri = __builtin_amdgcn_rsqf(r2);
  rr = r2 * ri;
}
__device__ __forceinline__ float pair_sqrt(float r2) { return __builtin_amdgcn_sqrtf(r2); }
#else
__device__ __forceinline__ void pair_root(float r2, float &rr, float &ri) {
  rr = sqrtf(r2);
  ri = 1.0f / rr;
}
__device__ __forceinline__ float pair_sqrt(float r2) { return sqrtf(r2); }
#endif

// Division and square root in the Morris branches of the pair pass (an IEEE
// division is ~10 instructions in float, more in double).  FP64 form: the
// hardware reciprocal refined by two Newton steps, each of which squares the
// relative error, times the numerator (within 2 ulp), and the pair roots'
// Newton-refined square root.  FP32 form: the hardware reciprocal v_rcp_f32
// and v_sqrt_f32 (1 ulp) unless SPH_FP32_IEEE_MATH is defined, as its pair
// roots do.
// (b = 0 and a = 0 keep the IEEE results: the Newton step would turn the
// reciprocal's infinity into 0 * inf = NaN.)
__device__ __forceinline__ double pair_div(double a, double b) {
  const double y0 = __builtin_amdgcn_rcp(b);
  double y = y0;
#pragma unroll
  for (int k = 0; k < 2; ++k) y = fma(y, fma(-b, y, 1.0), y);
  return a * (b == 0.0 ? y0 : y);
}
__device__ __forceinline__ double branch_sqrt(double a) { return a == 0.0 ? a : pair_sqrt(a); }
#ifndef SPH_FP32_IEEE_MATH
__device__ __forceinline__ float pair_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float branch_sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
#else
__device__ __forceinline__ float pair_div(float a, float b) { return a / b; }
__device__ __forceinline__ float branch_sqrt(float a) { return sqrtf(a); }
#endif

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  The pair kernels map workgroup b to
// tile xcd_tile(b) so that XCD k walks the k-th contiguous eighth of the
// cell-sorted particles and its L2 keeps the neighbour data of that slab.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb / kXcds, r = nb % kXcds;
  const int x = b % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

// Minimum resident wavefronts per SIMD requested for the fused pair kernel
// (caps its registers), per precision form; the FP64 form with walls at three
// (1M-particle walls + colloids step 1.49 -> 1.46 ms), the periodic FP64 form
// uncapped (its 64^3 step is 0.339 ms uncapped, 0.357 ms at three).
#ifndef SPH_PAIRS_WAVES_F32
#define SPH_PAIRS_WAVES_F32 1
#endif
#ifndef SPH_PAIRS_WAVES_F64
#define SPH_PAIRS_WAVES_F64 3
#endif
#ifndef SPH_PAIRS_WAVES_F64_PERIODIC
#define SPH_PAIRS_WAVES_F64_PERIODIC 1
#endif
// Record loads of the pair walk straight into LDS (0: through registers).
#ifndef SPH_PAIRS_LDS_LOADS
#define SPH_PAIRS_LDS_LOADS 1
#endif
// and for the density pass
#ifndef SPH_DENSITY_WAVES
#define SPH_DENSITY_WAVES 1
#endif

// max(x, 0) of an x <= 1: fmin(fmax(x, 0), 1) becomes the clamp output
// modifier of the instruction that forms x (no instruction of its own).
template <class R>
__device__ __forceinline__ R clamp01(R x) { return fmin(fmax(x, R(0)), R(1)); }

// The three spline terms below use, for 0 <= q3 < 3, a = 3 - q3, h = b / 2 =
// max(1 - q3 / 2, 0) and c = max(1 - q3, 0), both clamps free (clamp01).  A term
// past its range is an exact 0, and b^5 = 32 h^5, b^4 = 16 h^4 hold exactly
// (1 - q3 / 2 rounds to half of 2 - q3 rounded, products of h are normal
// numbers, and powers of two scale without rounding), so every result is
// bitwise the one of the piecewise form with b = 2 - q3.

// kernel_W (src/kernel_W.cu:13-59): cw*((a^5 - 6 b^5) + 15 c^5), the b term
// where q3 < 2 and the c term where q3 < 1, in the reference's order (each
// product rounded before its sum: no fused multiply-adds).
template <class R>
__device__ __forceinline__ R sph_W(R r, R rcut_inv, R cw) {
#pragma clang fp contract(off)
  const R q3 = r * rcut_inv * R(3);
  const R a = R(3) - q3, a2 = a * a;
  const R h = clamp01(fma(R(-0.5), q3, R(1))), h2 = h * h;
  const R c = clamp01(R(1) - q3), c2 = c * c;
  return cw * ((a2 * a2 * a - R(192) * (h2 * h2 * h)) + R(15) * (c2 * c2 * c));
}

// The spline polynomial of kernel_W without cw, from q3 = 3 r / rcut, so that a
// sum of W is cw times a sum of these: a^5 - 6 b^5 + 15 c^5.
template <class R>
__device__ __forceinline__ R sph_W_poly(R q3) {
  const R a = R(3) - q3, a2 = a * a;
  const R h = clamp01(fma(R(-0.5), q3, R(1))), h2 = h * h;
  const R c = clamp01(R(1) - q3), c2 = c * c;
  return fma(R(15), c2 * c2 * c, fma(R(-192), h2 * h2 * h, a2 * a2 * a));
}

// kernel_gradW (src/kernel_gradW.cu:13-53) is c_gradw*((a^4 - 6 b^4) + 15 c^4)
// piecewise in q3 = 3 r / rcut.  Its polynomial without c_gradw (0 <= q3 < 3).
template <class R>
__device__ __forceinline__ R sph_gradW_poly(R q3) {
  const R a = R(3) - q3, a2 = a * a;
  const R h = clamp01(fma(R(-0.5), q3, R(1))), h2 = h * h;
  const R c = clamp01(R(1) - q3), c2 = c * c;
  return fma(R(15), c2 * c2, fma(R(-96), h2 * h2, a2 * a2));
}

// Unclamped cell coordinates of a particle, as the pair kernels of the
// reference compute them (e.g. src/kernel_densities.cu:47-53).
template <int DIM, bool WALL>
__device__ __forceinline__ void cell_coords(const sph_params_t &P, double x, double y, double z,
                                            int &cx, int &cy, int &cz) {
  cx = (int)floor(x / P.cell_size[0]);
  cy = WALL ? (int)floor((y + P.wall_width) / P.cell_size[1]) : (int)floor(y / P.cell_size[1]);
  cz = DIM == 3 ? (int)floor(z / P.cell_size[2]) : 0;
}

// Calls fn(k0, k1, sx, sy, sz) for each contiguous sorted range covering the 9
// (2D: 3) neighbour strips of cell (cx, cy, cz), in the reference's cell order
// (dz outer, dy, then x-1, x, x+1).  Interior strips are one range of three
// consecutive cells; at the periodic x seam a strip splits in two.  Rows outside
// the wall-bounded y range are skipped.  (sx, sy, sz) is the periodic shift -
// 0 or +-L per axis - of the range's cells: for every pair closer than the
// cut-off, (x_i - x_j) + s equals the minimum image the reference computes
// pair by pair (cells are at least rcut wide and the box at least 3 cells).
template <int DIM, bool WALL, class Fn>
__device__ __forceinline__ void for_each_strip(const sph_params_t &P, const int *__restrict__ cs,
                                               int cx, int cy, int cz, Fn &&fn) {
  const int Nx = P.Ncells[0], Ny = P.Ncells[1], Nz = P.Ncells[2];
  // A particle outside the grid (already flagged by k_hash) has no neighbours;
  // this keeps every cs[] index in bounds.
  if (cx < 0 || cx > Nx || cy < -1 || cy > Ny || (DIM == 3 && (cz < 0 || cz > Nz))) return;
  // Cells cx-1..cx+1 of a row as x-ranges [a0, a1) then [b0, b1), with shifts
  // sa, sb.  (cx == Nx happens for x == L, whose neighbours are Nx-1, 0, 1.)
  const double Lx = P.L[0];
  int a0, a1, b0 = 0, b1 = 0;
  double sa = 0.0, sb = 0.0;
  if (cx >= 1 && cx <= Nx - 2) {
    a0 = cx - 1; a1 = cx + 2;
  } else if (cx == Nx - 1) {
    a0 = Nx - 2; a1 = Nx; b1 = 1; sb = -Lx;
  } else if (cx == 0) {
    a0 = Nx - 1; a1 = Nx; b1 = 2; sa = Lx;
  } else {
    a0 = Nx - 1; a1 = Nx; b1 = 2; sb = -Lx;
  }
  // One loop over (row, part) with a single call site keeps the caller's body
  // instantiated once.
  constexpr int kRows = DIM == 3 ? 9 : 3;
#pragma unroll 1
  for (int s = 0; s < 2 * kRows; ++s) {
    const int part = s & 1;
    if (part && b1 == 0) continue;
    const int row = s >> 1;
    int y = cy + row % 3 - 1;
    int z = DIM == 3 ? cz + row / 3 - 1 : 0;
    double sy = 0.0, sz = 0.0;
    if (DIM == 3) {
      if (z < 0) { z += Nz; sz = P.L[2]; }
      if (z >= Nz) { z -= Nz; sz = -P.L[2]; }
    }
    if (!WALL) {
      if (y < 0) { y += Ny; sy = P.L[1]; }
      if (y >= Ny) { y -= Ny; sy = -P.L[1]; }
    } else if (y < 0 || y >= Ny) {
      continue;
    }
    const int base = (z * Ny + y) * Nx;
    fn(cs[base + (part ? b0 : a0)], cs[base + (part ? b1 : a1)], part ? sb : sa, sy, sz);
  }
}

// The cells a Verlet list can reach: calls fn(k0, k1, sx, sy, sz) for each
// contiguous sorted range of cells cx - lo .. cx + hi per axis, rows in
// (z, y) order, split at the periodic seam as for_each_strip does.  lo and hi
// are 1, or 2 where the particle lies closer than rv - (cell size) to that
// face of its cell (cells are at least rcut wide and rv = rcut + skin <= 2
// rcut), so a build scans ~3 x 3 x 3 cells instead of 5 x 5 x 5; the hits
// (within rcut, so within the 3 x 3 x 3 block) come in the order of the strip
// scan.  Needs at least 5 cells along every periodic axis (checked on the
// host), so no cell is visited twice.
template <int DIM, bool WALL, class Fn>
__device__ __forceinline__ void for_each_reach(const sph_params_t &P, const int *__restrict__ cs,
                                               double x, double y, double z, int cx, int cy,
                                               int cz, double rv, Fn &&fn) {
  const int Nx = P.Ncells[0], Ny = P.Ncells[1], Nz = P.Ncells[2];
  if (cx < 0 || cx > Nx || cy < -1 || cy > Ny || (DIM == 3 && (cz < 0 || cz > Nz))) return;
  // reach below / above the particle's cell along one axis (a small margin
  // keeps the scan a superset under rounding)
  const double rvm = rv * (1.0 + 1e-6);
  auto reach = [&](double u, double h, int c, int &lo, int &hi) {
    lo = max(1, min(2, c - (int)floor((u - rvm) / h)));
    hi = max(1, min(2, (int)floor((u + rvm) / h) - c));
  };
  int xlo, xhi, ylo, yhi, zlo = 0, zhi = 0;
  reach(x, P.cell_size[0], cx, xlo, xhi);
  reach(WALL ? y + P.wall_width : y, P.cell_size[1], cy, ylo, yhi);
  if (DIM == 3) reach(z, P.cell_size[2], cz, zlo, zhi);
  const double Lx = P.L[0];
  int a0 = cx - xlo, a1 = cx + xhi + 1, b0 = 0, b1 = 0;
  double sa = 0.0, sb = 0.0;
  if (a0 < 0) {          // [Nx + a0, Nx) shifted +L, then [0, a1)
    b1 = a1;
    a1 = Nx;
    a0 = Nx + a0;
    sa = Lx;
  } else if (a1 > Nx) {  // [a0, Nx), then [0, a1 - Nx) shifted -L
    b1 = a1 - Nx;
    a1 = Nx;
    sb = -Lx;
  }
  const int ny = ylo + yhi + 1, nrows = ny * (zlo + zhi + 1);
#pragma unroll 1
  for (int s = 0; s < 2 * nrows; ++s) {
    const int part = s & 1;
    if (part && b1 == 0) continue;
    const int row = s >> 1;
    int y = cy + row % ny - ylo;
    int z = DIM == 3 ? cz + row / ny - zlo : 0;
    double sy = 0.0, sz = 0.0;
    if (DIM == 3) {
      if (z < 0) { z += Nz; sz = P.L[2]; }
      if (z >= Nz) { z -= Nz; sz = -P.L[2]; }
    }
    if (!WALL) {
      if (y < 0) { y += Ny; sy = P.L[1]; }
      if (y >= Ny) { y -= Ny; sy = -P.L[1]; }
    } else if (y < 0 || y >= Ny) {
      continue;
    }
    const int base = (z * Ny + y) * Nx;
    fn(cs[base + (part ? b0 : a0)], cs[base + (part ? b1 : a1)], part ? sb : sa, sy, sz);
  }
}

// Windows of the staged density walk.  At a list build each window tile
// (kWinTile consecutive particles of a simulation, from its first; a workgroup
// of the passes lies inside one) gets a window: the sorted-index ranges of the
// cells within one cell per axis of the cells its particles occupy (periodic
// images folded in, rows beyond a wall dropped), merged.  Every list entry also
// gets its slot in the window, or the far slot (kFarSlot) for the few entries
// that a build's reach adds from the second ring of cells.  The density pass copies the
// window's positions (FP32 form: fixed point; FP64 form: doubles) into LDS
// once per workgroup and reads each entry's position from there: the gathers
// no longer go through the vector L1
// one cache line per lane, and a position is fetched once per tile instead of
// once per particle that lists it.  Windows do not depend on the launch shape.
constexpr int kWinTile = 256;                 // particles per window tile
constexpr int kWinRanges = 24;                // merged index ranges per window
constexpr int kWinInts = 2 + 2 * kWinRanges;  // [ranges | -1, positions, (first, count)...]
constexpr int kStageCap = 3584;               // positions per window (3 x 4 B each in LDS)
constexpr int kStageCap2D = 1792;             // 2D: 2 x 4 B each (a 2D window holds ~800)
// FP64 form, 3D: 3 x 8 B each in LDS; two workgroups of the density pass per
// CU.  2D (staged only on request): 2 x 8 B each in the 12 KB that the
// unstaged walk's hand-over takes anyway, so the pass keeps its occupancy.
constexpr int kStageCap64 = 3376;
constexpr int kStageCap64_2D = 768;
template <class R>
constexpr int stage_cap_of(int dim) {
  return sizeof(R) == 4 ? (dim == 3 ? kStageCap : kStageCap2D)
                        : (dim == 3 ? kStageCap64 : kStageCap64_2D);
}
// The slot of a list entry outside its tile's window: the LDS element just past
// the window's capacity, which the walk reads like any other slot (and then
// replaces with the position from memory), so that slots need no clamp.  The
// walk's LDS arrays hold kStageLds(cap) elements (a multiple of 64, so that the
// coordinate arrays stay 256 B-aligned for paired LDS reads).
template <class R, int DIM>
constexpr unsigned kFarSlot = (unsigned)stage_cap_of<R>(DIM);
constexpr int kStageLds(int cap) { return (cap + 64) / 64 * 64; }

// The window of a tile whose particles occupy cells c0 .. c1 (the simulation's
// cell indices, sorted order), built by the 64 lanes of one wavefront: w[0] =
// the number of merged ranges, or -1 when the tile spans rows of cells whose
// neighbour rows give more than 64 unmerged ranges, or the window has more than
// kWinRanges ranges or `cap` positions; w[1] = positions; then the (first global
// index, count) of each range in increasing order; woff[g] = the slot of the
// first index of range g.
template <int DIM, bool WALL>
__device__ void window_build(const sph_params_t &P, const int *__restrict__ cs, int c0, int c1,
                             int cap, int *w, int *woff) {
  constexpr int kRows = DIM == 3 ? 9 : 3;
  const int lane = threadIdx.x & 63;
  const int Nx = P.Ncells[0], Ny = P.Ncells[1], Nz = DIM == 3 ? P.Ncells[2] : 1;
  const int r0 = c0 / Nx, r1 = c1 / Nx, nseg = r1 - r0 + 1;
  const bool fits = c0 >= 0 && c1 >= c0 && nseg * kRows <= 32;
  // lane L < 32 makes the (up to two) ranges of segment row r0 + L / kRows and
  // its neighbour row L % kRows
  int2 a = make_int2(INT_MAX, INT_MAX), b = a;
  if (fits && lane < nseg * kRows) {
    const int r = r0 + lane / kRows, nr = lane % kRows;
    const int xa = r == r0 ? c0 - r * Nx : 0, xb = r == r1 ? c1 - r * Nx : Nx - 1;
    int y = r % Ny + nr % 3 - 1, z = DIM == 3 ? r / Ny + nr / 3 - 1 : 0;
    bool row = true;
    if (WALL) row = y >= 0 && y < Ny; else y = (y + Ny) % Ny;
    if (DIM == 3) z = (z + Nz) % Nz;
    if (row) {
      const int base = (z * Ny + y) * Nx, x0 = xa - 1, x1 = xb + 1;
      int p0 = x0, p1 = x1, q0 = -1, q1 = -1;  // cells p0..p1, then q0..q1, of the row
      if (x1 - x0 + 1 >= Nx) { p0 = 0; p1 = Nx - 1; }
      else if (x0 < 0) { p0 = x0 + Nx; p1 = Nx - 1; q0 = 0; q1 = x1; }
      else if (x1 >= Nx) { p1 = Nx - 1; q0 = 0; q1 = x1 - Nx; }
      const int s0 = cs[base + p0], e0 = cs[base + p1 + 1];
      if (e0 > s0) a = make_int2(s0, e0);
      if (q0 >= 0) {
        const int s1 = cs[base + q0], e1 = cs[base + q1 + 1];
        if (e1 > s1) b = make_int2(s1, e1);
      }
    }
  }
  // candidate L = range L & 1 of lane L / 2
  const int src = lane >> 1;
  const int ax = __shfl(a.x, src), ay = __shfl(a.y, src), bx = __shfl(b.x, src), by = __shfl(b.y, src);
  const int cs0 = (lane & 1) ? bx : ax, ce0 = (lane & 1) ? by : ay;
  const bool live = cs0 != INT_MAX;
  // rank by (first index, candidate), then each lane takes the candidate of its rank
  int rank = 0;
  for (int o = 0; o < 64; ++o) {
    const int so = __shfl(cs0, o);
    rank += (so < cs0 || (so == cs0 && o < lane)) ? 1 : 0;
  }
  const int n = __popcll(__ballot(live));
  int s1 = INT_MAX, e1 = INT_MAX;
  for (int o = 0; o < 64; ++o) {
    const int ro = __shfl(rank, o), so = __shfl(cs0, o), eo = __shfl(ce0, o);
    if (__shfl(live ? 1 : 0, o) && ro == lane) { s1 = so; e1 = eo; }
  }
  // merge: a range starts a group unless it begins at or before the furthest
  // end of the ranges before it
  int pm = lane < n ? e1 : INT_MIN;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(pm, o);
    if (lane >= o) pm = max(pm, t);
  }
  int prev = __shfl_up(pm, 1);
  if (lane == 0) prev = INT_MIN;
  const bool head = lane < n && s1 > prev;
  const unsigned long long heads = __ballot(head);
  const int m = __popcll(heads);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const int g = __popcll(heads & upto) - 1;
  const bool last = lane < n && (lane == n - 1 || ((heads >> (lane + 1)) & 1ull));
  const int hl = 63 - __clzll(heads & upto);  // head lane of this lane's group
  const int gs = __shfl(s1, hl < 0 ? 0 : hl);
  const int len = last ? pm - gs : 0;
  int incl = len;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  const int total = __shfl(incl, 63);
  if (last && g < kWinRanges) {
    w[2 + 2 * g] = gs;
    w[3 + 2 * g] = len;
    woff[g] = incl - len;
  }
  if (lane == 0) {
    w[0] = fits && m <= kWinRanges && total <= cap ? m : -1;
    w[1] = total;
  }
}

// Window tile of workgroup tile b and the index of its window.
__device__ __forceinline__ int window_of(const SimTable &T, const int *wbase, int b, int sim) {
  return wbase[sim] + (T.blk_first[b] - T.pbeg[sim]) / kWinTile;
}

// r_ij of a candidate inside a strip segment: (x_i - x_j) + s in double, rounded to R.
template <class R, int DIM>
__device__ __forceinline__ void seg_sep(double xi, double yi, double zi, double xj, double yj,
                                        double zj, double sx, double sy, double sz, R r[3]) {
  r[0] = (R)((xi - xj) + sx);
  r[1] = (R)((yi - yj) + sy);
  r[2] = DIM == 3 ? (R)((zi - zj) + sz) : R(0);
}

// Cut-off test of a candidate in a strip segment.  FP64 form: (x_i - x_j) + s in
// double, as the reference.  FP32 form: the fixed-point separation (Fix; the
// 32-bit wraparound is the segment's periodic shift for every pair within
// reach, since a box holds at least three cells per axis), rounded to float -
// the separation every FP32 pass uses, so the strip scans, the Verlet build and
// the Verlet density walk agree on every hit.  Only pairs within ~1e-7 of the
// cut-off - where W and gradW vanish to fifth and fourth order - can fall on the
// other side of the reference's test.  Both forms give a pure function of the
// positions, identical in every pass that re-scans the strips.
template <class R, int DIM>
struct Cand;
template <int DIM>
struct Cand<double, DIM> {
  double x, y, z;
  // (a load per coordinate array: the list build measured 18 % slower reading
  // the packed double4 copy instead)
  __device__ __forceinline__ void load(const Particles<double> &S, int k) {
    x = S.x[k];
    y = S.y[k];
    z = DIM == 3 ? S.z[k] : 0.0;
  }
};
template <int DIM>
struct Cand<float, DIM> {
  uint4 q;
  __device__ __forceinline__ void load(const Particles<float> &S, int k) { q = S.pq[k]; }
};
template <class R, int DIM>
struct CandSelf;
template <int DIM>
struct CandSelf<double, DIM> {
  double x, y, z, sx, sy, sz;
  __device__ __forceinline__ void init(const Particles<double> &S, int i, const sph_params_t &) {
    x = S.x[i];
    y = S.y[i];
    z = DIM == 3 ? S.z[i] : 0.0;
  }
  __device__ __forceinline__ void shift(double a, double b, double c) { sx = a; sy = b; sz = c; }
  __device__ __forceinline__ double r2(const Cand<double, DIM> &c) const {
    double r[3];
    seg_sep<double, DIM>(x, y, z, c.x, c.y, c.z, sx, sy, sz, r);
    return DIM == 3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
  }
};
template <int DIM>
struct CandSelf<float, DIM> {
  uint4 q;
  float u[3];
  __device__ __forceinline__ void init(const Particles<float> &S, int i, const sph_params_t &P) {
    q = S.pq[i];
    const Fix f = fix_of(P);
    u[0] = f.u[0];
    u[1] = f.u[1];
    u[2] = f.u[2];
  }
  __device__ __forceinline__ void shift(double, double, double) {}
  __device__ __forceinline__ float r2(const Cand<float, DIM> &c) const {
    return fix_r2<DIM>(u, q, c.q);
  }
};

// r_ij = x_i - x_j with the minimum image (x, z, and y without walls), computed
// in double and rounded to R.
template <class R, int DIM, bool WALL>
__device__ __forceinline__ void pair_sep(const Box &X, double xi, double yi, double zi,
                                         double xj, double yj, double zj, R r[3]) {
  r[0] = (R)mimg_r(xi - xj, X.L[0], X.iL[0]);
  r[1] = WALL ? (R)(yi - yj) : (R)mimg_r(yi - yj, X.L[1], X.iL[1]);
  r[2] = DIM == 3 ? (R)mimg_r(zi - zj, X.L[2], X.iL[2]) : R(0);
}

// The parameters the pair body reads, copied out of the parameter block once
// per thread; read through the parameter pointer inside the neighbour loop
// they were re-loaded for every neighbour.
// Pair records with fixed-point positions: the FP32 form (the 2D record then
// fits 32 bytes; the separation of a pair is three integer differences, with
// the minimum image for free, instead of double arithmetic).  The FP64 form
// carries double positions.
template <class R, int DIM>
constexpr bool kFixRec = sizeof(R) == 4;


// (With fixed-point records only the fixed-point scale stays live; the box and
// the fixed-point origins are re-derived from P in the rare wall / colloid
// branches, which keeps the loop's scalar registers for the hot values.)
template <class R, int DIM>
struct PairK {
  const sph_params_t *P;
  Box X;       // double positions: the minimum image of every pair
  float u[3];  // fixed-point positions: span / 2^32
  double y_bottom, y_top;
  // q3s = 3 / rcut, ncg = -c_gradw, na = -a, nb = -b
  R q3s, ncg, bmax, nkT, na, nb, V_bottom, V_top, coll_R;
  __device__ explicit PairK(const sph_params_t &P_)
      : P(&P_), X(kFixRec<R, DIM> ? Box{} : box_of(P_)), y_bottom(P_.y_bottom),
        y_top(P_.y_top), q3s(R(3) * (R)P_.rcut_inv), ncg(-(R)P_.c_gradw),
        bmax((R)P_.beta_max), nkT((R)P_.nkT), na(-(R)P_.a), nb(-(R)P_.b),
        V_bottom((R)P_.V_bottom), V_top((R)P_.V_top), coll_R((R)P_.coll_R) {
    if constexpr (kFixRec<R, DIM>) {
      const Fix f = fix_of(P_);
#pragma unroll
      for (int k = 0; k < 3; ++k) u[k] = f.u[k];
    }
  }
  __device__ __forceinline__ Box box() const { return kFixRec<R, DIM> ? box_of(*P) : X; }
};

// What the pair pass needs of a neighbour j, in registers: position (the
// 32-bit fixed-point coordinates or doubles, as the record holds it), velocity,
// w = 1/d for fluid or -type for wall/colloid particles, and q = Pi/d^2
// (Pi = p I - nkT c; xx xy yy xz yz zz) for fluid or the raw conformation
// tensor c for wall/colloid particles (the only parts the reference's force and
// gradient kernels read of such a partner, src/kernel_forces.cu:196-233,
// 320-336).  Components a dimension lacks are zero.
template <class R>
struct PairVal;
template <>
struct PairVal<float> {
  unsigned X[3];  // fixed-point records
  double x[3];    // double-position records
  float w, v[3], q[6];
};
template <>
struct PairVal<double> {
  double x[3];
  double w, v[3], q[6];
};

// Coordinate a of neighbour j's position in double: the stored double, or the
// fixed-point coordinate mapped back (within 2^-33 of the span of the double
// position).
template <class R, int DIM>
__device__ __forceinline__ double pos_of(const PairK<R, DIM> &K, const PairVal<R> &rj, int a) {
  if constexpr (kFixRec<R, DIM>) {
    const Fix f = fix_of(*K.P);
    return fma((double)rj.X[a], f.ud[a], f.o[a]);
  } else {
    return rj.x[a];
  }
}

// Everything about particle i that the pair body reads.  Its position: the
// fixed-point coordinates or the doubles, as the records hold them.
template <class R>
struct PairSelf {
  double x, y, z;  // double-position records
  unsigned X[3];   // fixed-point records
  int t;
  R v[3];
  R dinv, p;
  // coordinate a as a double (fixed-point values mapped back)
  template <int DIM>
  __device__ __forceinline__ double pos(const PairK<R, DIM> &K, int a) const {
    if constexpr (kFixRec<R, DIM>) {
      const Fix f = fix_of(*K.P);
      return fma((double)X[a], f.ud[a], f.o[a]);
    } else {
      return a == 0 ? x : a == 1 ? y : z;
    }
  }
};

// The colloid state the Morris velocity of a wall/colloid pair reads (fields
// C_X .. C_OZ: position, velocity, angular velocity of colloid c of the pair's
// simulation): from the Bodies arrays, or from a copy the pair pass staged in
// LDS, whose reads do not wait for the record loads in flight.
// fixc(c): the centre in the fixed-point coordinates of the simulation's
// positions (for the forms whose pair records hold fixed-point positions).
struct CollFromBodies {
  const Bodies &B;
  const sph_params_t &P;
  __device__ __forceinline__ double operator()(int f, int c) const { return B.s[f][c]; }
  __device__ __forceinline__ uint4 fixc(int c) const {
    return fix_pos(fix_of(P), B.s[C_X][c], B.s[C_Y][c], B.s[C_Z][c]);
  }
};
constexpr int kCollStage = 32;  // colloids per simulation the pair pass stages in LDS
struct CollFromLds {
  const double (&t)[C_OZ + 1][kCollStage];
  const uint4 (&q)[kCollStage];
  __device__ __forceinline__ double operator()(int f, int c) const { return t[f][c]; }
  __device__ __forceinline__ uint4 fixc(int c) const { return q[c]; }
};

// Relative velocity v_ij with Morris boundary conditions for wall and colloid
// partners (src/kernel_vel_gradients.cu:157-332, 415-630; the same branches in
// src/kernel_forces.cu:196-419, 533-822).  Called only when ti == tj == 0 or
// ti != tj.
template <class R, int DIM, bool WALL, class CV>
__device__ __forceinline__ void morris_vij(const PairK<R, DIM> &K, const CV &B, int ti,
                                           int tj, double xi, double yi, double zi,
                                           double xj, double yj, double zj, uint4 Xi, uint4 Xj,
                                           const R vi[3], const R vj[3], const R rij[3],
                                           R v[3]) {
  const R bmax = K.bmax;
  if (ti == tj) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = vi[a] - vj[a];
    return;
  }
  if (ti == 0) {
    if (tj == 1 || tj == 2) {
      R di, dj, Vw;
      if (tj == 1) {
        di = (R)(yi - K.y_bottom); dj = (R)(K.y_bottom - yj); Vw = K.V_bottom;
      } else {
        di = (R)(K.y_top - yi); dj = (R)(yj - K.y_top); Vw = K.V_top;
      }
      R beta = R(1) + pair_div(dj, di);
      if (beta > bmax) beta = bmax;
      if (di > R(0)) {
        v[0] = beta * (vi[0] - Vw); v[1] = beta * vi[1]; v[2] = beta * vi[2];
      } else {
        v[0] = vi[0] - Vw; v[1] = vi[1]; v[2] = vi[2];
      }
    } else {
      const int c = tj - 3;
      R rc[3];
      if constexpr (kFixRec<R, DIM>)  // (fixed-point forms: the pairs' own arithmetic)
        fix_sep<DIM>(K.u, Xi, B.fixc(c), rc);
      else
        pair_sep<R, DIM, WALL>(K.box(), xi, yi, zi, B(C_X, c), B(C_Y, c),
                               DIM == 3 ? B(C_Z, c) : 0.0, rc);
      R dc = branch_sqrt(rc[0] * rc[0] + rc[1] * rc[1] + rc[2] * rc[2]);
      const R Rc = K.coll_R;
      R dis = dc - Rc;
      const R Vc[3] = {(R)B(C_VX, c), (R)B(C_VY, c), DIM == 3 ? (R)B(C_VZ, c) : R(0)};
      if (dis < R(0)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = vi[a] - Vc[a];
      } else {
        R sv[3] = {pair_div(rc[0], dc), pair_div(rc[1], dc), pair_div(rc[2], dc)};
        R xs = sv[0] * Rc, ys = sv[1] * Rc, zs = sv[2] * Rc;
        R ox = DIM == 3 ? (R)B(C_OX, c) : R(0), oy = DIM == 3 ? (R)B(C_OY, c) : R(0);
        R oz = (R)B(C_OZ, c);
        R vr[3] = {oy * zs - oz * ys, oz * xs - ox * zs, ox * ys - oy * xs};
        R djs = (rij[0] * sv[0] + rij[1] * sv[1] + rij[2] * sv[2]) - dis;
        R beta = R(1) + pair_div(djs, dis);
        if (beta > bmax) beta = bmax;
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = beta * (vi[a] - Vc[a] - vr[a]);
      }
    }
  } else {
    if (ti == 1 || ti == 2) {
      R di, dj, Vw;
      if (ti == 1) {
        di = (R)(K.y_bottom - yi); dj = (R)(yj - K.y_bottom); Vw = K.V_bottom;
      } else {
        di = (R)(yi - K.y_top); dj = (R)(K.y_top - yj); Vw = K.V_top;
      }
      R beta = R(1) + pair_div(di, dj);
      if (beta > bmax) beta = bmax;
      if (dj > R(0)) {
        v[0] = beta * (Vw - vj[0]); v[1] = -beta * vj[1]; v[2] = -beta * vj[2];
      } else {
        v[0] = Vw - vj[0]; v[1] = -vj[1]; v[2] = -vj[2];
      }
    } else {
      const int c = ti - 3;
      R rc[3];
      if constexpr (kFixRec<R, DIM>)
        fix_sep<DIM>(K.u, Xj, B.fixc(c), rc);
      else
        pair_sep<R, DIM, WALL>(K.box(), xj, yj, zj, B(C_X, c), B(C_Y, c),
                               DIM == 3 ? B(C_Z, c) : 0.0, rc);
      R dc = branch_sqrt(rc[0] * rc[0] + rc[1] * rc[1] + rc[2] * rc[2]);
      const R Rc = K.coll_R;
      R djs = dc - Rc;
      const R Vc[3] = {(R)B(C_VX, c), (R)B(C_VY, c), DIM == 3 ? (R)B(C_VZ, c) : R(0)};
      if (djs < R(0)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = Vc[a] - vj[a];
      } else {
        R sv[3] = {pair_div(rc[0], dc), pair_div(rc[1], dc), pair_div(rc[2], dc)};
        R xs = sv[0] * Rc, ys = sv[1] * Rc, zs = sv[2] * Rc;
        R ox = DIM == 3 ? (R)B(C_OX, c) : R(0), oy = DIM == 3 ? (R)B(C_OY, c) : R(0);
        R oz = (R)B(C_OZ, c);
        R vr[3] = {oy * zs - oz * ys, oz * xs - ox * zs, ox * ys - oy * xs};
        R dis = (-rij[0] * sv[0] - rij[1] * sv[1] - rij[2] * sv[2]) - djs;
        R beta = R(1) + pair_div(dis, djs);
        if (beta > bmax) beta = bmax;
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = beta * (Vc[a] + vr[a] - vj[a]);
      }
    }
  }
  if (DIM == 2) v[2] = R(0);
}

// Block-wide sum of `n` doubles per thread; result valid in thread 0.
template <int NV, int BS>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[BS / 64][NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    v[k] = t;
  }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = 0.0;
      for (int w = 0; w < BS / 64; ++w) t += red[w][k];
      v[k] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// neighbour search: hash, count, scan, scatter, stable rank + reorder
// ---------------------------------------------------------------------------
// kernel_cell_list (src/kernel_cell_list.cu:15-63): clamp c == Ncells for x, z
// and for y without walls.  A particle outside the grid (the reference's device
// assert) raises its simulation's flag and is filed under the nearest cell, so
// the sorted order stays a permutation within each simulation and the other
// simulations of a batch are unaffected.
// With `rebuild` (a Verlet step) a simulation that keeps its order is not
// hashed: its particle count goes to its first cell, so that the scan still
// gives every rebuilt simulation's cells their places (its own cell starts are
// not read until its next rebuild).
__global__ __launch_bounds__(kBlock) void k_hash(SimTable T, const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z, int *cell, int *slot,
                                                 int *count, int *flag, const int *rebuild,
                                                 const int *any) {
  if (any && !*any) return;  // no simulation rebuilds this step
  const int b = blockIdx.x, sim = T.blk_sim[b];
  if (rebuild && (!rebuild[sim] || T.halt[sim])) {
    if (threadIdx.x == 0 && T.blk_first[b] == T.pbeg[sim])
      atomicAdd(&count[T.cell_off[sim]], T.pbeg[sim + 1] - T.pbeg[sim]);
    return;
  }
  const sph_params_t &P = T.P[sim];
  const int i = T.blk_first[b] + threadIdx.x;
  int c = -2;  // lanes past the simulation's last particle
  if (i < T.pbeg[sim + 1]) {
    bool ok;
    c = T.cell_off[sim] + cell_of(P, x[i], y[i], z[i], ok);
    if (!ok) {
      atomicOr(&flag[sim], 1);
      atomicOr(&T.halt[sim], 1);
    }
    cell[i] = c;
  }
  const int s = count_into_cell(c, count);
  if (c >= 0) slot[i] = s;
}

// rebuild (when not null): only simulations with rebuild[s] set are reordered.
__global__ __launch_bounds__(kBlock) void k_scatter(SimTable T, const int *__restrict__ cell,
                                                    const int *__restrict__ slot,
                                                    const int *__restrict__ cs,
                                                    const int *__restrict__ id, int *stage_src,
                                                    int *stage_id, const int *rebuild,
                                                    const int *any) {
  if (any && !*any) return;
  const int b = blockIdx.x, sim = T.blk_sim[b];
  const int k = T.blk_first[b] + threadIdx.x;
  if (k >= T.pbeg[sim + 1] || (rebuild && !rebuild[sim]) || T.halt[sim]) return;
  int c = cell[k];
  if (c < 0) return;
  int p = cs[c] + slot[k];
  stage_src[p] = k;
  stage_id[p] = id[k];
}

// Final position = cell start + number of particles of the same cell with a
// smaller ID, i.e. the stable (cell, ID) order.  Gathers the moving state.
// A simulation that is not rebuilt (or is held) keeps its order: with copy_others its state
// is copied position for position into the other buffer set (a full reorder
// flips the sets); without, it is left alone (a Verlet step reorders in place:
// its list build copies the rebuilt simulations back, x v c m type ID; the
// packed position copies, shared by both sets, are written here in the new
// order).
template <class R>
__global__ __launch_bounds__(kBlock) void k_rank_gather(
    SimTable T, Particles<R> src, Particles<R> dst, const int *__restrict__ cell,
    const int *__restrict__ cs, const int *__restrict__ stage_src,
    const int *__restrict__ stage_id, int *rank, int *sorted_cell, int nc,
    const int *rebuild, int *newpos, int copy_others) {
  const int b = blockIdx.x, sim = T.blk_sim[b];
  const int p = T.blk_first[b] + threadIdx.x;
  if (p >= T.pbeg[sim + 1]) return;
  if ((rebuild && !rebuild[sim]) || T.halt[sim]) {
    if (!copy_others) return;
    const double xp = src.x[p], yp = src.y[p], zp = src.z[p];
    dst.x[p] = xp;
    dst.y[p] = yp;
    dst.z[p] = zp;
    store_pos(dst, T.P[sim], p, xp, yp, zp);
#pragma unroll
    for (int a = 0; a < 3; ++a) dst.v[a][p] = src.v[a][p];
    for (int a = 0; a < nc; ++a) dst.c[a][p] = src.c[a][p];
    dst.mass[p] = src.mass[p];
    dst.type[p] = src.type[p];
    dst.id[p] = src.id[p];
    if (newpos) newpos[p] = p;
    return;
  }
  const int k = stage_src[p];
  const int c = cell[k];
  const int s0 = cs[c], s1 = cs[c + 1];
  const int myid = stage_id[p];
  int r = 0;
  for (int q = s0; q < s1; ++q) r += stage_id[q] < myid;
  const int d = s0 + r;
  const double xk = src.x[k], yk = src.y[k], zk = src.z[k];
  dst.x[d] = xk;
  dst.y[d] = yk;
  dst.z[d] = zk;
  store_pos(dst, T.P[sim], d, xk, yk, zk);
#pragma unroll
  for (int a = 0; a < 3; ++a) dst.v[a][d] = src.v[a][k];
  for (int a = 0; a < nc; ++a) dst.c[a][d] = src.c[a][k];
  dst.mass[d] = src.mass[k];
  dst.type[d] = src.type[k];
  dst.id[d] = myid;
  rank[myid] = d;
  sorted_cell[d] = c;
  if (newpos) newpos[k] = d;
}

// tmp[newpos[k]] = a[k]: a per-particle field carried into the new order.
template <class T>
__global__ __launch_bounds__(kBlock) void k_permute(const T *__restrict__ a, T *__restrict__ tmp,
                                                    const int *__restrict__ newpos, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) tmp[newpos[k]] = a[k];
}

// ---------------------------------------------------------------------------
// pair pass 1: density (+ pressure) and the neighbour list
// ---------------------------------------------------------------------------
// Scans the candidate strips once, appending every other particle inside the
// cut-off to the particle's neighbour list in scan order; then sums W(r) over
// the particle itself and its list (src/kernel_densities.cu:16-193 sums the
// same set in cell order).  The list is column-major - entry k of particle i sits at
// nlist[k * N + i] - so a wavefront's 64 lanes write and read 256 contiguous
// bytes per entry.  `count` holds the true number of neighbours even past kmax;
// the pair pass re-scans the strips for the entries that did not fit.
struct NList {
  int4 *idx;   // (kmax / 4) * N groups of four entries
  int *count;  // N: true neighbour count (may exceed kmax)
  int *queue;  // particles whose neighbours did not all fit (k_pairs -> k_pairs_tail)
  int *qlen;   // number of queued particles
  int kmax;    // multiple of 4
};

// The stored record: the fields the dimension uses, packed so that a neighbour
// costs as few 16-byte chunks as possible (gathered as whole chunks):
//   FP32 3D  [X Y Z w] [vx vy vz qxx] [qxy qyy qxz qyz] [qzz]        64 B
//   FP32 2D  [X Y w vx] [vy qxx qxy qyy]                              32 B
//   FP64 3D  [x y] [z vx] [vy vz] [w qxx] [qxy qyy] [qxz qyz] [qzz]  128 B
//   FP64 2D  [x y] [vx vy] [w qxx] [qxy qyy]                          64 B
template <class R, int DIM>
struct PairRec;
template <>
struct alignas(64) PairRec<float, 3> {
  unsigned X[3];
  float w, v[3], q[6];
};
template <>
struct alignas(32) PairRec<float, 2> {
  unsigned X[2];
  float w, v[2], q[3];
};
template <>
struct alignas(128) PairRec<double, 3> {
  double x[3], v[3], w, q[6];
};
template <>
struct alignas(64) PairRec<double, 2> {
  double x[2], v[2], w, q[3];
};
static_assert(sizeof(PairRec<float, 3>) == 64 && sizeof(PairRec<float, 2>) == 32 &&
                  sizeof(PairRec<double, 3>) == 128 && sizeof(PairRec<double, 2>) == 64,
              "pair records are whole 16-byte chunks");

// 16-byte chunks holding the used part of a record (alignment padding skipped).
template <class R, int DIM>
constexpr int kRecVecs = DIM == 3 ? (sizeof(R) == 4 ? 4 : 7) : (sizeof(R) == 4 ? 2 : 4);

template <class R, int DIM>
__device__ __forceinline__ PairVal<R> unpack(const PairRec<R, DIM> &r) {
  PairVal<R> o;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if constexpr (kFixRec<R, DIM>) o.X[a] = a < DIM ? r.X[a] : 0u;
    else o.x[a] = a < DIM ? r.x[a] : 0.0;
    o.v[a] = a < DIM ? r.v[a] : R(0);
  }
  o.w = r.w;
#pragma unroll
  for (int k = 0; k < 6; ++k) o.q[k] = (DIM == 3 || k < 3) ? r.q[k] : R(0);
  return o;
}

// Record from its chunks (as loaded).
template <class R, int DIM, class V>
__device__ __forceinline__ PairVal<R> unpack_chunks(const V (&c)[kRecVecs<R, DIM>]) {
  PairRec<R, DIM> r;
  __builtin_memcpy(&r, c, sizeof(c));
  return unpack<R, DIM>(r);
}

// The record of particle j (global index).
template <class R, int DIM>
__device__ __forceinline__ PairVal<R> load_rec(const PairRec<R, DIM> *__restrict__ rec, int j) {
  const float4 *__restrict__ p = reinterpret_cast<const float4 *>(rec + j);
  float4 v[kRecVecs<R, DIM>];
#pragma unroll
  for (int k = 0; k < kRecVecs<R, DIM>; ++k) v[k] = p[k];
  return unpack_chunks<R, DIM>(v);
}

// Value of lane 4 (l / 4) + K of the quad of lane l (DPP quad_perm [K,K,K,K]).
template <int K>
__device__ __forceinline__ int quad_bcast(int x) {
  return __builtin_amdgcn_update_dpp(0, x, K * 0x55, 0xF, 0xF, false);
}
// DPP quad permutation with control byte C (lane 4a + b reads lane 4a + ((C >> 2b) & 3)).
template <int C>
__device__ __forceinline__ int quad_pick(int x) {
  return __builtin_amdgcn_update_dpp(0, x, C, 0xF, 0xF, false);
}

// A 16-byte buffer load per lane whose destination is LDS: lane L's chunk lands
// at dst + 16 L (dst wave-uniform).  Nothing waits for it: the caller waits
// (vmcnt) before reading the chunks.
typedef unsigned lds_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load_to_lds(__amdgpu_buffer_rsrc_t rs, lds_u4 *dst, int voff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)dst, 16,
                                           voff, 0, 0, 0);
}

// The records of one simulation as a buffer resource (a scalar descriptor:
// base, size in bytes, raw addressing; 0x00020000 is the gfx9 word 3).  Record
// loads through it are whole dwordx4 buffer loads, and an offset past the
// simulation's records reads zeros instead of faulting.  Offsets are 32-bit:
// the host limits a simulation to 2^32 - 1 bytes of records.
struct RecBuf {
  __amdgpu_buffer_rsrc_t rs;
};
template <class R, int DIM>
__device__ __forceinline__ RecBuf rec_buf(const PairRec<R, DIM> *rec, const SimTable &T,
                                          int sim) {
  // uniform by construction (one simulation per workgroup); readfirstlane keeps
  // the descriptor in scalar registers where the compiler cannot prove it
  const int b = __builtin_amdgcn_readfirstlane(T.pbeg[sim]);
  const int n = __builtin_amdgcn_readfirstlane(T.pbeg[sim + 1]) - b;
  RecBuf r;
  r.rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<PairRec<R, DIM> *>(rec + b), 0,
      (int)((unsigned)n * (unsigned)sizeof(PairRec<R, DIM>)), 0x00020000);
  return r;
}

// The record of particle jl (index within the simulation) through the buffer.
template <class R, int DIM>
__device__ __forceinline__ PairVal<R> load_rec(const RecBuf &rb, int jl) {
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  const unsigned off = (unsigned)jl * (unsigned)sizeof(PairRec<R, DIM>);
  u4 v[kRecVecs<R, DIM>];
#pragma unroll
  for (int k = 0; k < kRecVecs<R, DIM>; ++k)
    v[k] = __builtin_amdgcn_raw_buffer_load_b128(rb.rs, (int)(off + 16u * k), 0, 0);
  return unpack_chunks<R, DIM>(v);
}

// Pressure p = c^2 (m d - rho0) (src/kernel_pressures.cu:21), 1/d, and the pair
// record of particle i, with Pi/d^2 in the reference's arithmetic
// (src/kernel_forces.cu:64-88, 547-558).
template <class R, int DIM>
__device__ __forceinline__ void pressure_terms(const sph_params_t &P, const Particles<R> &S,
                                               const Derived<R> &D, PairRec<R, DIM> *rec, int i,
                                               R d) {
  const R p = (R)P.csq * (S.mass[i] * d - (R)P.rho0);
  const R dinv = R(1) / d;
  const R dinv2 = dinv * dinv;
  const R nkT = (R)P.nkT;
  D.press[i] = p;
  D.dinv[i] = dinv;
  const int t = S.type[i];
  constexpr bool d3 = DIM == 3;
  R c[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) c[k] = (d3 || k < 3) ? S.c[k][i] : R(0);
  PairRec<R, DIM> r;
  if constexpr (kFixRec<R, DIM>) {
    const uint4 X = S.pq[i];
    r.X[0] = X.x;
    r.X[1] = X.y;
    if constexpr (d3) r.X[2] = X.z;
  } else {
    r.x[0] = S.x[i];
    r.x[1] = S.y[i];
    if constexpr (d3) r.x[2] = S.z[i];
  }
#pragma unroll
  for (int a = 0; a < DIM; ++a) r.v[a] = S.v[a][i];
  if (t == 0) {
    r.w = dinv;
    r.q[0] = (p - nkT * c[0]) * dinv2;
    r.q[1] = (-nkT * c[1]) * dinv2;
    r.q[2] = (p - nkT * c[2]) * dinv2;
    if constexpr (d3) {
      r.q[3] = (-nkT * c[3]) * dinv2;
      r.q[4] = (-nkT * c[4]) * dinv2;
      r.q[5] = (p - nkT * c[5]) * dinv2;
    }
  } else {
    r.w = -(R)t;
#pragma unroll
    for (int k = 0; k < (d3 ? 6 : 3); ++k) r.q[k] = c[k];
  }
  rec[i] = r;
}

template <class R, int DIM, bool WALL>
__global__ __launch_bounds__(kBlock, SPH_DENSITY_WAVES) void k_density(SimTable T, Particles<R> S, Derived<R> D,
                                                    PairRec<R, DIM> *rec, const int *__restrict__ csg,
                                                    NList L, int write_dens, VList VL,
                                                    VerletCtl V, int verlet_done) {
  const int b = tile_of(T), sim = T.blk_sim[b];
  if (verlet_done && threadIdx.x == 0 && T.blk_first[b] == T.pbeg[sim]) {
    // After a Verlet step's list build (the previous launch): a simulation
    // whose list overflowed is rebuilt again next step (no thread of this pass
    // reads the flags).
    const int of = V.overflow[sim] && !T.halt[sim];
    V.rebuild[sim] = of ? 1 : 0;
    if (of) atomicOr(V.any, 1);  // (the next step's flag)
    V.overflow[sim] = 0;
  }
  const sph_params_t &P = T.P[sim];
  const int *__restrict__ cs = csg + T.cell_off[sim];
  const int N = T.ntotal;  // list stride
  const int iend = T.pbeg[sim + 1];
  const int i0 = T.blk_first[b] + threadIdx.x;
  if (i0 == 0) *L.qlen = 0;  // the overflow queue of the following pair pass
  if (T.halt[sim]) return;
  const bool live = i0 < iend;
  // Verlet walk: every lane of a wave takes part in the pair loads, so the
  // lanes past the simulation's end stay (aliasing its last particle for their
  // reads) and store nothing.
  if (!live && !VL.idx) return;
  const int i = live ? i0 : iend - 1;
  const double *__restrict__ X = S.x;
  const double *__restrict__ Y = S.y;
  const double *__restrict__ Z = S.z;
  int4 *__restrict__ out = L.idx + i;
  const double xi = X[i], yi = Y[i], zi = DIM == 3 ? Z[i] : 0.0;
  int cx, cy, cz;
  cell_coords<DIM, WALL>(P, xi, yi, zi, cx, cy, cz);
  const R rcutsq = (R)P.rcutsq, rinv = (R)P.rcut_inv, cw = (R)P.cw;
  const R rinv3 = R(3) * rinv;
  // The density sum term of a hit at r^2.  FP32 form: the spline polynomial
  // at q3 = r (3 / rcut), the sum scaled by cw once (d = cw (66 + sum); W(0) =
  // 66 cw), in every path alike (Verlet walk and strip scan round the same
  // way, so a particle's density does not depend on which path took it).
  // FP64 form: kernel_W itself, summed from W(0) as the reference does.
  auto w_term = [&](R r2) {
    if constexpr (sizeof(R) == 4) return sph_W_poly<R>(pair_sqrt(r2) * rinv3);
    else return sph_W<R>(pair_sqrt(r2), rinv, cw);
  };
  const Box bx = box_of(P);
  const int kmax = L.kmax;
  int cnt = 0;
  R d = R(0);
  const bool vok = VL.idx && live && VL.count[i] <= VL.cap;
  if constexpr (sizeof(R) == 4) {
    if (VL.idx) {
      // FP32 form, Verlet list: the hits in list order.  One 16-byte
      // fixed-point position per entry; the lanes step through the list groups
      // (four entries) together, so index loads are coalesced and their
      // addresses are uniform but for the lane offset.  A
      // lane past its list asks for an offset past the simulation's positions
      // (zeros, no memory access).  The spline polynomial of the hits is summed
      // and scaled by cw once (d = cw (66 + sum t), q = r (3 / rcut)); the hits
      // are marked in the mask words, stored as each fills, and counted from them.
      const int vc = vok ? VL.count[i] : 0;
      const int4 *__restrict__ vl = VL.idx + i;
      const int capg = VL.cap >> 2;  // groups per list: index loads stay inside it
      // (uniform by construction; readfirstlane keeps the descriptor scalar)
      const int pb = __builtin_amdgcn_readfirstlane(T.pbeg[sim]);
      const int pe = __builtin_amdgcn_readfirstlane(iend);
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<uint4 *>(S.pq + pb), 0, (int)((unsigned)(pe - pb) * 16u), 0x00020000);
      const unsigned oob = (unsigned)(pe - pb) * 16u;
      CandSelf<R, DIM> me;
      me.init(S, i, P);
      auto index_group = [&](int gi) { return ld_stream(&vl[(size_t)min(gi, capg - 1) * N]); };
      auto group_load = [&](const int4 &g4, int e0, uint4(&v)[4]) {
        const int jj[4] = {g4.x, g4.y, g4.z, g4.w};
#pragma unroll
        for (int u = 0; u < 4; ++u)
          v[u] = __builtin_bit_cast(
              uint4, __builtin_amdgcn_raw_buffer_load_b128(
                         rs, (int)(e0 + u < vc ? (unsigned)jj[u] * 16u : oob), 0, 0));
      };
      R tsum = R(0);
      unsigned mw = 0u;
      auto group_eval = [&](int g, const uint4(&c)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const R r2 = (R)me.r2(Cand<R, DIM>{c[u]});
          if (c[u].w != 0u && r2 < rcutsq) {  // (entries past the list read 0)
            tsum = tsum + w_term(r2);
            mw |= 1u << ((g + u) & 31);
          }
        }
        if (g < vc && (((g + 4) & 31) == 0 || g + 4 >= vc)) {
          VL.mask[(size_t)(g >> 5) * N + i] = mw;
          cnt += __builtin_popcount(mw);
          mw = 0u;
        }
      };
      // (uniform: the tile's window, or -1 for the unstaged walk)
      const int *__restrict__ gw =
          VL.slot ? VL.win + (size_t)window_of(T, VL.wbase, b, sim) * kWinInts : nullptr;
      const int nwin = gw ? gw[0] : -1;
      if (nwin >= 0) {
        // Staged walk: the window's positions are copied into LDS (coalesced
        // runs of the sorted order), then each entry's position is read from
        // its slot; entries of the second ring of cells (the far slot) take their
        // index from the list and their position from memory.  The same
        // positions in the same order as the unstaged walk: the same sums.
        constexpr int kCap = DIM == 3 ? kStageCap : kStageCap2D;
        constexpr int kLds = kStageLds(kCap);
        constexpr unsigned kFar = kFarSlot<R, DIM>;
        static_assert(kFar == (unsigned)kCap && kFar < (unsigned)kLds, "far slot in LDS");
        __shared__ unsigned wx[kLds], wy[kLds], wz[DIM == 3 ? kLds : 1];
        __shared__ int wr[kWinInts];
        if ((int)threadIdx.x < 2 + 2 * nwin) wr[threadIdx.x] = gw[threadIdx.x];
        __syncthreads();
        // slot p of the window: range r (held from the thread's previous slot,
        // which is lower) at offset p - ro; four loads in flight per thread
        const int total = wr[1];
        int r = 0, ro = 0;
#pragma unroll 1
        for (int p0 = (int)threadIdx.x; p0 < total; p0 += 4 * (int)blockDim.x) {
          uint4 q[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * (int)blockDim.x;
            if (p < total) {
              while (p >= ro + wr[3 + 2 * r]) ro += wr[3 + 2 * r++];
              q[u] = S.pq[wr[2 + 2 * r] + (p - ro)];
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int p = p0 + u * (int)blockDim.x;
            if (p < total) {
              wx[p] = q[u].x;
              wy[p] = q[u].y;
              if (DIM == 3) wz[p] = q[u].z;
            }
          }
        }
        __syncthreads();
        const uint4 *__restrict__ sl = VL.slot + i;
        const int capq = VL.cap >> 3;
        auto slot_group = [&](int gq) { return ld_stream(&sl[(size_t)min(gq, capq - 1) * N]); };
        uint4 sg = slot_group(0);
#pragma unroll 1
        for (int g = 0; g < vc; g += 8) {
          const uint4 sn = slot_group((g >> 3) + 1);
          const unsigned w4[4] = {sg.x, sg.y, sg.z, sg.w};
          // every slot is read (every slot is inside the LDS arrays: window
          // slots, the far slot, and the 0 that pads a list's last group of
          // eight, which is masked below); entries of the second ring of cells
          // are fetched from memory afterwards
          uint4 c[8];
          bool far = false;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const unsigned sk = (w4[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
            c[u] = make_uint4(wx[sk], wy[sk], DIM == 3 ? wz[sk] : 0u, 1u);
            far |= sk == kFar;
          }
          if (far) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const unsigned sk = (w4[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
              if (sk == kFar && g + u < vc) {
                const int4 q4 = vl[(size_t)((g + u) >> 2) * N];
                const int cq = (g + u) & 3;
                const int j = cq == 0 ? q4.x : cq == 1 ? q4.y : cq == 2 ? q4.z : q4.w;
                c[u] = __builtin_bit_cast(
                    uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((unsigned)j * 16u), 0, 0));
              }
            }
          }
          // (branch-free: the term of every entry is formed and a miss adds
          // +0, which leaves the sum bitwise as if the miss were skipped; the
          // group's eight hit bits are gathered at constant positions, then
          // placed in the mask word at once)
          unsigned m8 = 0u;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const R r2 = (R)me.r2(Cand<R, DIM>{c[u]});
            const R t = w_term(r2);
            const bool hit = g + u < vc && r2 < rcutsq;
            tsum = tsum + (hit ? t : R(0));
            m8 |= hit ? 1u << u : 0u;
          }
          mw |= m8 << (g & 31);
          if (((g + 8) & 31) == 0 || g + 8 >= vc) {
            VL.mask[(size_t)(g >> 5) * N + i] = mw;
            cnt += __builtin_popcount(mw);
            mw = 0u;
          }
          sg = sn;
        }
      } else {
        // The positions of the next group are in flight while a group is
        // evaluated, and its indices one group further ahead.  (Each lane
        // leaves the loop after its own list: finished lanes issue no loads;
        // the group counter is the same in every lane still in it.)
        int4 qn = index_group(1);
        uint4 cur[4];
        group_load(index_group(0), 0, cur);
#pragma unroll 1
        for (int g = 0; g < vc; g += 4) {
          uint4 nxt[4];
          group_load(qn, g + 4, nxt);
          qn = index_group((g >> 2) + 2);
          group_eval(g, cur);
#pragma unroll
          for (int u = 0; u < 4; ++u) cur[u] = nxt[u];
        }
      }
      if (!live) return;
      if (vok) d = cw * (R(66) + tsum);  // (W(0) = 66 cw: the particle itself)
    }
  } else {
    // (uniform: the tile's window, or -1 for the unstaged walk)
    const int *__restrict__ gw =
        VL.idx && VL.slot ? VL.win + (size_t)window_of(T, VL.wbase, b, sim) * kWinInts : nullptr;
    const int nwin = gw ? gw[0] : -1;
    // the LDS of the staged walk's window or of the unstaged walk's position
    // hand-over (a workgroup takes one of the two walks)
    constexpr int kCap = DIM == 3 ? kStageCap64 : kStageCap64_2D;
    constexpr int kLds = kCap + 1;  // (the far slot; 2 x 3 x kLds doubles fit one CU)
    constexpr unsigned kFar = kFarSlot<R, DIM>;
    static_assert(kFar == (unsigned)kCap && kFar < (unsigned)kLds, "far slot in LDS");
    __shared__ __attribute__((aligned(16))) double lds64[DIM * kLds];
    if (nwin >= 0) {
      // FP64 form, staged walk (as the FP32 one above): the window's positions
      // (the doubles of the packed position copy) are copied into LDS, then
      // each entry's position is read from its slot, entries of the second
      // ring of cells (the far slot) from memory.  The same positions, tests and
      // sums in the same order as the unstaged walk below: the same densities.
      double *wxd = lds64, *wyd = lds64 + kLds, *wzd = lds64 + (DIM == 3 ? 2 * kLds : 0);
      __shared__ int wr[kWinInts];
      if ((int)threadIdx.x < 2 + 2 * nwin) wr[threadIdx.x] = gw[threadIdx.x];
      __syncthreads();
      const int total = wr[1];
      int r = 0, ro = 0;
#pragma unroll 1
      for (int p0 = (int)threadIdx.x; p0 < total; p0 += 4 * (int)blockDim.x) {
        double4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int p = p0 + u * (int)blockDim.x;
          if (p < total) {
            while (p >= ro + wr[3 + 2 * r]) ro += wr[3 + 2 * r++];
            q[u] = S.pd[wr[2 + 2 * r] + (p - ro)];
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int p = p0 + u * (int)blockDim.x;
          if (p < total) {
            wxd[p] = q[u].x;
            wyd[p] = q[u].y;
            if (DIM == 3) wzd[p] = q[u].z;
          }
        }
      }
      __syncthreads();
      const int vc = vok ? VL.count[i] : 0;
      const int4 *__restrict__ vl = VL.idx + i;
      const int pb = T.pbeg[sim];
      const uint4 *__restrict__ sl = VL.slot + i;
      const int capq = VL.cap >> 3;
      auto slot_group = [&](int gq) { return ld_stream(&sl[(size_t)min(gq, capq - 1) * N]); };
      uint4 sg = slot_group(0);
      if (vok) d = sph_W<R>(R(0), rinv, cw);  // the particle itself (src/kernel_densities.cu:95-99)
      unsigned mw = 0u;
#pragma unroll 1
      for (int g = 0; g < vc; g += 8) {
        const uint4 sn = slot_group((g >> 3) + 1);
        const unsigned w4[4] = {sg.x, sg.y, sg.z, sg.w};
        double px[8], py[8], pz[8];
        bool far = false;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const unsigned sk = (w4[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
          px[u] = wxd[sk];
          py[u] = wyd[sk];
          pz[u] = DIM == 3 ? wzd[sk] : 0.0;
          far |= sk == kFar;
        }
        if (far) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const unsigned sk = (w4[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
            if (sk == kFar && g + u < vc) {
              const int4 q4 = vl[(size_t)((g + u) >> 2) * N];
              const int cq = (g + u) & 3;
              const int j = cq == 0 ? q4.x : cq == 1 ? q4.y : cq == 2 ? q4.z : q4.w;
              const double4 pj = S.pd[pb + j];
              px[u] = pj.x;
              py[u] = pj.y;
              pz[u] = pj.z;
            }
          }
        }
        // (branch-free as in the FP32 form: a miss, or an entry past the list,
        // adds +0 to the positive d, which leaves it bitwise unchanged)
        unsigned m8 = 0u;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          R rs[3];
          pair_sep<R, DIM, WALL>(bx, xi, yi, zi, px[u], py[u], DIM == 3 ? pz[u] : 0.0, rs);
          const R r2 = DIM == 3 ? rs[0] * rs[0] + rs[1] * rs[1] + rs[2] * rs[2]
                                : rs[0] * rs[0] + rs[1] * rs[1];
          const R t = sph_W<R>(pair_sqrt(r2), rinv, cw);
          const bool hit = g + u < vc && r2 < rcutsq;
          d = d + (hit ? t : R(0));
          m8 |= hit ? 1u << u : 0u;
        }
        mw |= m8 << (g & 31);
        cnt += __builtin_popcount(m8);
        if (((g + 8) & 31) == 0 || g + 8 >= vc) {
          VL.mask[(size_t)(g >> 5) * N + i] = mw;
          mw = 0u;
        }
        sg = sn;
      }
      if (!live) return;
    } else if (VL.idx) {
      // FP64 form, Verlet list: the hits in list order, W summed as they come; all
      // lanes of the wave step through the list groups (four entries) together (a lane
      // past its list or not on the Verlet path has no entries), so index loads
      // are coalesced.  The hits are marked in the mask words, stored as each
      // fills.  Positions (double4, 32 bytes) move in two stages: the two lanes
      // of pair a load, in instruction k, the two 16-byte halves of the position
      // lane 2a + k needs (one line per lane pair instead of one per lane), then
      // the halves go through LDS to the lane that needs them.  The loads of
      // group g + 1 are issued before group g is evaluated; a lane without an
      // entry asks for an offset past the simulation's positions (zeros, no
      // memory access).
      typedef unsigned u4 __attribute__((ext_vector_type(4)));
      constexpr int kSlot = 3;  // 48-byte slots: conflict-free LDS rows
      static_assert(sizeof(u4) * kBlock * kSlot <= sizeof(lds64), "the hand-over fits");
      u4(&st)[64][kSlot] = reinterpret_cast<u4(*)[64][kSlot]>(lds64)[threadIdx.x >> 6];
      const int lane = threadIdx.x & 63, hb = lane & 1, pa = lane & ~1;
      const int vc = vok ? VL.count[i] : 0;
      int vmax = vc;
#pragma unroll
      for (int o = 32; o; o >>= 1) vmax = max(vmax, __shfl_xor(vmax, o));
      vmax = __builtin_amdgcn_readfirstlane(vmax);
      const int4 *__restrict__ vl = VL.idx + i;
      // (uniform by construction; readfirstlane keeps the descriptor scalar)
      const int pb = __builtin_amdgcn_readfirstlane(T.pbeg[sim]);
      const int pe = __builtin_amdgcn_readfirstlane(iend);
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<double4 *>(S.pd + pb), 0, (int)((unsigned)(pe - pb) * 32u), 0x00020000);
      const unsigned oob = (unsigned)(pe - pb) * 32u;
      // (a lane past its list re-reads its last group: no new lines)
      const int lastg = max(vc - 1, 0) >> 2;
      auto group_load = [&](const int4 &g4, int e0, u4(&v)[4][2]) {
        const int jj[4] = {g4.x, g4.y, g4.z, g4.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          // lane pa + k's position offset (broadcast in the pair) plus this lane's half
          const int mine = (int)(e0 + u < vc ? (unsigned)jj[u] * 32u : oob), ho = 16 * hb;
          v[u][0] = __builtin_amdgcn_raw_buffer_load_b128(rs, quad_pick<0xA0>(mine) + ho, 0, 0);
          v[u][1] = __builtin_amdgcn_raw_buffer_load_b128(rs, quad_pick<0xF5>(mine) + ho, 0, 0);
        }
      };
      // q: indices of the current group, qn: of the following group
      int4 q = ld_stream(&vl[0]);
      int4 qn = ld_stream(&vl[(size_t)min(1, lastg) * N]);
      u4 cur[4][2];
      group_load(q, 0, cur);
      if (vok) d = sph_W<R>(R(0), rinv, cw);  // the particle itself (src/kernel_densities.cu:95-99)
      unsigned mw = 0u;
#pragma unroll 1
      for (int g = 0; g < vmax; g += 4) {
        u4 nxt[4][2];
        group_load(qn, g + 4, nxt);
        q = qn;
        qn = ld_stream(&vl[(size_t)min((g + 8) >> 2, lastg) * N]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          // half hb of the position of lane pa + k -> slot pa + k
          st[pa][hb] = cur[u][0];
          st[pa + 1][hb] = cur[u][1];
          __builtin_amdgcn_wave_barrier();
          u4 mine[2];
          mine[0] = st[lane][0];
          mine[1] = st[lane][1];
          __builtin_amdgcn_wave_barrier();
          if (g + u < vc) {
            double4 pj;
            __builtin_memcpy(&pj, mine, sizeof(pj));
            R r[3];
            pair_sep<R, DIM, WALL>(bx, xi, yi, zi, pj.x, pj.y, DIM == 3 ? pj.z : 0.0, r);
            const R r2 =
                DIM == 3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
            if (r2 < rcutsq) {
              d = d + sph_W<R>(pair_sqrt(r2), rinv, cw);
              mw |= 1u << ((g + u) & 31);
              cnt++;
            }
          }
        }
        if (g < vc && (((g + 4) & 31) == 0 || g + 4 >= vc)) {
          VL.mask[(size_t)(g >> 5) * N + i] = mw;
          mw = 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          cur[u][0] = nxt[u][0];
          cur[u][1] = nxt[u][1];
        }
      }
      if (!live) return;
    }
  }
  if (!vok) {
    R d_over = R(0);  // terms of the hits past the list capacity, in scan order
    int4 buf = make_int4(0, 0, 0, 0);  // last four hits, oldest in .x
    // Phase 1: the candidate test only.  Four candidates' positions are loaded
    // before any is tested, so their latencies overlap.
    CandSelf<R, DIM> me;
    me.init(S, i, P);
    auto test = [&](int k, const Cand<R, DIM> &c) {
      const R r2 = (R)me.r2(c);
      if (r2 < rcutsq && k != i) {
        buf = make_int4(buf.y, buf.z, buf.w, k);
        cnt++;
        if (cnt > kmax) d_over = d_over + w_term(r2);
        else if ((cnt & 3) == 0) out[(size_t)((cnt >> 2) - 1) * N] = buf;
      }
    };
    for_each_strip<DIM, WALL>(P, cs, cx, cy, cz, [&](int k0, int k1, double sx, double sy,
                                                      double sz) {
      me.shift(sx, sy, sz);
      int k = k0;
#pragma unroll 1
      for (; k + 4 <= k1; k += 4) {
        Cand<R, DIM> c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u].load(S, k + u);
#pragma unroll
        for (int u = 0; u < 4; ++u) test(k + u, c[u]);
      }
#pragma unroll 1
      for (; k < k1; ++k) {
        Cand<R, DIM> c;
        c.load(S, k);
        test(k, c);
      }
    });
    const int tail = cnt & 3;
    if (tail && cnt < kmax) {
      for (int q = tail; q < 4; ++q) buf = make_int4(buf.y, buf.z, buf.w, 0);
      out[(size_t)(cnt >> 2) * N] = buf;
    }
    // Phase 2: W over the stored hits with converged lanes (in phase 1 a wave
    // would evaluate W under a mask for every candidate any lane hits).  The
    // minimum image reproduces the strip segment's shift bit for bit.
    // the particle itself (src/kernel_densities.cu:95-99): W(0) = 66 cw
    d = sizeof(R) == 4 ? R(0) : sph_W<R>(R(0), rinv, cw);
    if (write_dens) {
      const int stored = min(cnt, kmax);
#pragma unroll 1
      for (int g = 0; g < stored; g += 4) {
        int4 q = out[(size_t)(g >> 2) * N];
        const int m = min(4, stored - g);
        const int jj[4] = {q.x, q.y, q.z, q.w};
        if constexpr (sizeof(R) == 4) {
          uint4 pj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) pj[u] = S.pq[u < m ? jj[u] : i];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (u >= m) break;
            d = d + w_term((R)me.r2(Cand<R, DIM>{pj[u]}));
          }
        } else {
          double xj[4], yj[4], zj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = u < m ? jj[u] : i;
            xj[u] = X[j];
            yj[u] = Y[j];
            zj[u] = DIM == 3 ? Z[j] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (u >= m) break;
            R r[3];
            pair_sep<R, DIM, WALL>(bx, xi, yi, zi, xj[u], yj[u], zj[u], r);
            const R r2 =
                DIM == 3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
            d = d + w_term(r2);
          }
        }
      }
      if constexpr (sizeof(R) == 4) d = cw * ((R(66) + d) + d_over);
      else d = d + d_over;
    }
  }
  L.count[i] = cnt;
  if (write_dens) {
    D.dens[i] = d;
    pressure_terms<R, DIM>(P, S, D, rec, i, d);
  }
}

// Verlet list build of the simulations being rebuilt: every other particle
// within rcut + skin over the cells it can reach (for_each_reach), in scan
// order; the positions at the build become the reference of the displacement
// checks.  Lists that overflow their
// capacity raise overflow[sim]: those particles use the strip scan, and the
// simulation is rebuilt again next step.
// Step counter and magnetic moment m(t) = (cos w t, sin w t, 0) of simulation
// s at t = t0 + dt k (src/main.cu:290-293, 395-400); `step` >= 0 sets k, -1
// advances the device counter.
__device__ __forceinline__ void set_step_of(const SimTable &T, int *steps, double *mag, int step,
                                            int s) {
  const sph_params_t &P = T.P[s];
  const int k = step >= 0 ? step : steps[s] + 1;
  steps[s] = k;
  const double ang = P.omega_magnet * (P.t0 + P.dt * (double)k);
  mag[3 * s] = cos(ang);
  mag[3 * s + 1] = sin(ang);
  mag[3 * s + 2] = 0.0;
}

// `steps` non-null: the first nsim threads of the grid also do k_set_step's
// work for this step (one launch fewer per step).
template <int DIM, bool WALL, class R>
__global__ __launch_bounds__(kBlock) void k_verlet_build(SimTable T, Particles<R> S,
                                                         Particles<R> G, int nc, int *any,
                                                         const int *__restrict__ csg, VList VL,
                                                         VerletCtl V, int *steps, double *mag,
                                                         int step) {
  if (steps) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < T.nsim) set_step_of(T, steps, mag, step, s);
  }
  // (every reader of the batch rebuild flag in this step has run)
  if (blockIdx.x == 0 && threadIdx.x == 0) *any = 0;
  const int b = tile_of(T), sim = T.blk_sim[b];
  if (!V.rebuild[sim] || T.halt[sim]) return;
  const int i = T.blk_first[b] + threadIdx.x;
  if (i == T.pbeg[sim]) atomicAdd(&V.rebuilds[sim], 1ull);
  const sph_params_t &P = T.P[sim];
  const int *__restrict__ cs = csg + T.cell_off[sim];
  const int N = T.ntotal;
  const int iend = T.pbeg[sim + 1];
  // The reorder of this step gathered the simulation into the other buffer set
  // G in the new order (the packed position copies are shared by both sets):
  // each thread copies its particle back into S, and every position the build
  // reads comes from G, which no thread of this launch writes.
  if (i < iend) {
    S.x[i] = G.x[i];
    S.y[i] = G.y[i];
    S.z[i] = G.z[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) S.v[a][i] = G.v[a][i];
    for (int a = 0; a < nc; ++a) S.c[a][i] = G.c[a][i];
    S.mass[i] = G.mass[i];
    S.type[i] = G.type[i];
    S.id[i] = G.id[i];
  }
  // With the staged density walk: the window of this workgroup's
  // window tile, from the cells of the tile's first and last particle (the
  // reorder of this step sorted them by these cells), before any lane leaves;
  // the tile's first workgroup stores it
  __shared__ int wwin[kWinInts];
  __shared__ int woff[kWinRanges];
  const bool slots = VL.slot != nullptr;
  if (slots) {
    if (threadIdx.x < 64) {
      const int f = T.pbeg[sim] + (T.blk_first[b] - T.pbeg[sim]) / kWinTile * kWinTile;
      const int l = min(f + kWinTile, iend) - 1;
      bool ok;
      const int c0 = cell_of(P, G.x[f], G.y[f], DIM == 3 ? G.z[f] : 0.0, ok);
      const int c1 = cell_of(P, G.x[l], G.y[l], DIM == 3 ? G.z[l] : 0.0, ok);
      window_build<DIM, WALL>(P, cs, c0, c1, VL.stage_cap, wwin, woff);
    }
    __syncthreads();
    if (T.blk_first[b] % kWinTile == T.pbeg[sim] % kWinTile && threadIdx.x < kWinInts)
      VL.win[(size_t)window_of(T, VL.wbase, b, sim) * kWinInts + threadIdx.x] = wwin[threadIdx.x];
  }
  if (i >= iend) return;
  const bool staged = slots && wwin[0] >= 0;
  const int wm = staged ? wwin[0] : 0;
  // the window range of the last slot looked up: [wlo, whi), slot = k + wsb
  int wlo = 0, whi = 0, wsb = 0;
  auto slot_of = [&](int k) -> unsigned {
    if (k >= wlo && k < whi) return (unsigned)(k + wsb);
    int lo = 0, hi = wm - 1;
    while (lo < hi) {  // the last range whose first index is <= k
      const int mid = (lo + hi + 1) >> 1;
      if (wwin[2 + 2 * mid] <= k) lo = mid; else hi = mid - 1;
    }
    const int f = wwin[2 + 2 * lo], n = wwin[3 + 2 * lo];
    if (k < f || k >= f + n) return kFarSlot<R, DIM>;
    wlo = f;
    whi = f + n;
    wsb = woff[lo] - f;
    return (unsigned)(k + wsb);
  };
  const double xi = G.x[i], yi = G.y[i], zi = DIM == 3 ? G.z[i] : 0.0;
  int cx, cy, cz;
  cell_coords<DIM, WALL>(P, xi, yi, zi, cx, cy, cz);
  const double rv = P.rcut * (1.0 + (double)VL.skin_frac);
  const R rv2 = (R)(rv * rv);
  int4 *__restrict__ out = VL.idx + i;
  const int cap = VL.cap;
  const int pb = T.pbeg[sim];
  int cnt = 0;
  int4 buf = make_int4(0, 0, 0, 0);
  // slots of the current group of eight entries, two per word (entry 2k + h
  // in bits 16 h of word k)
  unsigned sw[4] = {0u, 0u, 0u, 0u};
  uint4 *__restrict__ sout = slots ? VL.slot + i : nullptr;
  // The cut-off test of the density pass (float positions in the FP32 form),
  // eight candidates' loads in flight at a time, tested into a hit mask; only
  // the hits are appended (a wave runs the append as often as its lane with
  // the most hits in the group, not once per candidate).  Entries are indices
  // within the simulation (k - pbeg), the offsets the passes' buffer loads
  // take.  The slots of the current group of eight entries are a 128-bit shift
  // register: each hit enters at the top, so after eight the oldest is in the
  // low half of sw[0].
  CandSelf<R, DIM> me;
  me.init(G, i, P);
  auto hit = [&](const Cand<R, DIM> &c) { return (R)me.r2(c) < rv2; };
  auto push_slot = [&](unsigned v) {
    sw[0] = __builtin_amdgcn_alignbit(sw[1], sw[0], 16);
    sw[1] = __builtin_amdgcn_alignbit(sw[2], sw[1], 16);
    sw[2] = __builtin_amdgcn_alignbit(sw[3], sw[2], 16);
    sw[3] = __builtin_amdgcn_alignbit(v, sw[3], 16);
  };
  auto append = [&](int k) {
    buf = make_int4(buf.y, buf.z, buf.w, k - pb);
    if (staged) push_slot(slot_of(k));
    cnt++;
    if ((cnt & 3) == 0 && cnt <= cap) out[(size_t)((cnt >> 2) - 1) * N] = buf;
    if (staged && (cnt & 7) == 0 && cnt <= cap)
      sout[(size_t)((cnt >> 3) - 1) * N] = make_uint4(sw[0], sw[1], sw[2], sw[3]);
  };
  for_each_reach<DIM, WALL>(P, cs, xi, yi, zi, cx, cy, cz, rv, [&](int k0, int k1, double sx,
                                                                    double sy, double sz) {
    me.shift(sx, sy, sz);
    int k = k0;
#pragma unroll 1
    for (; k + 8 <= k1; k += 8) {
      Cand<R, DIM> c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) c[u].load(G, k + u);
      unsigned h = 0u;
#pragma unroll
      for (int u = 0; u < 8; ++u) h |= (hit(c[u]) ? 1u : 0u) << u;
      h &= ~(((unsigned)(i - k) < 8u ? 1u : 0u) << ((i - k) & 7));  // (not the particle itself)
#pragma unroll 1
      while (h) {  // the hits in index order
        const int u = __builtin_ctz(h);
        h &= h - 1u;
        append(k + u);
      }
    }
    if (k + 4 <= k1) {
      Cand<R, DIM> c[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u].load(G, k + u);
      unsigned h = 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u) h |= (hit(c[u]) ? 1u : 0u) << u;
      h &= ~(((unsigned)(i - k) < 4u ? 1u : 0u) << ((i - k) & 3));
#pragma unroll 1
      while (h) {
        const int u = __builtin_ctz(h);
        h &= h - 1u;
        append(k + u);
      }
      k += 4;
    }
#pragma unroll 1
    for (; k < k1; ++k) {
      Cand<R, DIM> c;
      c.load(G, k);
      if (hit(c) && k != i) append(k);
    }
  });
  const int tail = cnt & 3;
  if (tail && cnt < cap) {
    for (int q = tail; q < 4; ++q) buf = make_int4(buf.y, buf.z, buf.w, 0);
    out[(size_t)(cnt >> 2) * N] = buf;
  }
  if (staged && (cnt & 7) && cnt < cap) {
    // the group's entries down to the low slots, zeros above
    for (int q = cnt & 7; q < 8; ++q) push_slot(0u);
    sout[(size_t)(cnt >> 3) * N] = make_uint4(sw[0], sw[1], sw[2], sw[3]);
  }
  VL.count[i] = cnt;
  // the displacement checks compare float positions with this reference
  V.ref[i] = make_float4((float)xi, (float)yi, (float)zi, 0.0f);
  if (cnt > cap) atomicOr(&V.overflow[sim], 1);
}


__global__ void k_fill(int *a, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Float copies of the positions (after an upload of x, y or z).
template <class R>
__global__ __launch_bounds__(kBlock) void k_pack_pos(SimTable T, Particles<R> S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.ntotal) return;
  store_pos(S, T.P[sim_of(T, i)], i, S.x[i], S.y[i], S.z[i]);
}

template <class R, int DIM>
__global__ __launch_bounds__(kBlock) void k_pressure(SimTable T, Particles<R> S, Derived<R> D,
                                                     PairRec<R, DIM> *rec) {
  const int b = blockIdx.x, sim = T.blk_sim[b];
  const int i = T.blk_first[b] + threadIdx.x;
  if (i >= T.pbeg[sim + 1]) return;
  pressure_terms<R, DIM>(T.P[sim], S, D, rec, i, D.dens[i]);
}

// Pi / d^2 with Pi = p I - nkT c, components xx xy yy xz yz zz
// (src/kernel_forces.cu:64-88, 547-558).
template <class R, int DIM>
__device__ __forceinline__ void pi_over_d2(const R *const c[6], int k, R nkT, R pr, R dinv2,
                                           R out[6]) {
  out[0] = (pr - nkT * c[0][k]) * dinv2;
  out[1] = (-nkT * c[1][k]) * dinv2;
  out[2] = (pr - nkT * c[2][k]) * dinv2;
  if (DIM == 3) {
    out[3] = (-nkT * c[3][k]) * dinv2;
    out[4] = (-nkT * c[4][k]) * dinv2;
    out[5] = (pr - nkT * c[5][k]) * dinv2;
  } else {
    out[3] = out[4] = out[5] = R(0);
  }
}

template <class R, int DIM>
__device__ __forceinline__ void ext_force(const sph_params_t &P, R m, double y, R f[3]) {
  // kernel_ext_force (src/kernel_ext_force.cu:22-72), fluid particles only.
  if (P.ext_force_type == 1) {
    f[0] += m * (R)P.ext_force[0];
    f[1] += m * (R)P.ext_force[1];
    if (DIM == 3) f[2] += m * (R)P.ext_force[2];
  } else if (P.ext_force_type == 2) {
    f[0] += m * (R)P.ext_force[0] * (R)sin(P.ky * y);
  } else if (P.ext_force_type == 3) {
    if (y < 0.5 * P.L[1]) f[0] += m * (R)P.ext_force[0];
    else f[0] += -(m * (R)P.ext_force[0]);
  }
}

// Per-particle accumulators of the fused pair pass, as sums of ge = -gradW(r) e
// (e the unit vector from j to i):
//   velocity gradient  (src/kernel_vel_gradients.cu:632-640)
//     gradv_ab = -(1/d_i) sum gradW e_a v_b = (1/d_i) g_ab,  g_ab = sum ge_a v_b;
//   reversible force   (src/kernel_forces.cu:429-433)
//     -sum gradW (Pi_i/d_i^2 + Pi_j/d_j^2).e = (Pi_i/d_i^2) s + sum (Pi_j/d_j^2) ge,
//     s = sum ge: particle i's own stress enters once per particle instead of once per pair;
//   irreversible force (src/kernel_forces.cu:436-440) summed into f per pair.
// pair_finish scales g and adds (Pi_i/d_i^2) s.
template <class R>
struct PairAcc {
  R g[3][3];  // sum of ge_a v_b
  R s[3];     // sum of ge over the force pairs
  R f[3];     // force sum without the (Pi_i/d_i^2) s part
};


// One neighbour j of i: velocity-gradient and force contributions with the
// reference's pair selection, Morris velocities and Pi/d^2 substitutions
// (src/kernel_vel_gradients.cu:123-640, src/kernel_forces.cu:154-850).
// A fluid - fluid pair (j's w = 1/d > 0) is told apart by one test; j's type
// and the pair selection are worked out only for the others.
template <class R, int DIM, bool WALL, bool DO_GV, bool DO_F, class CV>
__device__ __forceinline__ void pair_body(const PairK<R, DIM> &K, const PairVal<R> &rj,
                                          const CV &B, const PairSelf<R> &me,
                                          PairAcc<R> &acc) {
  const bool ff = me.t == 0 && rj.w > R(0);
  int tj = 0;
  bool on_gv = DO_GV, on_f = DO_F;
  if (!ff) {
    tj = rj.w > R(0) ? 0 : (int)(-rj.w);
    on_gv = DO_GV && (me.t == 0 || tj != me.t);
    on_f = DO_F && (me.t == 0 || tj == 0);
  }
  if (!on_gv && !on_f) return;
  R r[3];
  if constexpr (kFixRec<R, DIM>) {
    const uint4 a = make_uint4(me.X[0], me.X[1], me.X[2], 0u);
    const uint4 b = make_uint4(rj.X[0], rj.X[1], rj.X[2], 0u);
    fix_sep<DIM>(K.u, a, b, r);
  } else {
    pair_sep<R, DIM, WALL>(K.X, me.x, me.y, me.z, rj.x[0], rj.x[1], rj.x[2], r);
  }
  const R r2 = DIM == 3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
  R rr, ri;
  pair_root(r2, rr, ri);
  // ge = -gradW(r) e = gr r, gradW = c_gradw * poly(3 r / rcut) (src/kernel_gradW.cu:13-53)
  const R gr = sph_gradW_poly<R>(rr * K.q3s) * (ri * K.ncg);
  const R ge[3] = {gr * r[0], gr * r[1], DIM == 3 ? gr * r[2] : R(0)};
  const R vj[3] = {rj.v[0], rj.v[1], rj.v[2]};
  R v[3];
  if (ff) {
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = me.v[a] - vj[a];
  } else if constexpr (kFixRec<R, DIM>) {
    // (the fixed-point coordinates mapped back; used only by wall and colloid pairs)
    morris_vij<R, DIM, WALL>(K, B, me.t, tj, me.pos(K, 0), me.pos(K, 1),
                             DIM == 3 ? me.pos(K, 2) : 0.0, pos_of(K, rj, 0), pos_of(K, rj, 1),
                             DIM == 3 ? pos_of(K, rj, 2) : 0.0,
                             make_uint4(me.X[0], me.X[1], me.X[2], 0u),
                             make_uint4(rj.X[0], rj.X[1], rj.X[2], 0u), me.v, vj, r, v);
  } else {
    morris_vij<R, DIM, WALL>(K, B, me.t, tj, me.x, me.y, me.z, rj.x[0], rj.x[1], rj.x[2],
                             make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u), me.v, vj,
                             r, v);
  }
  if (on_gv) {
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) acc.g[a][b] = fma(ge[a], v[b], acc.g[a][b]);
  }
  if (on_f) {
    // Q = Pi_j/d_j^2 as the pair's case has it, dd = 1/(d_i d_j)
    // (src/kernel_forces.cu:196-233, 319-336, 429-440).  Fluid - fluid: the
    // record's Pi_j/d_j^2 as it is.
    R Q[6], dd;
#pragma unroll
    for (int q = 0; q < 6; ++q) Q[q] = rj.q[q];
    if (ff) {
      dd = me.dinv * rj.w;
    } else if (me.t == 0) {  // wall/colloid j takes i's density and pressure, keeps its c
      const R nkT = K.nkT, d2 = me.dinv * me.dinv;
      Q[0] = (me.p - nkT * Q[0]) * d2;
      Q[1] = (-nkT * Q[1]) * d2;
      Q[2] = (me.p - nkT * Q[2]) * d2;
      if (DIM == 3) {
        Q[3] = (-nkT * Q[3]) * d2;
        Q[4] = (-nkT * Q[4]) * d2;
        Q[5] = (me.p - nkT * Q[5]) * d2;
      }
      dd = d2;
    } else {  // wall/colloid i: both sides use fluid j's d, p and c
#pragma unroll
      for (int q = 0; q < 6; ++q) Q[q] = Q[q] + Q[q];
      dd = rj.w * rj.w;
    }
#pragma unroll
    for (int a = 0; a < DIM; ++a) acc.s[a] += ge[a];
    if (DIM == 3) {
      acc.f[0] = fma(Q[0], ge[0], fma(Q[1], ge[1], fma(Q[3], ge[2], acc.f[0])));
      acc.f[1] = fma(Q[1], ge[0], fma(Q[2], ge[1], fma(Q[4], ge[2], acc.f[1])));
      acc.f[2] = fma(Q[3], ge[0], fma(Q[4], ge[1], fma(Q[5], ge[2], acc.f[2])));
    } else {
      acc.f[0] = fma(Q[0], ge[0], fma(Q[1], ge[1], acc.f[0]));
      acc.f[1] = fma(Q[1], ge[0], fma(Q[2], ge[1], acc.f[1]));
    }
    // gradW/(d_i d_j r) (a v + b (e.v) e) = -(gr dd)(a v + b (r.v)/r^2 r)
    const R rv = DIM == 3 ? r[0] * v[0] + r[1] * v[1] + r[2] * v[2] : r[0] * v[0] + r[1] * v[1];
    const R h = gr * dd;
    const R ca = h * K.na, cb = (h * K.nb) * (rv * (ri * ri));
#pragma unroll
    for (int a = 0; a < DIM; ++a) acc.f[a] = fma(ca, v[a], fma(cb, r[a], acc.f[a]));
  }
}

template <class R, int DIM>
__device__ __forceinline__ void load_self(const sph_params_t &P, const Particles<R> &S,
                                          const Derived<R> &D, int i, bool forces,
                                          PairSelf<R> &me) {
  if constexpr (kFixRec<R, DIM>) {
    const uint4 X = S.pq[i];
    me.X[0] = X.x;
    me.X[1] = X.y;
    me.X[2] = X.z;
  } else {
    me.x = S.x[i];
    me.y = S.y[i];
    me.z = DIM == 3 ? S.z[i] : 0.0;
  }
  me.t = S.type[i];
  me.v[0] = S.v[0][i];
  me.v[1] = S.v[1][i];
  me.v[2] = DIM == 3 ? S.v[2][i] : R(0);
  me.dinv = D.dinv[i];
  me.p = forces ? D.press[i] : R(0);
}

// Adds (Pi_i/d_i^2) s to the force sum and clears s (fluid i; a wall or colloid
// particle's force holds its partners' stresses only).
template <class R, int DIM>
__device__ __forceinline__ void add_self_stress(const PairRec<R, DIM> *rec, int i,
                                                const PairSelf<R> &me, PairAcc<R> &acc) {
  if (me.t == 0) {
    R q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = DIM == 3 || k < 3 ? rec[i].q[k] : R(0);
    if (DIM == 3) {
      acc.f[0] = fma(q[0], acc.s[0], fma(q[1], acc.s[1], fma(q[3], acc.s[2], acc.f[0])));
      acc.f[1] = fma(q[1], acc.s[0], fma(q[2], acc.s[1], fma(q[4], acc.s[2], acc.f[1])));
      acc.f[2] = fma(q[3], acc.s[0], fma(q[4], acc.s[1], fma(q[5], acc.s[2], acc.f[2])));
    } else {
      acc.f[0] = fma(q[0], acc.s[0], fma(q[1], acc.s[1], acc.f[0]));
      acc.f[1] = fma(q[1], acc.s[0], fma(q[2], acc.s[1], acc.f[1]));
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) acc.s[a] = R(0);
}

// Writes the results of particle i: force (+ external force), gradient tensor
// (store_gv), dc/dt of fluid particles (write_dcdt; src/kernel_dcdt.cu:76-102).
template <class R, int DIM, bool DO_GV, bool DO_F>
__device__ __forceinline__ void pair_finish(const sph_params_t &P, const Particles<R> &S,
                                            const Derived<R> &D,
                                            const PairRec<R, DIM> *rec, int i,
                                            const PairSelf<R> &me, PairAcc<R> &acc,
                                            int store_gv, int write_dcdt, int with_ext,
                                            int kick) {
  if (DO_F) {
    add_self_stress<R, DIM>(rec, i, me, acc);
    if (with_ext && me.t == 0) ext_force<R, DIM>(P, S.mass[i], S.y[i], acc.f);
    D.f[0][i] = acc.f[0];
    D.f[1][i] = acc.f[1];
    D.f[2][i] = DIM == 3 ? acc.f[2] : R(0);
  }
  if (!DO_GV) return;
  R g[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) g[a][b] = me.dinv * acc.g[a][b];
  if (store_gv) {
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b) D.gv[a * 3 + b][i] = g[a][b];
  }
  if (!write_dcdt) return;
  if (me.t != 0) {
#pragma unroll
    for (int k = 0; k < (DIM == 3 ? 6 : 3); ++k) D.dcdt[k][i] = R(0);
    return;
  }
  const R tinv = R(1) / (R)P.tau;
  R c[6], dc[6];
#pragma unroll
  for (int k = 0; k < (DIM == 3 ? 6 : 3); ++k) c[k] = S.c[k][i];
  const R cxx = c[0], cxy = c[1], cyy = c[2];
  if (DIM == 2) {
    R cgxx = cxx * g[0][0] + cxy * g[1][0], cgxy = cxx * g[0][1] + cxy * g[1][1];
    R cgyx = cxy * g[0][0] + cyy * g[1][0], cgyy = cxy * g[0][1] + cyy * g[1][1];
    dc[0] = R(2) * cgxx + tinv * (R(1) - cxx);
    dc[1] = cgxy + cgyx - tinv * cxy;
    dc[2] = R(2) * cgyy + tinv * (R(1) - cyy);
  } else {
    const R cxz = c[3], cyz = c[4], czz = c[5];
    R cgxx = cxx * g[0][0] + cxy * g[1][0] + cxz * g[2][0];
    R cgxy = cxx * g[0][1] + cxy * g[1][1] + cxz * g[2][1];
    R cgxz = cxx * g[0][2] + cxy * g[1][2] + cxz * g[2][2];
    R cgyx = cxy * g[0][0] + cyy * g[1][0] + cyz * g[2][0];
    R cgyy = cxy * g[0][1] + cyy * g[1][1] + cyz * g[2][1];
    R cgyz = cxy * g[0][2] + cyy * g[1][2] + cyz * g[2][2];
    R cgzx = cxz * g[0][0] + cyz * g[1][0] + czz * g[2][0];
    R cgzy = cxz * g[0][1] + cyz * g[1][1] + czz * g[2][1];
    R cgzz = cxz * g[0][2] + cyz * g[1][2] + czz * g[2][2];
    dc[0] = R(2) * cgxx + tinv * (R(1) - cxx);
    dc[1] = cgxy + cgyx - tinv * cxy;
    dc[2] = R(2) * cgyy + tinv * (R(1) - cyy);
    dc[3] = cgxz + cgzx - tinv * cxz;
    dc[4] = cgyz + cgzy - tinv * cyz;
    dc[5] = R(2) * cgzz + tinv * (R(1) - czz);
  }
#pragma unroll
  for (int k = 0; k < (DIM == 3 ? 6 : 3); ++k) D.dcdt[k][i] = dc[k];
  if (DO_F && kick) {
    // V_Verlet_step2 of this fluid particle (src/main.cu step order; the same
    // arithmetic as k_vv2), from the force and dc/dt still in registers: the
    // step needs no separate kick pass.  Only this thread touches particle i's
    // v and c here; the neighbours' were read from the pair records.
    const R hdt = (R)(0.5 * P.dt);
    const R hdtm = hdt / S.mass[i];
#pragma unroll
    for (int a = 0; a < DIM; ++a) S.v[a][i] = me.v[a] + hdtm * acc.f[a];
#pragma unroll
    for (int k = 0; k < (DIM == 3 ? 6 : 3); ++k) S.c[k][i] = c[k] + hdt * dc[k];
  }
}

// ---------------------------------------------------------------------------
// pair pass 2: velocity gradient, dc/dt, forces and external force, fused
// ---------------------------------------------------------------------------
// One walk over the neighbour list built by k_density.  The reference runs
// kernel_vel_gradients (src/kernel_vel_gradients.cu:15-662) and kernel_forces
// (src/kernel_forces.cu:15-865) as separate walks; forces need neighbour
// densities and pressures but not the velocity gradient, so both are gathered
// here from one evaluation of r, e, gradW and the Morris relative velocity.
// DO_GV / DO_F select the parts.  A particle with more neighbours than the list
// holds parks its partial sums in gv/f and joins the queue for k_pairs_tail.
// CST: the simulations' colloid states are staged in LDS (every simulation of the
// batch has at most kCollStage colloids).
template <class R, int DIM, bool WALL, bool DO_GV, bool DO_F, bool CST>
__global__ __launch_bounds__(kBlock, sizeof(R) == 4 ? SPH_PAIRS_WAVES_F32
                                    : WALL ? SPH_PAIRS_WAVES_F64 : SPH_PAIRS_WAVES_F64_PERIODIC) void k_pairs(
    SimTable T, Particles<R> S, Derived<R> D, Bodies Ball,
    const PairRec<R, DIM> *__restrict__ rec,
    NList L, int store_gv, int write_dcdt, int with_ext, int kick, VList VL) {
  const int b = tile_of(T), sim = T.blk_sim[b];
  if (T.halt[sim]) return;  // (the whole workgroup: one simulation)
  const int iend = T.pbeg[sim + 1];
  const int i0 = T.blk_first[b] + threadIdx.x;
  const bool live = i0 < iend;
  const sph_params_t &P = T.P[sim];
  const Bodies B = bodies_of(Ball, T, sim);
  // the colloid states of the block's simulation, staged before any thread leaves
  __shared__ double cst[C_OZ + 1][kCollStage];
  __shared__ uint4 cstq[kCollStage];
  if constexpr (CST) {
    const int nc = P.N_colloids;  // (uniform: one simulation per workgroup)
    if (nc > 0) {
      // (fields unrolled: a run-time field index would put B.s in scratch)
      for (int c = threadIdx.x; c < nc; c += blockDim.x) {
#pragma unroll
        for (int f = 0; f <= C_OZ; ++f) cst[f][c] = B.s[f][c];
        if constexpr (kFixRec<R, DIM>)
          cstq[c] = fix_pos(fix_of(P), B.s[C_X][c], B.s[C_Y][c], B.s[C_Z][c]);
      }
      __syncthreads();
    }
  }
  const auto cv = [&] {
    if constexpr (CST) return CollFromLds{cst, cstq};
    else return CollFromBodies{B, P};
  }();
  // Verlet walk: every lane of a wave takes part in the quad loads, so the
  // lanes past the simulation's end stay (aliasing its last particle for their
  // reads) and store nothing.
  if (!live && !VL.idx) return;
  const int i = live ? i0 : iend - 1;
  const int N = T.ntotal;  // list stride
  PairSelf<R> me;
  load_self<R, DIM>(P, S, D, i, DO_F, me);
  PairAcc<R> acc;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    acc.f[a] = R(0);
    acc.s[a] = R(0);
#pragma unroll
    for (int b = 0; b < 3; ++b) acc.g[a][b] = R(0);
  }
  const PairK<R, DIM> K(P);
  const bool vok = VL.idx && live && VL.count[i] <= VL.cap;
  if (VL.idx) {
    // The Verlet entries the density pass marked, in list order, all lanes of
    // the wave stepping through entry e together (a lane past its list or not
    // on the Verlet path has no hits).  Index and mask loads are coalesced.
    // Records move in two stages: (1) the G lanes of group a (G = 4, or 2 for
    // the two-chunk 2D FP32 record) load the record lane Ga + k needs as
    // 16-byte chunks, lane Ga + b taking chunk b (the 128-byte FP64 3D record
    // takes two instructions, chunks b and 4 + b), so an instruction touches
    // one line per group instead of one per lane; (2) the chunks go through LDS
    // to the lane that needs them.  The group loads of entry e + 1 are issued
    // before entry e is evaluated.  A lane without a hit asks for an offset past
    // the simulation's records, which the buffer returns as zeros without a
    // memory access.
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    using Rec = PairRec<R, DIM>;
    constexpr int kChunks = sizeof(Rec) / 16;          // 2, 4 or 8 (FP64 3D: last is padding)
    constexpr int kG = kChunks < 4 ? kChunks : 4;      // lanes per group
    constexpr int kLoads = kChunks;                    // group instructions per entry
    constexpr int kPer = kChunks / kG;                 // instructions per record
    constexpr int kSlot = kChunks + 1;  // padded LDS slots: conflict-free rows
    // The loader runs kAhead entries ahead of the evaluation (below).  In the
    // FP64 form the group loads go straight into LDS (buffer loads with the
    // LDS as destination: instruction k of a wave writes its 64 lanes' chunks
    // to row k of the wave's buffer, lane after lane), so no registers hold
    // the chunks in flight (64 fewer: three wavefronts per SIMD instead of
    // two, the FP64 step 5 % faster) and nothing is written to LDS by the wave
    // itself.  The compiler does not wait for such loads before reading the
    // LDS they write, so the walk waits itself: for every vector memory load
    // before it reads an entry (vmcnt 0), and for its LDS reads before it
    // reloads the buffer (lgkmcnt 0).  (The FP32 form keeps its loads in
    // registers: its 32 freed registers did not buy a wavefront, and the
    // waits made the pass 5 % slower, 14 % at five wavefronts per SIMD.)
    constexpr int kAhead = kChunks <= 2 ? 2 : 1;
    constexpr bool kDma = SPH_PAIRS_LDS_LOADS && kAhead == 1 && sizeof(R) == 8;
    constexpr int kRow = 64 + 1;  // (a padded row: the four lanes of a group read distinct banks)
    // (the LDS of the walk not taken shrinks to one element)
    constexpr int kStW = kDma ? 1 : kBlock / 64, kStL = kDma ? 1 : 64, kStS = kDma ? 1 : kSlot;
    constexpr int kDbW = kDma ? kBlock / 64 : 1, kDbL = kDma ? kLoads : 1, kDbR = kDma ? kRow : 1;
    __shared__ u4 stage[kStW][kStL][kStS];
    u4(&st)[kStL][kStS] = stage[kDma ? 0 : threadIdx.x >> 6];
    __shared__ u4 dmab[kDbW][kDbL][kDbR];
    u4(&db)[kDbL][kDbR] = dmab[kDma ? threadIdx.x >> 6 : 0];
    const int lane = threadIdx.x & 63, qb = lane & (kG - 1), qa = lane & ~(kG - 1);
    const int vc = vok ? VL.count[i] : 0;
    int vmax = vc;
#pragma unroll
    for (int o = 32; o; o >>= 1) vmax = max(vmax, __shfl_xor(vmax, o));
    vmax = __builtin_amdgcn_readfirstlane(vmax);
    const int4 *__restrict__ vl = VL.idx + i;
    const unsigned *__restrict__ vm = VL.mask + i;
    const RecBuf rb = rec_buf(rec, T, sim);
    const int pb = __builtin_amdgcn_readfirstlane(T.pbeg[sim]);
    const unsigned oob =
        (unsigned)(__builtin_amdgcn_readfirstlane(iend) - pb) * (unsigned)sizeof(Rec);
    // (a lane past its list re-reads its last index group and mask word: every
    // lane walks to the wave's longest list, and new lines would cost traffic)
    const int lastg = max(vc - 1, 0) >> 2, lastw = max(vc - 1, 0) >> 5;
    // q: indices of the loader's current group of four; qn: the following
    // group.  m: hit bits of the current 32 entries, bit 0 the loader's entry
    // (masked to zero past the list); mn, mn2: the
    // following two words (unmasked until used; two ahead so that the word
    // taken over is never one still being loaded).
    int4 q = ld_stream(&vl[0]);
    int4 qn = ld_stream(&vl[(size_t)min(1, lastg) * N]);
    unsigned m = ld_stream(&vm[0]);
    unsigned mn = ld_stream(&vm[(size_t)min(1, lastw) * N]);
    unsigned mn2 = ld_stream(&vm[(size_t)min(2, lastw) * N]);
    if (vc < 32) m &= (1u << vc) - 1u;  // (vc = 0: no hits)
    // instruction k of a group load: the record of lane qa + src(k), chunk chk(k)
    auto src = [](int k) { return k / kPer; };
    auto chk = [&](int k) { return qb + kG * (k % kPer); };
    auto quad_load = [&](unsigned off, u4(&v)[kLoads]) {
      const int mine = (int)off;
#pragma unroll
      for (int k = 0; k < kLoads; ++k) {
        int o;
        if constexpr (kG == 4)
          o = src(k) == 0 ? quad_bcast<0>(mine)
            : src(k) == 1 ? quad_bcast<1>(mine)
            : src(k) == 2 ? quad_bcast<2>(mine) : quad_bcast<3>(mine);
        else  // lane pairs: lanes 2a, 2a + 1 read lane 2a + k
          o = src(k) == 0 ? quad_pick<0xA0>(mine) : quad_pick<0xF5>(mine);
        v[k] = __builtin_amdgcn_raw_buffer_load_b128(rb.rs, o + 16 * chk(k), 0, 0);
      }
    };
    // The loader (q, m) runs kAhead entries ahead of the evaluation: its record
    // loads are in flight for kAhead entries of work.  Two ahead for the
    // two-chunk record (2D FP32: its batched sweep steps 2 % faster); the larger
    // records measured no faster two ahead, for more registers.
    // Entries per trip, and chunk buffers (each entry's buffer is fixed within
    // a trip; only kAhead + 1 of the four are live at a time).  Four entries
    // per trip start every trip at a multiple of four, so that each entry's
    // place in its group of four indices is known at compile time (a rotating
    // group would cost moves per entry, and with three entries per trip copies
    // onto registers still being loaded).
    constexpr int kTrip = 4;
    // loader: from entry k - 1 to entry k (bit 0 of m then describes k; its
    // index is component k & 3 of the group q)
    auto advance = [&](int k) {
      m >>= 1;
      if ((k & 3) == 0) {
        q = qn;
        qn = ld_stream(&vl[(size_t)min((k + 4) >> 2, lastg) * N]);
      }
      if ((k & 31) == 0) {
        m = mn;
        mn = mn2;
        mn2 = ld_stream(&vm[(size_t)min((k + 64) >> 5, lastw) * N]);
        const int left = vc - k;
        if (left < 32) m &= left > 0 ? (1u << left) - 1u : 0u;
      }
    };
    auto load_at = [&](int k, u4(&v)[kLoads]) {
      const int c = k & 3;  // (known at compile time, see kTrip)
      const int j = c == 0 ? q.x : c == 1 ? q.y : c == 2 ? q.z : q.w;
      quad_load((m & 1u) ? (unsigned)j * (unsigned)sizeof(Rec) : oob, v);
    };
    if constexpr (kDma) {
      constexpr int kVm0 = 0x0F70, kLgkm0 = 0xC07F;  // s_waitcnt vmcnt(0) / lgkmcnt(0)
      // lane qa + qb's chunk c sits in row qb kPer + c / kG at lane qa + c % kG
      const u4 *mrow = &db[qb * kPer][qa];
      auto dma_at = [&](int k) {
        const int c = k & 3;  // (known at compile time, see kTrip)
        const int j = c == 0 ? q.x : c == 1 ? q.y : c == 2 ? q.z : q.w;
        const int mine = (int)((m & 1u) ? (unsigned)j * (unsigned)sizeof(Rec) : oob);
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
          const int o = src(u) == 0 ? quad_bcast<0>(mine)
                      : src(u) == 1 ? quad_bcast<1>(mine)
                      : src(u) == 2 ? quad_bcast<2>(mine) : quad_bcast<3>(mine);
          load_to_lds(rb.rs, &db[u][0], o + 16 * chk(u));
        }
      };
      // hit: the hit bit of the entry whose loads are in flight
      unsigned hitb = m & 1u;
      dma_at(0);
      auto entry_lds = [&](int e) {
        const bool hit = hitb;
        __builtin_amdgcn_s_waitcnt(kVm0);  // entry e has landed
        u4 mine[kRecVecs<R, DIM>];
#pragma unroll
        for (int c = 0; c < kRecVecs<R, DIM>; ++c) mine[c] = mrow[(c / kG) * kRow + c % kG];
        __builtin_amdgcn_s_waitcnt(kLgkm0);  // read before the buffer is reloaded
        advance(e + 1);
        hitb = m & 1u;
        dma_at(e + 1);
        if (hit) pair_body<R, DIM, WALL, DO_GV, DO_F>(K, unpack_chunks<R, DIM>(mine), cv, me, acc);
      };
#pragma unroll 1
      for (int e = 0; e < vmax; e += kTrip) {
#pragma unroll
        for (int t = 0; t < kTrip; ++t) entry_lds(e + t);
      }
      // (the loads of the entry past the walk are waited for: nothing of this
      // wave writes the buffer after it leaves)
      __builtin_amdgcn_s_waitcnt(kVm0);
    } else {
    // pend: hit bits of the entries loaded but not yet evaluated (bit 0: the
    // next to evaluate)
    u4 buf[kTrip][kLoads];
    unsigned pend = m & 1u;
    load_at(0, buf[0]);
#pragma unroll
    for (int j = 1; j < kAhead; ++j) {
      advance(j);
      pend |= (m & 1u) << j;
      load_at(j, buf[j]);
    }
    // entry e: load entry e + kAhead into nxt, evaluate the chunks in cur
    auto entry = [&](int e, const u4(&cur)[kLoads], u4(&nxt)[kLoads]) {
      advance(e + kAhead);
      pend |= (m & 1u) << kAhead;
      const bool hit = pend & 1u;
      pend >>= 1;
      load_at(e + kAhead, nxt);
#pragma unroll
      for (int k = 0; k < kLoads; ++k) st[qa + src(k)][chk(k)] = cur[k];
      __builtin_amdgcn_wave_barrier();
      u4 mine[kRecVecs<R, DIM>];
#pragma unroll
      for (int k = 0; k < kRecVecs<R, DIM>; ++k) mine[k] = st[lane][k];
      __builtin_amdgcn_wave_barrier();
      if (hit) pair_body<R, DIM, WALL, DO_GV, DO_F>(K, unpack_chunks<R, DIM>(mine), cv, me, acc);
    };
    // kTrip entries per trip with the chunk buffers rotating roles (no copies).
    // (Always all entries of a trip: with a trip that could end early, the
    // loads of its last entries would still be pending at the loop head and
    // every trip would wait for them before issuing its own.  Entries past
    // vmax are past every lane's list: no hit, no record access.)
#pragma unroll 1
    for (int e = 0; e < vmax; e += kTrip) {
#pragma unroll
      for (int t = 0; t < kTrip; ++t) entry(e + t, buf[t], buf[(t + kAhead) % kTrip]);
    }
    }
    if (!live) return;
    if (vok) {
      pair_finish<R, DIM, DO_GV, DO_F>(P, S, D, rec, i, me, acc, store_gv, write_dcdt,
                                       with_ext, kick);
      return;
    }
  }
  const int total = L.count[i];
  const int cnt = min(total, L.kmax);
  const int4 *__restrict__ nl = L.idx + i;
  const int pbn = T.pbeg[sim];
  const RecBuf rbn = rec_buf(rec, T, sim);
#pragma unroll 1
  for (int k = 0; k < cnt; k += 4) {
    int4 q = nl[(size_t)(k >> 2) * N];
    const int m = min(4, cnt - k);
#pragma unroll 1
    for (int u = 0; u < m; ++u) {
      pair_body<R, DIM, WALL, DO_GV, DO_F>(K, load_rec<R, DIM>(rbn, q.x - pbn), cv, me, acc);
      q = make_int4(q.y, q.z, q.w, 0);
    }
  }
  if (total > cnt) {
    // Park the partial sums this pass computes (and only those: a gradient-only
    // pass leaves the forces of an earlier force pass alone, and a force-only
    // pass the gradient); k_pairs_tail continues from them.  The force is
    // parked with i's own stress term of the pairs so far, the gradient
    // unscaled.
    if (DO_F) add_self_stress<R, DIM>(rec, i, me, acc);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (DO_F) D.f[a][i] = acc.f[a];
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (DO_GV) D.gv[a * 3 + b][i] = acc.g[a][b];
    }
    L.queue[atomicAdd(L.qlen, 1)] = i;
    return;
  }
  pair_finish<R, DIM, DO_GV, DO_F>(P, S, D, rec, i, me, acc, store_gv, write_dcdt, with_ext,
                                   kick);
}

// Particles whose neighbours did not all fit in the list: re-scan the strips
// with the density pass's arithmetic (same hits, same order), skip the hits the
// list held and add the rest, then finish as k_pairs would have.
template <class R, int DIM, bool WALL, bool DO_GV, bool DO_F>
__global__ __launch_bounds__(kBlock) void k_pairs_tail(SimTable T, Particles<R> S,
                                                       Derived<R> D, Bodies Ball,
                                                       const PairRec<R, DIM> *__restrict__ rec,
                                                       NList L,
                                                       const int *__restrict__ csg, int store_gv,
                                                       int write_dcdt, int with_ext, int kick) {
  const int n = *L.qlen;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int i = L.queue[q];
    const int sim = sim_of(T, i);
    const sph_params_t &P = T.P[sim];
    const Bodies B = bodies_of(Ball, T, sim);
    const int *__restrict__ cs = csg + T.cell_off[sim];
    PairSelf<R> me;
    load_self<R, DIM>(P, S, D, i, DO_F, me);
    PairAcc<R> acc;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      acc.f[a] = DO_F ? D.f[a][i] : R(0);
      acc.s[a] = R(0);
#pragma unroll
      for (int b = 0; b < 3; ++b) acc.g[a][b] = DO_GV ? D.gv[a * 3 + b][i] : R(0);
    }
    const R rcutsq = (R)P.rcutsq;
    const PairK<R, DIM> K(P);
    const int skip = L.kmax;
    int hit = 0;
    int cx, cy, cz;
    cell_coords<DIM, WALL>(P, S.x[i], S.y[i], DIM == 3 ? S.z[i] : 0.0, cx, cy, cz);
    CandSelf<R, DIM> cme;
    cme.init(S, i, P);
    for_each_strip<DIM, WALL>(P, cs, cx, cy, cz, [&](int k0, int k1, double sx, double sy,
                                                      double sz) {
      cme.shift(sx, sy, sz);
      for (int k = k0; k < k1; ++k) {
        Cand<R, DIM> c;
        c.load(S, k);
        const R r2 = (R)cme.r2(c);
        if (r2 < rcutsq && k != i) {
          if (hit >= skip)
            pair_body<R, DIM, WALL, DO_GV, DO_F>(K, load_rec(rec, k), CollFromBodies{B, P}, me,
                                                 acc);
          hit++;
        }
      }
    });
    pair_finish<R, DIM, DO_GV, DO_F>(P, S, D, rec, i, me, acc, store_gv, write_dcdt,
                                     with_ext, kick);
  }
}

template <class R, int DIM>
__global__ __launch_bounds__(kBlock) void k_ext(SimTable T, Particles<R> S, Derived<R> D) {
  const int b = blockIdx.x, sim = T.blk_sim[b];
  const sph_params_t &P = T.P[sim];
  const int i = T.blk_first[b] + threadIdx.x;
  if (i >= T.pbeg[sim + 1] || S.type[i] != 0) return;
  R f[3] = {D.f[0][i], D.f[1][i], DIM == 3 ? D.f[2][i] : R(0)};
  ext_force<R, DIM>(P, S.mass[i], S.y[i], f);
  D.f[0][i] = f[0];
  D.f[1][i] = f[1];
  if (DIM == 3) D.f[2][i] = f[2];
}

// ---------------------------------------------------------------------------
// wall and colloid force collection (FP64, deterministic block reductions)
// ---------------------------------------------------------------------------
// Wall forces: each wall's list is cut into chunks of kWallChunk particles, one
// workgroup per chunk (a single workgroup per wall waits out one dependent
// gather per iteration).  Each workgroup writes its partial sum; the last one of
// a wall to finish (a counter) adds the partials in chunk order and resets the
// counter.  The chunking depends only on the wall's particle count, so the sums
// are deterministic and the same in a batch as alone.
constexpr int kWallBlock = 256, kWallChunk = 1024;

// One workgroup of the wall sums: chunk `ch` of wall w of simulation `sim`
// (the list holds global IDs).
template <class R>
__device__ __forceinline__ void walls_block(const SimTable &T, const Derived<R> &D, const Bodies &B,
                                            const int *__restrict__ rank, double *part,
                                            unsigned *cnt, int maxch, int ch, int w, int sim) {
  const int j0 = T.wrange[2 * sim + w], j1 = T.wrange[2 * sim + w + 1];
  const int nch = max(1, (j1 - j0 + kWallChunk - 1) / kWallChunk);
  if (ch >= nch) return;
  const int c0 = j0 + ch * kWallChunk, c1 = min(j1, c0 + kWallChunk);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int j = c0 + threadIdx.x; j < c1; j += kWallBlock) {
    int p = rank[B.wlist[j]];
    acc[0] += (double)D.f[0][p];
    acc[1] += (double)D.f[1][p];
    acc[2] += (double)D.f[2][p];
  }
  block_sum<3, kWallBlock>(acc);
  const int q = 2 * sim + w;
  if (nch == 1) {
    // a wall of one chunk (every 2D wall of a sweep): its workgroup is the only
    // one, so no counter and no device-scope fence (which writes back the L2 of
    // its XCD; thousands of them cost a batch more than the sums)
    if (threadIdx.x != 0) return;
    double *fw = B.fwall + 6 * sim;
    fw[0 * 2 + w] = 0.0 + acc[0];
    fw[1 * 2 + w] = 0.0 + acc[1];
    fw[2 * 2 + w] = T.P[sim].dim == 3 ? 0.0 + acc[2] : 0.0;
    return;
  }
  __shared__ int last;
  if (threadIdx.x == 0) {
    double *pp = part + ((size_t)q * maxch + ch) * 3;
    pp[0] = acc[0];
    pp[1] = acc[1];
    pp[2] = acc[2];
    __threadfence();  // the partial is visible before the count
    last = atomicAdd(&cnt[q], 1u) == (unsigned)(nch - 1);
  }
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  __threadfence();  // the other workgroups' partials are read after their counts
  const double *pp = part + (size_t)q * maxch * 3;
  double t[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nch; ++k)
#pragma unroll
    for (int a = 0; a < 3; ++a) t[a] += pp[3 * k + a];
  cnt[q] = 0u;
  double *fw = B.fwall + 6 * sim;
  fw[0 * 2 + w] = t[0];
  fw[1 * 2 + w] = t[1];
  fw[2 * 2 + w] = T.P[sim].dim == 3 ? t[2] : 0.0;
}

template <class R>
__global__ __launch_bounds__(kWallBlock) void k_walls(SimTable T, Derived<R> D, Bodies B,
                                                      const int *__restrict__ rank, double *part,
                                                      unsigned *cnt, int maxch) {
  // block (chunk, w, s)
  walls_block<R>(T, D, B, rank, part, cnt, maxch, blockIdx.x, blockIdx.y, blockIdx.z);
}

// One workgroup of the colloid pass: colloid g of the batch.
template <class R>
__device__ __forceinline__ void colloid_block(const SimTable &T, const Particles<R> &S,
                                              const Derived<R> &D, const Bodies &Ball,
                                              const int *__restrict__ rank,
                                              const double *__restrict__ mag_all, int kick,
                                              int g) {
  // ci is the colloid's index in its simulation
  const int sim = T.coll_sim[g];
  const sph_params_t &P = T.P[sim];
  const Bodies B = bodies_of(Ball, T, sim);
  const int ci = g - T.coll_off[sim], ncoll = P.N_colloids;
  const bool d3 = P.dim == 3;
  const int j0 = Ball.cstart[g], j1 = Ball.cstart[g + 1];
  const double Xc = B.s[C_X][ci], Yc = B.s[C_Y][ci], Zc = B.s[C_Z][ci];
  // fx fy fz tx ty tz over the colloid's particles, then the repulsion and
  // magnetic forces from the other colloids (one colloid j per thread), all
  // summed by one fixed-order block reduction
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = j0 + threadIdx.x; j < j1; j += kBlock) {
    int p = rank[B.clist[j]];
    double Fx = D.f[0][p], Fy = D.f[1][p], Fz = d3 ? (double)D.f[2][p] : 0.0;
    double rx = mimg(S.x[p] - Xc, P.L[0], 0.5 * P.L[0]);
    double ry = S.y[p] - Yc;
    if (P.wall == 0) ry = mimg(ry, P.L[1], 0.5 * P.L[1]);
    double rz = d3 ? mimg(S.z[p] - Zc, P.L[2], 0.5 * P.L[2]) : 0.0;
    acc[0] += Fx;
    acc[1] += Fy;
    acc[2] += Fz;
    acc[3] += ry * Fz - rz * Fy;
    acc[4] += rz * Fx - rx * Fz;
    acc[5] += rx * Fy - ry * Fx;
  }
  if (ncoll >= 2) {
    const double mag[3] = {mag_all[3 * sim], mag_all[3 * sim + 1], mag_all[3 * sim + 2]};
    // Repulsion and magnetic force from every other colloid within their
    // cut-offs (the set the reference finds through its colloid cell list).
    for (int j = threadIdx.x; j < ncoll; j += kBlock) {
      if (j == ci) continue;
      double r[3];
      r[0] = mimg(Xc - B.s[C_X][j], P.L[0], 0.5 * P.L[0]);
      r[1] = Yc - B.s[C_Y][j];
      if (P.wall == 0) r[1] = mimg(r[1], P.L[1], 0.5 * P.L[1]);
      r[2] = d3 ? mimg(Zc - B.s[C_Z][j], P.L[2], 0.5 * P.L[2]) : 0.0;
      const double r2 = d3 ? r[0] * r[0] + r[1] * r[1] + r[2] * r[2] : r[0] * r[0] + r[1] * r[1];
      double rs2 = r2;
      if (rs2 < P.rcuton_coll_rep_sq) rs2 = P.rcuton_coll_rep_sq;
      if (rs2 <= P.rcutoff_coll_rep_sq) {
        double rr = sqrt(rs2);
        double ts = -P.tau_rep * (rr - 2.0 * P.coll_R) / P.coll_R;
        double ex = exp(ts);
        double Fm = (P.F0_rep * P.tau_rep) * ex / (1.0 - ex);
        for (int a = 0; a < P.dim; ++a) acc[6 + a] += Fm * (r[a] / rr);
      }
      if (r2 <= P.r0_magnet_sq) {
        double rr = d3 ? sqrt(r2) / P.coll_R : sqrt(r2);
        double e[3] = {r[0] / rr, r[1] / rr, d3 ? r[2] / rr : 0.0};
        double me = d3 ? mag[0] * e[0] + mag[1] * e[1] + mag[2] * e[2]
                       : mag[0] * e[0] + mag[1] * e[1];
        double me2 = me * me;
        double F0r4 = P.F0_magnet / (r2 * r2);
        for (int a = 0; a < P.dim; ++a)
          acc[6 + a] += F0r4 * (2.0 * me * mag[a] - (5.0 * me2 - 1.0) * e[a]);
      }
    }
  }
  block_sum<9, kBlock>(acc);
  if (threadIdx.x != 0) return;
  double F[3] = {acc[0] + acc[6], acc[1] + acc[7], acc[2] + acc[8]};
  B.s[C_FX][ci] = F[0];
  B.s[C_FY][ci] = F[1];
  B.s[C_FZ][ci] = d3 ? F[2] : 0.0;
  B.s[C_TX][ci] = d3 ? acc[3] : 0.0;
  B.s[C_TY][ci] = d3 ? acc[4] : 0.0;
  B.s[C_TZ][ci] = acc[5];
  if (kick && P.coll_move == 0) {
    // the colloid's V_Verlet_step2 half-kick (k_coll_vv2's arithmetic) from the
    // force and torque just stored; the other blocks read colloid positions only
    const double hdtm = 0.5 * P.dt / P.coll_mass, hdtI = 0.5 * P.dt / P.coll_I;
    B.s[C_VX][ci] += hdtm * F[0];
    B.s[C_VY][ci] += hdtm * F[1];
    if (d3) B.s[C_VZ][ci] += hdtm * F[2];
    B.s[C_OZ][ci] += hdtI * acc[5];
    if (d3) {
      B.s[C_OX][ci] += hdtI * acc[3];
      B.s[C_OY][ci] += hdtI * acc[4];
    }
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_colloids(SimTable T, Particles<R> S, Derived<R> D,
                                                     Bodies Ball, const int *__restrict__ rank,
                                                     const double *__restrict__ mag_all,
                                                     int kick) {
  colloid_block<R>(T, S, D, Ball, rank, mag_all, kick, blockIdx.x);  // one block per colloid
}

// The wall sums and the colloid pass of a step in one launch (they read the
// forces of disjoint particle sets and write disjoint outputs; both are
// latency-bound, so side by side they take about the time of one): workgroups
// [0, ncoll) are the colloids', the others the wall chunks (chunk fastest,
// then wall, then simulation).
static_assert(kWallBlock == kBlock, "one workgroup size for both");
template <class R>
__global__ __launch_bounds__(kBlock) void k_bodies(SimTable T, Particles<R> S, Derived<R> D,
                                                   Bodies Ball, const int *__restrict__ rank,
                                                   const double *__restrict__ mag_all, int kick,
                                                   int ncoll, double *part, unsigned *cnt,
                                                   int maxch) {
  const int b = blockIdx.x;
  if (b < ncoll) {
    colloid_block<R>(T, S, D, Ball, rank, mag_all, kick, b);
  } else {
    const int k = b - ncoll;
    walls_block<R>(T, D, Ball, rank, part, cnt, maxch, k % maxch, (k / maxch) & 1,
                   k / (2 * maxch));
  }
}

// ---------------------------------------------------------------------------
// integrators
// ---------------------------------------------------------------------------
// True when the particle at sorted position i has moved more than skin / 2 from
// its position at the last rebuild (minimum image; float), or its position is
// not finite (the rebuild's hash then flags its simulation).
__device__ __forceinline__ bool displaced(const VerletCtl &V, const sph_params_t &P, int i,
                                          double x, double y, double z) {
  const float4 r = V.ref[i];
  float dx = (float)x - r.x, dy = (float)y - r.y, dz = P.dim == 3 ? (float)z - r.z : 0.0f;
  const float Lx = (float)P.L[0], Ly = (float)P.L[1], Lz = (float)P.L[2];
  dx -= Lx * rintf(dx / Lx);
  if (!P.wall) dy -= Ly * rintf(dy / Ly);
  if (P.dim == 3) dz -= Lz * rintf(dz / Lz);
  const float lim = V.half_skin_frac * (float)P.rcut;
  return !(dx * dx + dy * dy + dz * dz <= lim * lim);
}

// Raises rebuild[sim] and the batch flag for a displaced particle.
__device__ __forceinline__ void displacement_check(const VerletCtl &V, const sph_params_t &P,
                                                   int sim, int i, double x, double y, double z) {
  if (!V.on || !displaced(V, P, i, x, y, z)) return;
  atomicOr(&V.rebuild[sim], 1);
  atomicOr(V.any, 1);
}

__device__ __forceinline__ void quat_mul(const double c[4], const double d[4], double r[4]) {
  // kernel_quaternion_product (src/kernel_quaternion_product.cu:15-32)
  r[0] = c[0] * d[0] - c[1] * d[1] - c[2] * d[2] - c[3] * d[3];
  r[1] = c[0] * d[1] + c[1] * d[0] + c[2] * d[3] - c[3] * d[2];
  r[2] = c[0] * d[2] - c[1] * d[3] + c[2] * d[0] + c[3] * d[1];
  r[3] = c[0] * d[3] + c[1] * d[2] - c[2] * d[1] + c[3] * d[0];
}

// Colloid half-kick, drift and rotation (one thread per colloid of the batch;
// simulations with coll_move != 0 keep their colloids fixed).  A 3D colloid
// with |omega| == 0 keeps its orientation (the reference divides by zero).
__device__ __forceinline__ void coll_vv1_of(const SimTable &T, Bodies B, int i) {
  const sph_params_t &P = T.P[T.coll_sim[i]];
  if (P.coll_move != 0 || T.halt[T.coll_sim[i]]) return;
  double **C = B.s;
  const bool d3 = P.dim == 3;
  const double hdtm = 0.5 * P.dt / P.coll_mass, hdtI = 0.5 * P.dt / P.coll_I;
  C[C_VX][i] += hdtm * C[C_FX][i];
  C[C_VY][i] += hdtm * C[C_FY][i];
  if (d3) C[C_VZ][i] += hdtm * C[C_FZ][i];
  C[C_OZ][i] += hdtI * C[C_TZ][i];
  double omega;
  if (!d3) {
    omega = C[C_OZ][i];
  } else {
    C[C_OX][i] += hdtI * C[C_TX][i];
    C[C_OY][i] += hdtI * C[C_TY][i];
    omega = sqrt(C[C_OX][i] * C[C_OX][i] + C[C_OY][i] * C[C_OY][i] + C[C_OZ][i] * C[C_OZ][i]);
  }
  double cx = C[C_X][i] + C[C_VX][i] * P.dt;
  double cy = C[C_Y][i] + C[C_VY][i] * P.dt;
  double cz = d3 ? C[C_Z][i] + C[C_VZ][i] * P.dt : 0.0;
  if (cx < 0) cx = cx + P.L[0];
  if (cx > P.L[0]) cx = cx - P.L[0];
  if (P.wall == 0) {
    if (cy < 0) cy = cy + P.L[1];
    if (cy > P.L[1]) cy = cy - P.L[1];
  }
  if (d3) {
    if (cz < 0) cz = cz + P.L[2];
    if (cz > P.L[2]) cz = cz - P.L[2];
  }
  C[C_X][i] = cx;
  C[C_Y][i] = cy;
  if (d3) C[C_Z][i] = cz;
  if (!d3) {
    C[C_THETA][i] = C[C_THETA][i] + C[C_OZ][i] * P.dt;
  } else {
    double h = 0.5 * omega * P.dt;
    double dq[4] = {1.0, 0.0, 0.0, 0.0};
    if (omega != 0.0) {
      double sh = sin(h);
      dq[0] = cos(h);
      dq[1] = sh * (C[C_OX][i] / omega);
      dq[2] = sh * (C[C_OY][i] / omega);
      dq[3] = sh * (C[C_OZ][i] / omega);
    }
    double q[4] = {C[C_Q0][i], C[C_QX][i], C[C_QY][i], C[C_QZ][i]}, qn[4];
    quat_mul(dq, q, qn);
    C[C_Q0][i] = qn[0];
    C[C_QX][i] = qn[1];
    C[C_QY][i] = qn[2];
    C[C_QZ][i] = qn[3];
  }
}

__global__ void k_coll_vv1(SimTable T, Bodies B, int ncoll) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncoll) coll_vv1_of(T, B, i);
}

template <class R, int DIM, bool WALL>
__global__ __launch_bounds__(kBlock) void k_vv1(SimTable T, Particles<R> S, Derived<R> D,
                                                VerletCtl V, Bodies B, int ncoll) {
  // ncoll > 0: the first ncoll threads of the grid also do k_coll_vv1's work
  // (the colloid bodies; this kernel leaves colloid particles to k_coll_place)
  if (ncoll) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ncoll) coll_vv1_of(T, B, g);
  }
  const int b = blockIdx.x, sim = T.blk_sim[b];
  const sph_params_t &P = T.P[sim];
  const int i = T.blk_first[b] + threadIdx.x;
  if (i >= T.pbeg[sim + 1] || T.halt[sim]) return;
  const int t = S.type[i];
  if (t > 2) return;
  double xi = S.x[i], yi = S.y[i], zi = DIM == 3 ? S.z[i] : 0.0;
  R v[3] = {S.v[0][i], S.v[1][i], DIM == 3 ? S.v[2][i] : R(0)};
  if (t == 0) {
    const R hdt = (R)(0.5 * P.dt);
    const R hdtm = hdt / S.mass[i];
    v[0] = v[0] + hdtm * D.f[0][i];
    v[1] = v[1] + hdtm * D.f[1][i];
    if (DIM == 3) v[2] = v[2] + hdtm * D.f[2][i];
#pragma unroll
    for (int k = 0; k < (DIM == 3 ? 6 : 3); ++k) S.c[k][i] = S.c[k][i] + hdt * D.dcdt[k][i];
  } else {
    v[0] = (R)(t == 1 ? P.V_bottom : P.V_top);
    v[1] = R(0);
    v[2] = R(0);
  }
  S.v[0][i] = v[0];
  S.v[1][i] = v[1];
  if (DIM == 3) S.v[2][i] = v[2];
  xi = xi + (double)v[0] * P.dt;
  yi = yi + (double)v[1] * P.dt;
  if (DIM == 3) zi = zi + (double)v[2] * P.dt;
  if (xi < 0) xi = xi + P.L[0];
  if (!WALL && yi < 0) yi = yi + P.L[1];
  if (DIM == 3 && zi < 0) zi = zi + P.L[2];
  if (xi > P.L[0]) xi = xi - P.L[0];
  if (!WALL && yi > P.L[1]) yi = yi - P.L[1];
  if (DIM == 3 && zi > P.L[2]) zi = zi - P.L[2];
  S.x[i] = xi;
  S.y[i] = yi;
  if (DIM == 3) S.z[i] = zi;
  store_pos(S, P, i, xi, yi, zi);
  if (V.on) {
    // one flag update per wavefront (its lanes belong to one simulation): a
    // moving wall passes the limit with all its particles in the same step
    const bool far = displaced(V, P, i, xi, yi, zi);
    const unsigned long long w = __ballot(far);
    if (far && (int)(threadIdx.x & 63) == __ffsll((long long)w) - 1) {
      atomicOr(&V.rebuild[sim], 1);
      atomicOr(V.any, 1);
    }
  }
}

template <class R, int DIM>
__global__ __launch_bounds__(kBlock) void k_vv2(SimTable T, Particles<R> S, Derived<R> D) {
  const int b = blockIdx.x, sim = T.blk_sim[b];
  const sph_params_t &P = T.P[sim];
  const int i = T.blk_first[b] + threadIdx.x;
  if (i >= T.pbeg[sim + 1] || T.halt[sim] || S.type[i] != 0) return;
  const R hdt = (R)(0.5 * P.dt);
  const R hdtm = hdt / S.mass[i];
  S.v[0][i] = S.v[0][i] + hdtm * D.f[0][i];
  S.v[1][i] = S.v[1][i] + hdtm * D.f[1][i];
  if (DIM == 3) S.v[2][i] = S.v[2][i] + hdtm * D.f[2][i];
#pragma unroll
  for (int k = 0; k < (DIM == 3 ? 6 : 3); ++k) S.c[k][i] = S.c[k][i] + hdt * D.dcdt[k][i];
}



// Rigidly re-places every colloid boundary particle (one thread per list entry).
template <class R>
__global__ __launch_bounds__(kBlock) void k_coll_place(SimTable T, Particles<R> S, Bodies B,
                                                       const int *__restrict__ rank, int nlist,
                                                       VerletCtl V) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlist) return;
  const int ci = B.cowner[j];  // colloid of the batch
  const int sim = T.coll_sim[ci];
  const sph_params_t &P = T.P[sim];
  if (P.coll_move != 0 || T.halt[sim]) return;
  double *const *C = B.s;
  double px, py, pz = 0.0;
  if (P.dim == 2) {
    double th = C[C_THETA][ci];
    double sa = sin(th), ca = cos(th);
    px = B.xc[j] * ca - B.yc[j] * sa;
    py = B.xc[j] * sa + B.yc[j] * ca;
  } else {
    double q[4] = {C[C_Q0][ci], C[C_QX][ci], C[C_QY][ci], C[C_QZ][ci]};
    double qi[4] = {q[0], -q[1], -q[2], -q[3]};
    double pt[4] = {0.0, B.xc[j], B.yc[j], B.zc[j]}, t1[4], t2[4];
    quat_mul(q, pt, t1);
    quat_mul(t1, qi, t2);
    px = t2[1];
    py = t2[2];
    pz = t2[3];
  }
  double X = px + C[C_X][ci], Y = py + C[C_Y][ci], Z = P.dim == 3 ? pz + C[C_Z][ci] : 0.0;
  if (X < 0) X = X + P.L[0];
  if (X > P.L[0]) X = X - P.L[0];
  if (P.wall == 0) {
    if (Y < 0) Y = Y + P.L[1];
    if (Y > P.L[1]) Y = Y - P.L[1];
  }
  if (P.dim == 3) {
    if (Z < 0) Z = Z + P.L[2];
    if (Z > P.L[2]) Z = Z - P.L[2];
  }
  const int p = rank[B.clist[j]];
  S.x[p] = X;
  S.y[p] = Y;
  if (P.dim == 3) S.z[p] = Z;
  store_pos(S, P, p, X, Y, Z);
  displacement_check(V, P, sim, p, X, Y, Z);
}

__global__ void k_coll_vv2(SimTable T, Bodies B, int ncoll) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncoll) return;
  const sph_params_t &P = T.P[T.coll_sim[i]];
  if (P.coll_move != 0 || T.halt[T.coll_sim[i]]) return;
  double **C = B.s;
  const double hdtm = 0.5 * P.dt / P.coll_mass, hdtI = 0.5 * P.dt / P.coll_I;
  C[C_VX][i] += hdtm * C[C_FX][i];
  C[C_VY][i] += hdtm * C[C_FY][i];
  if (P.dim == 3) C[C_VZ][i] += hdtm * C[C_FZ][i];
  C[C_OZ][i] += hdtI * C[C_TZ][i];
  if (P.dim == 3) {
    C[C_OX][i] += hdtI * C[C_TX][i];
    C[C_OY][i] += hdtI * C[C_TY][i];
  }
}

// ---------------------------------------------------------------------------
// kinetic energy: fixed partition + fixed reduction tree -> bitwise repeatable
// ---------------------------------------------------------------------------
// Partial sums of a simulation of n particles: np = min(kKeBlocks, n / kBlock
// rounded up) blocks, block b summing particles b*kBlock + t + k*np*kBlock;
// the partition depends on n only, so a simulation sums identically alone and
// in a batch.
__device__ __forceinline__ int n_parts(int n) { return min(kKeBlocks, (n + kBlock - 1) / kBlock); }

template <class R>
__global__ __launch_bounds__(kBlock) void k_ke_partial(SimTable T, Particles<R> S, double *part) {
  const int sim = blockIdx.y;
  const int p0 = T.pbeg[sim], n = T.pbeg[sim + 1] - p0;
  const int np = n_parts(n);
  if ((int)blockIdx.x >= np) return;
  const bool d3 = T.P[sim].dim == 3;
  double acc[1] = {0.0};
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += np * kBlock) {
    const int i = p0 + k;
    double vx = S.v[0][i], vy = S.v[1][i], vz = d3 ? (double)S.v[2][i] : 0.0;
    acc[0] += 0.5 * (double)S.mass[i] * (vx * vx + vy * vy + vz * vz);
  }
  block_sum<1, kBlock>(acc);
  if (threadIdx.x == 0) part[sim * kKeBlocks + blockIdx.x] = acc[0];
}

// Block s: the kinetic energy of simulation s from its partial sums.
__global__ __launch_bounds__(kBlock) void k_ke_final(SimTable T, const double *part, double *out) {
  const int sim = blockIdx.x;
  const int np = n_parts(T.pbeg[sim + 1] - T.pbeg[sim]);
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < np; i += kBlock) acc[0] += part[sim * kKeBlocks + i];
  block_sum<1, kBlock>(acc);
  if (threadIdx.x == 0) out[sim] = acc[0];
}

// Step counter and magnetic moment m = (cos a, sin a, 0), a = omega (t0 + dt step)
// of every simulation (src/main.cu:290-293, 395-400).  `step` >= 0 sets the
// counter; -1 advances it by one (the form a captured graph replays).
__global__ void k_set_step(SimTable T, int *steps, double *mag, int step) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < T.nsim) set_step_of(T, steps, mag, step, s);
}

// ---------------------------------------------------------------------------
// rheology observables (sweep outputs; no counterpart kernel in the reference)
// ---------------------------------------------------------------------------
// Per block: sum over fluid particles of (cxx - cyy), cxy and the count.
template <class R>
__global__ __launch_bounds__(kBlock) void k_polymer_partial(SimTable T, Particles<R> S,
                                                            double *part) {
  const int sim = blockIdx.y;
  const int p0 = T.pbeg[sim], n = T.pbeg[sim + 1] - p0;
  const int np = n_parts(n);
  if ((int)blockIdx.x >= np) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += np * kBlock) {
    const int i = p0 + k;
    if (S.type[i] != 0) continue;
    acc[0] += (double)S.c[0][i] - (double)S.c[2][i];
    acc[1] += (double)S.c[1][i];
    acc[2] += 1.0;
  }
  block_sum<3, kBlock>(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * (sim * kKeBlocks + blockIdx.x) + k] = acc[k];
}

// Block s: finishes the polymer sums of simulation s and its colloid chain
// orientation, the nematic mean of the in-plane angle of every colloid pair
// closer than the magnetic cut-off.  out[6 s + k]: [0] sum(cxx-cyy) [1] sum(cxy)
// [2] fluid count [3] sum sin(2 phi) [4] sum cos(2 phi) [5] pair count.
__global__ __launch_bounds__(kBlock) void k_observables_final(SimTable T, Bodies Ball,
                                                              const double *part, double *out) {
  const int sim = blockIdx.x;
  const sph_params_t &P = T.P[sim];
  const Bodies B = bodies_of(Ball, T, sim);
  const int nparts = n_parts(T.pbeg[sim + 1] - T.pbeg[sim]);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nparts; b += kBlock)
    for (int k = 0; k < 3; ++k) acc[k] += part[3 * (sim * kKeBlocks + b) + k];
  const int nc = P.N_colloids;
  const long long npairs = (long long)nc * nc;
  for (long long q = threadIdx.x; q < npairs; q += kBlock) {
    const int i = (int)(q / nc), j = (int)(q % nc);
    if (j <= i) continue;
    double dx = mimg(B.s[C_X][j] - B.s[C_X][i], P.L[0], 0.5 * P.L[0]);
    double dy = B.s[C_Y][j] - B.s[C_Y][i];
    if (P.wall == 0) dy = mimg(dy, P.L[1], 0.5 * P.L[1]);
    double dz = P.dim == 3 ? mimg(B.s[C_Z][j] - B.s[C_Z][i], P.L[2], 0.5 * P.L[2]) : 0.0;
    if (dx * dx + dy * dy + dz * dz > P.r0_magnet_sq) continue;
    const double phi = atan2(dy, dx);
    acc[3] += sin(2.0 * phi);
    acc[4] += cos(2.0 * phi);
    acc[5] += 1.0;
  }
  block_sum<6, kBlock>(acc);
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) out[6 * sim + k] = acc[k];
}

// Sums of the neighbour counts of the last density pass and of the Verlet
// list lengths over particles [p0, p1) (instrumentation: the pair work of a
// step).
__global__ __launch_bounds__(kBlock) void k_count_pairs(const int *__restrict__ nl_count,
                                                        const int *__restrict__ vl_count, int p0,
                                                        int p1, unsigned long long *out) {
  unsigned long long a = 0, v = 0;
  for (int i = p0 + blockIdx.x * blockDim.x + threadIdx.x; i < p1; i += gridDim.x * blockDim.x) {
    a += (unsigned)nl_count[i];
    if (vl_count) v += (unsigned)vl_count[i];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    v += __shfl_down(v, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], a);
    atomicAdd(&out[1], v);
  }
}

// ---------------------------------------------------------------------------
// owner of the device state
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    n = count;
    return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
  }
};

// Where one simulation of the batch sits in the concatenated arrays.
struct SimInfo {
  sph_params_t p;
  int pbeg, n;            // particles
  int cell_off;           // cells
  int coll_off;           // colloids
  int clist_off, wlist_off;
};

template <class R>
class DeviceImpl final : public Device {
 public:
  ~DeviceImpl() override {
    // Queued work (sph_step returns with graph launches in flight) must end
    // before the graph execs that hold its kernel arguments, the buffers it
    // reads and the stream it runs on are released.
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (auto &e : events_) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int nsim() const override { return (int)sims_.size(); }

  int upload(const std::vector<const HostState *> &hs, std::string &err) override {
    if (hs.empty()) {
      err = "no simulation to upload";
      return SPH_ERR_ARG;
    }
    P_ = hs[0]->p;
    sims_.clear();
    N_ = ncells_ = ncoll_ = nclist_ = nwlist_ = 0;
    any_walls_ = false;
    coll_lds_ = true;
    for (const HostState *h : hs) {
      const sph_params_t &p = h->p;
      if (p.dim != P_.dim || p.wall != P_.wall) {
        err = "all simulations of a batch must share dim and wall";
        return SPH_ERR_ARG;
      }
      if ((unsigned long long)p.N * sizeof(PairRec<R, 3>) > 0xFFFFFFFFull) {
        // record loads address a simulation's records with 32-bit offsets
        err = "a simulation of " + std::to_string(p.N) + " particles exceeds the " +
              std::to_string(0xFFFFFFFFull / sizeof(PairRec<R, 3>)) +
              "-particle limit of this precision form";
        return SPH_ERR_ARG;
      }
      SimInfo si;
      si.p = p;
      si.pbeg = N_;
      si.n = p.N;
      si.cell_off = ncells_;
      si.coll_off = ncoll_;
      si.clist_off = nclist_;
      si.wlist_off = nwlist_;
      sims_.push_back(si);
      N_ += p.N;
      ncells_ += p.Ntotal_cells;
      ncoll_ += p.N_colloids;
      nclist_ += p.Nlist_colloids;
      nwlist_ += p.wall ? p.Nlist_walls : 0;
      any_walls_ = any_walls_ || (p.wall && p.Nwalls > 0);
      coll_lds_ = coll_lds_ && p.N_colloids <= kCollStage;
    }
    // SPH_COLLOID_LDS=0 keeps the colloid states in global memory (same results)
    if (const char *e = std::getenv("SPH_COLLOID_LDS")) coll_lds_ = coll_lds_ && std::atoi(e) != 0;
    const int S = nsim();
    nc_ = P_.dim == 3 ? 6 : 3;
    max_parts_ = 1;
    for (const SimInfo &si : sims_)
      max_parts_ = std::max(max_parts_, std::min(kKeBlocks, (si.n + kBlock - 1) / kBlock));
    if (!stream_) HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      HIPCHK(x_[b].alloc(N_)); HIPCHK(y_[b].alloc(N_)); HIPCHK(z_[b].alloc(N_));
      for (int a = 0; a < 3; ++a) HIPCHK(v_[b][a].alloc(N_));
      for (int a = 0; a < 6; ++a) HIPCHK(c_[b][a].alloc(N_));
      HIPCHK(mass_[b].alloc(N_)); HIPCHK(type_[b].alloc(N_)); HIPCHK(id_[b].alloc(N_));
    }
    if (sizeof(R) == 4) HIPCHK(pq_.alloc(N_));
    else HIPCHK(pd_.alloc(N_));
    HIPCHK(newpos_.alloc(N_)); HIPCHK(ptmp_.alloc(N_));
    HIPCHK(dens_.alloc(N_)); HIPCHK(press_.alloc(N_));
    for (int a = 0; a < 3; ++a) HIPCHK(f_[a].alloc(N_));
    for (int a = 0; a < 6; ++a) HIPCHK(dcdt_[a].alloc(N_));
    for (int a = 0; a < 9; ++a) HIPCHK(gv_[a].alloc(N_));
    HIPCHK(cell_.alloc(N_)); HIPCHK(slot_.alloc(N_)); HIPCHK(rank_.alloc(N_));
    HIPCHK(stage_src_.alloc(N_)); HIPCHK(stage_id_.alloc(N_)); HIPCHK(sorted_cell_.alloc(N_));
    if (ncells_ > kScanTile * kScanTile) {
      err = "too many cells in the batch (at most 8192 * 8192)";
      return SPH_ERR_ARG;
    }
    // Tile offsets are added by k_scan_add_sums up to this many tiles, above it
    // by k_scan_tile_sums + k_scan_add (SPH_SCAN_SUM_TILES lowers the threshold,
    // so that small cell grids can take the second path too; same results).
    scan_sum_tiles_ = kScanSumTiles;
    if (const char *e = std::getenv("SPH_SCAN_SUM_TILES"))
      scan_sum_tiles_ = std::min(kScanSumTiles, std::max(1, std::atoi(e)));
    HIPCHK(count_.alloc(ncells_)); HIPCHK(cs_.alloc(ncells_ + 1));
    HIPCHK(tsum_.alloc((ncells_ + kScanTile - 1) / kScanTile));
    HIPCHK(flag_.alloc(S));
    HIPCHK(halt_.alloc(S));
    // Neighbour-list capacity: 1.5x the mean number of particles inside the
    // cut-off sphere (disc in 2D) at the initial number density, plus 64, for
    // the densest simulation of the batch.
    kmax_ = 4;
    for (const SimInfo &si : sims_) {
      const sph_params_t &p = si.p;
      double vol = p.L[0] * (p.L[1] + (p.wall ? 2.0 * p.wall_width : 0.0)) *
                   (p.dim == 3 ? p.L[2] : 1.0);
      double n0 = p.N / vol;
      double nb = p.dim == 3 ? n0 * 4.0 / 3.0 * M_PI * p.rcut * p.rcut * p.rcut
                             : n0 * M_PI * p.rcut * p.rcut;
      kmax_ = std::max(kmax_, ((int)(1.5 * nb) + 64 + 31) / 32 * 32);
    }
    // SPH_NEIGHBOUR_CAPACITY overrides it (smaller lists save memory; the pair
    // pass re-scans for particles whose neighbours do not fit).
    if (const char *e = std::getenv("SPH_NEIGHBOUR_CAPACITY")) kmax_ = std::max(1, std::atoi(e));
    kmax_ = (kmax_ + 3) / 4 * 4;
    HIPCHK(nl_idx_.alloc((size_t)(kmax_ / 4) * N_));
    // Verlet lists: skin as a fraction of rcut (SPH_VERLET_SKIN, 0 disables),
    // capacity 1.5x the mean count inside rcut + skin plus 64.
    // 0.04 rcut: measured best on the 1M-particle walls + colloids flow (0.02 to
    // 0.05 within 1 %: shorter lists against more frequent rebuilds).
    skin_frac_ = 0.04f;
    if (const char *e = std::getenv("SPH_VERLET_SKIN")) skin_frac_ = (float)std::atof(e);
    verlet_ok_ = skin_frac_ > 0.0f && skin_frac_ <= 1.0f;
    double vmean = 0.0;
    for (const SimInfo &si : sims_) {
      const sph_params_t &p = si.p;
      if (p.Ncells[0] < 5 || (!p.wall && p.Ncells[1] < 5) || (p.dim == 3 && p.Ncells[2] < 5))
        verlet_ok_ = false;
      double vol = p.L[0] * (p.L[1] + (p.wall ? 2.0 * p.wall_width : 0.0)) *
                   (p.dim == 3 ? p.L[2] : 1.0);
      const double rv = p.rcut * (1.0 + skin_frac_);
      vmean = std::max(vmean, p.N / vol * (p.dim == 3 ? 4.0 / 3.0 * M_PI * rv * rv * rv
                                                       : M_PI * rv * rv));
    }
    // (a multiple of 8: the staged walk's slots come in groups of eight)
    vcap_ = ((int)(1.5 * vmean) + 64 + 7) / 8 * 8;
    if (const char *e = std::getenv("SPH_VERLET_CAPACITY"))
      vcap_ = std::max(8, (std::atoi(e) + 7) / 8 * 8);
    // Staged density walk: 3D, and 2D in the FP32 form (SPH_DENSITY_STAGE=0
    // turns it off, =2 turns it on for the FP64 form in 2D too; SPH_STAGE_CAP
    // lowers the positions a window may hold, so that more tiles take the
    // unstaged walk; results are the same either way).  The 256-simulation 2D
    // sweep batch steps 1-2 % faster staged.
    int stage_mode = 1;
    if (const char *e = std::getenv("SPH_DENSITY_STAGE")) stage_mode = std::atoi(e);
    stage_ = verlet_ok_ && stage_mode != 0 &&
             (P_.dim == 3 || sizeof(R) == 4 || stage_mode == 2);
    stage_cap_ = stage_cap_of<R>(P_.dim);
    if (const char *e = std::getenv("SPH_STAGE_CAP"))
      stage_cap_ = std::min(stage_cap_of<R>(P_.dim), std::max(0, std::atoi(e)));
    if (verlet_ok_) {
      HIPCHK(vl_idx_.alloc((size_t)(vcap_ / 4) * N_));
      if (stage_) {
        HIPCHK(vl_slot_.alloc((size_t)(vcap_ / 8) * N_));
        // window tiles of kWinTile particles from each simulation's first
        std::vector<int> wb(S);
        int nw = 0;
        for (int k = 0; k < S; ++k) {
          wb[k] = nw;
          nw += (sims_[k].n + kWinTile - 1) / kWinTile;
        }
        HIPCHK(wbase_.alloc(S));
        HIPCHK(hipMemcpy(wbase_.p, wb.data(), S * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(vl_win_.alloc((size_t)std::max(nw, 1) * kWinInts));
      }
      HIPCHK(vl_mask_.alloc((size_t)((vcap_ + 31) / 32) * N_));
      HIPCHK(vl_count_.alloc(N_));
      HIPCHK(ref4_.alloc(N_));
      HIPCHK(vrebuild_.alloc(S));
      HIPCHK(voverflow_.alloc(S));
      HIPCHK(hipMemset(vrebuild_.p, 0, S * sizeof(int)));
      HIPCHK(hipMemset(voverflow_.p, 0, S * sizeof(int)));
      HIPCHK(any_.alloc(1));
      HIPCHK(hipMemset(any_.p, 0, sizeof(int)));
    }
    HIPCHK(rebuilds_.alloc(S));
    HIPCHK(hipMemset(rebuilds_.p, 0, S * sizeof(unsigned long long)));
    HIPCHK(count_out_.alloc(2));
    verlet_valid_ = verlet_dirty_ = in_verlet_ = false;
    HIPCHK(nl_count_.alloc(N_));
    HIPCHK(nl_queue_.alloc(N_));
    HIPCHK(nl_qlen_.alloc(1));
    HIPCHK(hipMemset(nl_qlen_.p, 0, sizeof(int)));
    qlen_clear_ = true;
    HIPCHK(dinv_.alloc(N_));
    HIPCHK(rec_.alloc((size_t)N_ * sizeof(PairRec<R, 3>)));  // the larger record
    HIPCHK(ke_part_.alloc((size_t)kKeBlocks * S)); HIPCHK(ke_out_.alloc(S));
    HIPCHK(obs_part_.alloc((size_t)3 * kKeBlocks * S)); HIPCHK(obs_out_.alloc(6 * S));
    for (int k = 0; k < C_NUM; ++k) HIPCHK(coll_[k].alloc(ncoll_));
    HIPCHK(clist_.alloc(nclist_)); HIPCHK(cstart_.alloc(ncoll_ + 1));
    HIPCHK(cowner_.alloc(nclist_));
    HIPCHK(xc_.alloc(nclist_)); HIPCHK(yc_.alloc(nclist_)); HIPCHK(zc_.alloc(nclist_));
    HIPCHK(wlist_.alloc(nwlist_));
    HIPCHK(fwall_.alloc(6 * S));
    HIPCHK(wcnt_.alloc(2 * S));
    HIPCHK(hipMemset(wcnt_.p, 0, 2 * S * sizeof(unsigned)));
    HIPCHK(steps_.alloc(S)); HIPCHK(mag_.alloc(3 * S));

    // particle state, ID order = initial order; IDs are global (pbeg + local)
    std::vector<double> xyz[3];
    std::vector<R> real(N_);
    for (int a = 0; a < 3; ++a) xyz[a].resize(N_);
    std::vector<int> ityp(N_), iota(N_);
    std::iota(iota.begin(), iota.end(), 0);
    for (int s = 0; s < S; ++s) {
      const HostState &h = *hs[s];
      const int b = sims_[s].pbeg;
      for (int i = 0; i < h.p.N; ++i) {
        xyz[0][b + i] = h.pf[F_X][i];
        xyz[1][b + i] = h.pf[F_Y][i];
        xyz[2][b + i] = h.pf[F_Z][i];
        ityp[b + i] = h.type[i];
      }
    }
    HIPCHK(hipMemcpy(x_[0].p, xyz[0].data(), N_ * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(y_[0].p, xyz[1].data(), N_ * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(z_[0].p, xyz[2].data(), N_ * sizeof(double), hipMemcpyHostToDevice));
    auto put_all = [&](R *dst, int field) -> int {
      for (int s = 0; s < S; ++s) {
        const HostState &h = *hs[s];
        for (int i = 0; i < h.p.N; ++i) real[sims_[s].pbeg + i] = (R)h.pf[field][i];
      }
      HIPCHK(hipMemcpy(dst, real.data(), N_ * sizeof(R), hipMemcpyHostToDevice));
      return SPH_OK;
    };
    int rc;
    if ((rc = put_all(v_[0][0].p, F_VX)) || (rc = put_all(v_[0][1].p, F_VY)) ||
        (rc = put_all(v_[0][2].p, F_VZ)))
      return rc;
    static const int cfield[6] = {F_CXX, F_CXY, F_CYY, F_CXZ, F_CYZ, F_CZZ};
    for (int a = 0; a < 6; ++a)
      if ((rc = put_all(c_[0][a].p, cfield[a]))) return rc;
    if ((rc = put_all(mass_[0].p, F_MASS))) return rc;
    HIPCHK(hipMemcpy(type_[0].p, ityp.data(), N_ * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(id_[0].p, iota.data(), N_ * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(rank_.p, iota.data(), N_ * sizeof(int), hipMemcpyHostToDevice));
    id_host_ = iota;
    // derived fields as the host state holds them: zero after initialize
    // (reference: cudaMemset, system_copy_pointers_to_device.cpp:190-256), the
    // saved values after a checkpoint is read (the next VV1 kicks with the
    // saved forces and dc/dt)
    static const int ffield[3] = {F_FX, F_FY, F_FZ};
    for (int a = 0; a < 3; ++a)
      if ((rc = put_all(f_[a].p, ffield[a]))) return rc;
    for (int a = 0; a < 6; ++a)
      if ((rc = put_all(dcdt_[a].p, F_DCDT_XX + a))) return rc;
    for (int a = 0; a < 9; ++a)
      if ((rc = put_all(gv_[a].p, F_GXX + a))) return rc;
    // The second buffer set starts as a copy, so rows a failed reorder leaves
    // unwritten still hold finite values.
    HIPCHK(hipMemcpy(x_[1].p, x_[0].p, N_ * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(y_[1].p, y_[0].p, N_ * sizeof(double), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(z_[1].p, z_[0].p, N_ * sizeof(double), hipMemcpyDeviceToDevice));
    for (int a = 0; a < 3; ++a)
      HIPCHK(hipMemcpy(v_[1][a].p, v_[0][a].p, N_ * sizeof(R), hipMemcpyDeviceToDevice));
    for (int a = 0; a < 6; ++a)
      HIPCHK(hipMemcpy(c_[1][a].p, c_[0][a].p, N_ * sizeof(R), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(mass_[1].p, mass_[0].p, N_ * sizeof(R), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(type_[1].p, type_[0].p, N_ * sizeof(int), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(id_[1].p, id_[0].p, N_ * sizeof(int), hipMemcpyDeviceToDevice));
    if ((rc = put_all(dens_.p, F_DENS)) || (rc = put_all(press_.p, F_PRESS))) return rc;
    HIPCHK(hipMemset(count_.p, 0, ncells_ * sizeof(int)));
    HIPCHK(hipMemset(cs_.p, 0, (ncells_ + 1) * sizeof(int)));
    HIPCHK(hipMemset(flag_.p, 0, S * sizeof(int)));
    HIPCHK(hipMemset(halt_.p, 0, S * sizeof(int)));
    HIPCHK(hipMemset(fwall_.p, 0, 6 * S * sizeof(double)));
    HIPCHK(hipMemset(steps_.p, 0, S * sizeof(int)));
    HIPCHK(hipMemset(mag_.p, 0, 3 * S * sizeof(double)));

    // colloids and lists, concatenated; list entries hold global particle IDs
    std::vector<double> cf(ncoll_);
    for (int k = 0; k < C_NUM; ++k) {
      for (int s = 0; s < S; ++s)
        for (int c = 0; c < hs[s]->p.N_colloids; ++c) cf[sims_[s].coll_off + c] = hs[s]->cf[k][c];
      if (ncoll_)
        HIPCHK(hipMemcpy(coll_[k].p, cf.data(), ncoll_ * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<int> clist(nclist_), cowner(nclist_), cst(ncoll_ + 1), csim(ncoll_),
        coff(S + 1), pbeg(S + 1), coffc(S), wlist(nwlist_), wr(2 * S + 1);
    std::vector<double> xc(nclist_), yc(nclist_), zc(nclist_);
    for (int s = 0; s < S; ++s) {
      const HostState &h = *hs[s];
      const SimInfo &si = sims_[s];
      const int nc = h.p.N_colloids;
      for (int c = 0; c < nc; ++c) {
        const int g = si.coll_off + c;
        csim[g] = s;
        cst[g] = si.clist_off + (h.p.Nlist_colloids ? h.colloids_start[c] : 0);
        const int e = c + 1 < nc ? h.colloids_start[c + 1] : h.p.Nlist_colloids;
        for (int j = h.colloids_start[c]; j < e; ++j) cowner[si.clist_off + j] = g;
      }
      for (int j = 0; j < h.p.Nlist_colloids; ++j) {
        clist[si.clist_off + j] = si.pbeg + h.colloids_list[j];
        xc[si.clist_off + j] = h.x_center[j];
        yc[si.clist_off + j] = h.y_center[j];
        zc[si.clist_off + j] = h.z_center[j];
      }
      const int nw = h.p.wall ? h.p.Nlist_walls : 0;
      for (int j = 0; j < nw; ++j) wlist[si.wlist_off + j] = si.pbeg + h.walls_list[j];
      wr[2 * s] = si.wlist_off + (nw && h.p.Nwalls > 0 ? h.walls_start[0] : 0);
      wr[2 * s + 1] = si.wlist_off + (nw && h.p.Nwalls > 1 ? h.walls_start[1] : nw);
      coff[s] = si.coll_off;
      pbeg[s] = si.pbeg;
      coffc[s] = si.cell_off;
    }
    cst[ncoll_] = nclist_;
    coff[S] = ncoll_;
    pbeg[S] = N_;
    wr[2 * S] = nwlist_;
    auto up = [&](DevBuf<int> &d, const std::vector<int> &v) -> int {
      HIPCHK(d.alloc(v.size()));
      if (!v.empty())
        HIPCHK(hipMemcpy(d.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
      return SPH_OK;
    };
    if ((rc = up(pbeg_, pbeg)) || (rc = up(cell_off_, coffc)) || (rc = up(coll_off_, coff)) ||
        (rc = up(coll_sim_, csim)) || (rc = up(wrange_, wr)))
      return rc;
    // wall-force partial sums: chunks per wall (at least one)
    wmaxch_ = 1;
    for (int q = 0; q < 2 * S; ++q)
      wmaxch_ = std::max(wmaxch_, (wr[q + 1] - wr[q] + kWallChunk - 1) / kWallChunk);
    HIPCHK(wpart_.alloc((size_t)3 * 2 * S * wmaxch_));
    if (nclist_) {
      HIPCHK(hipMemcpy(clist_.p, clist.data(), nclist_ * sizeof(int), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(cowner_.p, cowner.data(), nclist_ * sizeof(int), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(xc_.p, xc.data(), nclist_ * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(yc_.p, yc.data(), nclist_ * sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(zc_.p, zc.data(), nclist_ * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(cstart_.p, cst.data(), (ncoll_ + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (nwlist_)
      HIPCHK(hipMemcpy(wlist_.p, wlist.data(), nwlist_ * sizeof(int), hipMemcpyHostToDevice));
    std::vector<sph_params_t> ps(S);
    for (int s = 0; s < S; ++s) ps[s] = sims_[s].p;
    HIPCHK(simP_.alloc(S));
    HIPCHK(hipMemcpy(simP_.p, ps.data(), S * sizeof(sph_params_t), hipMemcpyHostToDevice));
    if ((rc = build_blocks(err))) return rc;
    cur_ = 0;
    have_cells_ = false;
    k_pack_pos<R><<<blocks(N_), kBlock, 0, stream_>>>(table(), parts(0));  // position copies
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return SPH_OK;
  }

  int set_launch(int block, int xcd, std::string &err) override {
    if (block != 64 && block != 128 && block != 256) {
      err = "block must be 64, 128 or 256 threads";
      return SPH_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(stream_));
    block_ = block;
    xcd_ = xcd ? 1 : 0;
    drop_graphs();
    return sims_.empty() ? SPH_OK : build_blocks(err);
  }

  void get_launch(int *block, int *xcd) const override {
    *block = block_;
    *xcd = xcd_;
  }

  // ---- per-step phases ----
  int neighbours_search(std::string &err) override {
    verlet_valid_ = false;
    return reorder(nullptr, true, err);
  }

  // Hash, scan, reorder.  Without `gate` every simulation is reordered into the
  // other buffer set, which becomes the current one.  With `gate` (a Verlet
  // step) only simulations whose gate[s] is set are reordered, in place
  // (gathered into the other set and copied back), and the batch rebuild flag
  // lets hash, scan and scatter exit at once when there are none.  `carry`
  // moves the derived fields along (not needed inside a step, which recomputes
  // them).
  int reorder(const int *gate, bool carry, std::string &err) {
    Mark m(this, PH_NS);
    Particles<R> S = parts(cur_), T = parts(cur_ ^ 1);
    int *any = gate ? any_.p : nullptr;
    k_hash<<<nblk_, block_, 0, stream_>>>(table(), S.x, S.y, S.z, cell_.p, slot_.p, count_.p,
                                          flag_.p, gate, any);
    const int ntiles = (ncells_ + kScanTile - 1) / kScanTile;
    k_scan_tiles<<<ntiles, 1024, 0, stream_>>>(count_.p, cs_.p, tsum_.p, ncells_, any);
    if (ntiles > scan_sum_tiles_) {
      k_scan_tile_sums<<<1, 1024, 0, stream_>>>(tsum_.p, ntiles, cs_.p, ncells_, any);
      k_scan_add<<<blocks(ncells_), kBlock, 0, stream_>>>(cs_.p, tsum_.p, ncells_, any);
    } else if (ntiles > 1) {
      k_scan_add_sums<<<blocks(ncells_), kBlock, 0, stream_>>>(cs_.p, tsum_.p, ntiles, ncells_,
                                                               any);
    }
    k_scatter<<<nblk_, block_, 0, stream_>>>(table(), cell_.p, slot_.p, cs_.p, S.id,
                                             stage_src_.p, stage_id_.p, gate, any);
    k_rank_gather<R><<<nblk_, block_, 0, stream_>>>(table(), S, T, cell_.p, cs_.p, stage_src_.p,
                                                    stage_id_.p, rank_.p, sorted_cell_.p, nc_,
                                                    gate, carry ? newpos_.p : nullptr,
                                                    gate ? 0 : 1);
    // (gated: the list build that follows copies the rebuilt simulations back)
    HIPCHK(hipGetLastError());
    if (carry) {
      // A stand-alone neighbour search: the fields computed before it (density,
      // pressure, forces, dc/dt, gradient) follow their particles, so any
      // method sequence reads them in the new order.  (Inside a step they are
      // all recomputed after the reorder.)
      std::vector<R *> fields = {dens_.p, press_.p, dinv_.p};
      for (int a = 0; a < 3; ++a) fields.push_back(f_[a].p);
      for (int a = 0; a < 6; ++a) fields.push_back(dcdt_[a].p);
      for (int a = 0; a < 9; ++a) fields.push_back(gv_[a].p);
      for (R *f : fields) {
        k_permute<R><<<blocks(N_), kBlock, 0, stream_>>>(f, ptmp_.p, newpos_.p, N_);
        HIPCHK(hipMemcpyAsync(f, ptmp_.p, N_ * sizeof(R), hipMemcpyDeviceToDevice, stream_));
      }
      HIPCHK(hipGetLastError());
    }
    if (!gate) cur_ ^= 1;
    list_valid_ = false;
    rec_valid_ = false;
    have_cells_ = true;
    id_dirty_ = true;
    return SPH_OK;
  }

  // calculate_density: density, pressure and the neighbour list.
  int density(std::string &err) override { return build_list(1, err); }

  // After steps with Verlet lists the sorted order of a simulation is that of
  // its last rebuild, not of the current cells; a method that scans cells
  // first re-sorts (a full neighbour search: same results, current order).
  int settle(std::string &err) {
    if (in_verlet_ || !verlet_dirty_) return SPH_OK;
    verlet_dirty_ = false;
    return neighbours_search(err);
  }

  int build_list(int write_dens, std::string &err) {
    int rc = settle(err);
    if (rc) return rc;
    Mark m(this, PH_DENS);
    dispatch3([&](auto dimc, auto wallc) {
      constexpr int DIM = decltype(dimc)::value;
      constexpr bool WALL = decltype(wallc)::value;
      k_density<R, DIM, WALL><<<nblk_, block_, 0, stream_>>>(
          table(), parts(cur_), derived(), recs<DIM>(), cs_.p, nlist(), write_dens, vlist(),
          vctl(), in_verlet_ ? 1 : 0);
    });
    HIPCHK(hipGetLastError());
    list_valid_ = true;
    qlen_clear_ = true;  // k_density empties the overflow queue of the next pair pass
    if (write_dens) rec_valid_ = true;
    return SPH_OK;
  }

  // The list depends on positions only; it is rebuilt (without touching the
  // densities) when positions changed since the last density pass.
  int ensure_list(std::string &err) { return list_valid_ ? SPH_OK : build_list(0, err); }

  // calculate_pressure: pressure, 1/d and the pair records from the current densities.
  int pressure(std::string &err) override {
    dispatch3([&](auto dimc, auto) {
      constexpr int DIM = decltype(dimc)::value;
      k_pressure<R, DIM><<<nblk_, block_, 0, stream_>>>(table(), parts(cur_), derived(),
                                                         recs<DIM>());
    });
    HIPCHK(hipGetLastError());
    rec_valid_ = true;
    return SPH_OK;
  }

  template <bool GV, bool F>
  int pairs(int store_gv, int write_dcdt, int with_ext, int phase, std::string &err,
            int kick = 0) {
    int rc = settle(err);
    if (rc) return rc;
    rc = ensure_list(err);
    if (rc) return rc;
    // Records hold positions, velocities and stresses: refresh them if any of
    // those changed since the last density/pressure pass.
    if (!rec_valid_ && (rc = pressure(err))) return rc;
    Mark m(this, phase);
    // Each pair pass starts with an empty overflow queue.  The density pass
    // empties it; a pair pass that reuses a list without one (e.g.
    // calculate_vel_gradient after calculate_forces) empties it first, or its
    // k_pairs_tail would take the earlier pass's particles a second time.
    if (!qlen_clear_) k_fill<<<1, 64, 0, stream_>>>(nl_qlen_.p, 1, 0);
    qlen_clear_ = false;
    dispatch3([&](auto dimc, auto wallc) {
      constexpr int DIM = decltype(dimc)::value;
      constexpr bool WALL = decltype(wallc)::value;
      if (coll_lds_)
        k_pairs<R, DIM, WALL, GV, F, true><<<nblk_, block_, 0, stream_>>>(
            table(), parts(cur_), derived(), bodies(), recs<DIM>(), nlist(), store_gv,
            write_dcdt, with_ext, kick, vlist());
      else
        k_pairs<R, DIM, WALL, GV, F, false><<<nblk_, block_, 0, stream_>>>(
            table(), parts(cur_), derived(), bodies(), recs<DIM>(), nlist(), store_gv,
            write_dcdt, with_ext, kick, vlist());
      k_pairs_tail<R, DIM, WALL, GV, F><<<kTailBlocks, kBlock, 0, stream_>>>(
          table(), parts(cur_), derived(), bodies(), recs<DIM>(), nlist(), cs_.p, store_gv,
          write_dcdt, with_ext, kick);
    });
    HIPCHK(hipGetLastError());
    if (kick) rec_valid_ = false;  // velocities and conformations moved
    return SPH_OK;
  }

  // calculate_vel_gradient: the gradient tensor only.
  int vel_gradient(bool store_gradv, std::string &err) override {
    return pairs<true, false>(store_gradv ? 1 : 0, 0, 0, PH_GRADV, err);
  }

  int pair_forces(bool with_ext, std::string &err) override {
    return pairs<false, true>(0, 0, with_ext ? 1 : 0, PH_FORCES, err);
  }

  // gradient + dc/dt + forces + external force in one walk.
  // `kick`: a step's fused pass also applies V_Verlet_step2 to fluid particles.
  int pairs_all(bool store_gradv, std::string &err, bool kick = false) {
    return pairs<true, true>(store_gradv ? 1 : 0, 1, 1, PH_FORCES, err, kick ? 1 : 0);
  }

  int external_forces(std::string &err) override {
    if (P_.dim == 3)
      k_ext<R, 3><<<nblk_, block_, 0, stream_>>>(table(), parts(cur_), derived());
    else
      k_ext<R, 2><<<nblk_, block_, 0, stream_>>>(table(), parts(cur_), derived());
    HIPCHK(hipGetLastError());
    return SPH_OK;
  }

  int force_walls(std::string &err) override {
    if (!any_walls_) return SPH_OK;
    Mark m(this, PH_BODIES);
    k_walls<R><<<dim3(wmaxch_, 2, nsim()), kWallBlock, 0, stream_>>>(
        table(), derived(), bodies(), rank_.p, wpart_.p, wcnt_.p, wmaxch_);
    HIPCHK(hipGetLastError());
    return SPH_OK;
  }

  int force_colloids(std::string &err) override { return colloid_forces(false, err); }

  // `kick`: a step's colloid pass also applies the colloids' V_Verlet_step2.
  int colloid_forces(bool kick, std::string &err) {
    if (ncoll_ == 0) return SPH_OK;
    Mark m(this, PH_BODIES);
    k_colloids<R><<<ncoll_, kBlock, 0, stream_>>>(table(), parts(cur_), derived(), bodies(),
                                                  rank_.p, mag_.p, kick ? 1 : 0);
    HIPCHK(hipGetLastError());
    return SPH_OK;
  }

  // calculate_forces as one sequence (src/system_calculate_forces.cu:86-170).
  int forces_all(bool store_gradv, std::string &err) override {
    int rc;
    if ((rc = density(err))) return rc;
    if ((rc = pairs_all(store_gradv, err))) return rc;
    if ((rc = force_walls(err))) return rc;
    return force_colloids(err);
  }

  int vv1(std::string &err) override {
    if (!in_verlet_) verlet_valid_ = false;
    Mark m(this, PH_VV1);
    list_valid_ = false;
    rec_valid_ = false;
    const VerletCtl V = vctl();
    // the particle launch also moves the colloids when its grid has a thread each
    const bool fuse = (long long)nblk_ * block_ >= ncoll_;
    dispatch3([&](auto dimc, auto wallc) {
      constexpr int DIM = decltype(dimc)::value;
      constexpr bool WALL = decltype(wallc)::value;
      k_vv1<R, DIM, WALL><<<nblk_, block_, 0, stream_>>>(table(), parts(cur_), derived(), V,
                                                           bodies(), fuse ? ncoll_ : 0);
    });
    if (ncoll_ > 0) {
      if (!fuse) k_coll_vv1<<<blocks(ncoll_), kBlock, 0, stream_>>>(table(), bodies(), ncoll_);
      if (nclist_ > 0)
        k_coll_place<R><<<blocks(nclist_), kBlock, 0, stream_>>>(table(), parts(cur_), bodies(),
                                                                 rank_.p, nclist_, V);
    }
    HIPCHK(hipGetLastError());
    return SPH_OK;
  }

  // calculate_forces and V_Verlet_step2 of a step: the fused pair pass applies
  // the fluid kick as it writes each particle's force and dc/dt (the wall and
  // colloid sums after it read forces and colloid particles only), and the
  // colloid pass the colloids' kick as it writes their force and torque.  Same
  // arithmetic as forces_all + vv2.
  int forces_and_kick(std::string &err) {
    int rc;
    if ((rc = density(err))) return rc;
    if ((rc = pairs_all(false, err, true))) return rc;
    if (any_walls_ && ncoll_ > 0) {
      Mark m(this, PH_BODIES);
      k_bodies<R><<<ncoll_ + wmaxch_ * 2 * nsim(), kBlock, 0, stream_>>>(
          table(), parts(cur_), derived(), bodies(), rank_.p, mag_.p, 1, ncoll_, wpart_.p,
          wcnt_.p, wmaxch_);
      HIPCHK(hipGetLastError());
      return SPH_OK;
    }
    if ((rc = force_walls(err))) return rc;
    return colloid_forces(true, err);
  }

  int vv2(std::string &err) override {
    Mark m(this, PH_VV2);
    rec_valid_ = false;
    if (P_.dim == 3)
      k_vv2<R, 3><<<nblk_, block_, 0, stream_>>>(table(), parts(cur_), derived());
    else
      k_vv2<R, 2><<<nblk_, block_, 0, stream_>>>(table(), parts(cur_), derived());
    if (ncoll_ > 0) k_coll_vv2<<<blocks(ncoll_), kBlock, 0, stream_>>>(table(), bodies(), ncoll_);
    HIPCHK(hipGetLastError());
    return SPH_OK;
  }

  int set_step(int step, std::string &err) override {
    if (ncoll_ == 0) return SPH_OK;  // the moment only enters the colloid forces
    k_set_step<<<blocks(nsim()), kBlock, 0, stream_>>>(table(), steps_.p, mag_.p, step);
    HIPCHK(hipGetLastError());
    return SPH_OK;
  }

  // One step in the order of src/main.cu:361-459.  `step` >= 0 sets the step
  // counter, -1 advances the device counter (captured form).
  int one_step(int step, std::string &err) {
    int rc;
    if (verlet_ok_) {
      in_verlet_ = true;
      rc = one_step_verlet(step, err);
      in_verlet_ = false;
      return rc;
    }
    if ((rc = vv1(err))) return rc;
    if ((rc = reorder(nullptr, false, err))) return rc;
    if ((rc = set_step(step, err))) return rc;
    return forces_and_kick(err);
  }

  // The same step with Verlet lists: the displacement checks of VV1 decide, per
  // simulation, whether this step reorders and rebuilds its lists.
  int one_step_verlet(int step, std::string &err) {
    int rc;
    if ((rc = vv1(err))) return rc;
    if ((rc = reorder(vrebuild_.p, false, err))) return rc;
    {
      Mark m(this, PH_NS);
      dispatch3([&](auto dimc, auto wallc) {
        constexpr int DIM = decltype(dimc)::value;
        constexpr bool WALL = decltype(wallc)::value;
        // with colloids, the launch also sets the step counter and m(t)
        k_verlet_build<DIM, WALL><<<nblk_, block_, 0, stream_>>>(
            table(), parts(cur_), parts(cur_ ^ 1), nc_, any_.p, cs_.p, vlist(), vctl(),
            ncoll_ ? steps_.p : nullptr, mag_.p, step);
      });
      // (the density pass that follows turns overflows into next step's rebuilds)
      HIPCHK(hipGetLastError());
    }
    return forces_and_kick(err);
  }

  int run_steps(int nsteps, int first_step, bool graph, std::string &err) override {
    int rc;
    if (verlet_ok_ && nsteps > 0) {
      if (!verlet_valid_) {  // lists of a different order or positions: rebuild all
        k_fill<<<blocks(nsim()), kBlock, 0, stream_>>>(vrebuild_.p, nsim(), 1);
        k_fill<<<1, kBlock, 0, stream_>>>(any_.p, 1, 1);
        HIPCHK(hipGetLastError());
      }
      verlet_valid_ = true;
      verlet_dirty_ = true;
    }
    if (!graph || timing_) {
      for (int n = 0; n < nsteps; ++n)
        if ((rc = one_step(first_step + n, err))) return rc;
      return SPH_OK;
    }
    if (nsteps > 0 && (rc = set_step(first_step - 1, err))) return rc;
    for (int n = 0; n < nsteps; ++n) {
      // A full reorder flips the buffer set (a Verlet step keeps it), so the
      // captured launches differ by the parity of the current set: one graph
      // per parity, and the parity the step leaves, recorded at capture.
      const int par = cur_;
      if (!exec_[par]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        rc = one_step(-1, err);
        hipError_t e = hipStreamEndCapture(stream_, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) {
          err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
          return SPH_ERR_DEVICE;
        }
        e = hipGraphInstantiate(&exec_[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
          exec_[par] = nullptr;
          err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
          return SPH_ERR_DEVICE;
        }
        exec_next_[par] = cur_;
        cur_ = par;  // capture recorded the launches without running them
      }
      HIPCHK(hipGraphLaunch(exec_[par], stream_));
      cur_ = exec_next_[par];
    }
    // state after a step: positions unchanged since the density pass, velocities moved
    list_valid_ = true;
    rec_valid_ = false;
    have_cells_ = true;
    id_dirty_ = true;
    return SPH_OK;
  }

  int observable_sums(double *out, std::string &err) override {
    const int S = nsim();
    k_polymer_partial<R><<<dim3(max_parts_, S), kBlock, 0, stream_>>>(table(), parts(cur_),
                                                                       obs_part_.p);
    k_observables_final<<<S, kBlock, 0, stream_>>>(table(), bodies(), obs_part_.p, obs_out_.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, obs_out_.p, 6 * S * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return SPH_OK;
  }

  int wall_forces(double *out, std::string &err) override {
    HIPCHK(hipMemcpyAsync(out, fwall_.p, 6 * (size_t)nsim() * sizeof(double),
                          hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return SPH_OK;
  }

  int kinetic_energy(double *ke, std::string &err) override {
    const int S = nsim();
    k_ke_partial<R><<<dim3(max_parts_, S), kBlock, 0, stream_>>>(table(), parts(cur_),
                                                                  ke_part_.p);
    k_ke_final<<<S, kBlock, 0, stream_>>>(table(), ke_part_.p, ke_out_.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ke, ke_out_.p, S * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return SPH_OK;
  }

  int synchronize(std::string &err) override {
    HIPCHK(hipStreamSynchronize(stream_));
    return check_flags(err);
  }

  int check_flags(std::string &err) override {
    std::vector<int> f;
    int rc = poll_flags(f, err);
    if (rc) return rc;
    for (int v : f)
      if (v) {
        err = "kernel cell ranges error: a particle has not a proper cell assigned";
        return SPH_ERR_OUT_OF_GRID;
      }
    return SPH_OK;
  }

  int poll_flags(std::vector<int> &flags, std::string &err) override {
    const int S = nsim();
    flags.assign(S, 0);
    HIPCHK(hipMemcpyAsync(flags.data(), flag_.p, S * sizeof(int), hipMemcpyDeviceToHost,
                          stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int v : flags)
      if (v) {
        HIPCHK(hipMemsetAsync(flag_.p, 0, S * sizeof(int), stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        break;
      }
    return SPH_OK;
  }

  // ---- field exchange (ID order on the host side) ----
  int download_field(int sim, HostState &hs, const std::string &name,
                     std::string &err) override {
    HIPCHK(hipStreamSynchronize(stream_));
    int rc = refresh_ids(err);
    if (rc) return rc;
    const SimInfo &si = sims_[sim];
    const int b = si.pbeg, n = si.n;
    for (int k = 0; k < F_NUM; ++k) {
      if (name != kPFieldNames[k]) continue;
      std::vector<double> &dst = hs.pf[k];
      dst.assign(std::max(hs.p.Nmax, n), 0.0);
      if (k == F_X || k == F_Y || k == F_Z) {
        std::vector<double> tmp(n);
        const double *src = k == F_X ? x_[cur_].p : (k == F_Y ? y_[cur_].p : z_[cur_].p);
        HIPCHK(hipMemcpy(tmp.data(), src + b, n * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < n; ++p)
          if (const int l = local_id(b, n, p); l >= 0) dst[l] = tmp[p];
        return SPH_OK;
      }
      const R *src = real_ptr(k);
      std::vector<R> tmp(n);
      HIPCHK(hipMemcpy(tmp.data(), src + b, n * sizeof(R), hipMemcpyDeviceToHost));
      for (int p = 0; p < n; ++p)
        if (const int l = local_id(b, n, p); l >= 0) dst[l] = (double)tmp[p];
      return SPH_OK;
    }
    if (name == "type") {
      std::vector<int> tmp(n);
      HIPCHK(hipMemcpy(tmp.data(), type_[cur_].p + b, n * sizeof(int), hipMemcpyDeviceToHost));
      hs.type.assign(std::max(hs.p.Nmax, n), 0);
      for (int p = 0; p < n; ++p)
        if (const int l = local_id(b, n, p); l >= 0) hs.type[l] = tmp[p];
      return SPH_OK;
    }
    for (int k = 0; k < C_NUM; ++k) {
      if (name != kCFieldNames[k]) continue;
      const int nc = si.p.N_colloids;
      hs.cf[k].assign(nc, 0.0);
      if (nc)
        HIPCHK(hipMemcpy(hs.cf[k].data(), coll_[k].p + si.coll_off, nc * sizeof(double),
                         hipMemcpyDeviceToHost));
      return SPH_OK;
    }
    if (name == "fx_wall" || name == "fy_wall" || name == "fz_wall") {
      double fw[6];
      HIPCHK(hipMemcpy(fw, fwall_.p + 6 * sim, 6 * sizeof(double), hipMemcpyDeviceToHost));
      int a = name[1] - 'x';
      hs.fwall[a][0] = fw[a * 2 + 0];
      hs.fwall[a][1] = fw[a * 2 + 1];
      return SPH_OK;
    }
    return SPH_OK;  // lists and body offsets never change on the device
  }

  int upload_field(int sim, const HostState &hs, const std::string &name,
                   std::string &err) override {
    HIPCHK(hipStreamSynchronize(stream_));
    // new host state releases a held simulation (the next hash holds it again
    // if a particle is still outside the grid)
    HIPCHK(hipMemset(halt_.p + sim, 0, sizeof(int)));
    rec_valid_ = false;
    int rc = refresh_ids(err);
    if (rc) return rc;
    const SimInfo &si = sims_[sim];
    const int b = si.pbeg, n = si.n;
    for (int k = 0; k < F_NUM; ++k) {
      if (name != kPFieldNames[k]) continue;
      if (k == F_X || k == F_Y || k == F_Z) {
        verlet_valid_ = false;
        std::vector<double> tmp(n);
        for (int p = 0; p < n; ++p) {
          const int l = local_id(b, n, p);
          tmp[p] = l >= 0 ? hs.pf[k][l] : 0.0;
        }
        double *dst = k == F_X ? x_[cur_].p : (k == F_Y ? y_[cur_].p : z_[cur_].p);
        HIPCHK(hipMemcpy(dst + b, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
        k_pack_pos<R><<<blocks(N_), kBlock, 0, stream_>>>(table(), parts(cur_));
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream_));
        list_valid_ = false;
        return SPH_OK;
      }
      std::vector<R> tmp(n);
      for (int p = 0; p < n; ++p) {
        const int l = local_id(b, n, p);
        tmp[p] = l >= 0 ? (R)hs.pf[k][l] : R(0);
      }
      HIPCHK(hipMemcpy(real_ptr(k) + b, tmp.data(), n * sizeof(R), hipMemcpyHostToDevice));
      return SPH_OK;
    }
    if (name == "type") {
      list_valid_ = false;
      verlet_valid_ = false;
      std::vector<int> tmp(n);
      for (int p = 0; p < n; ++p) {
        const int l = local_id(b, n, p);
        tmp[p] = l >= 0 ? hs.type[l] : 0;
      }
      HIPCHK(hipMemcpy(type_[cur_].p + b, tmp.data(), n * sizeof(int), hipMemcpyHostToDevice));
      return SPH_OK;
    }
    for (int k = 0; k < C_NUM; ++k) {
      if (name != kCFieldNames[k]) continue;
      const int nc = si.p.N_colloids;
      if (nc)
        HIPCHK(hipMemcpy(coll_[k].p + si.coll_off, hs.cf[k].data(), nc * sizeof(double),
                         hipMemcpyHostToDevice));
      return SPH_OK;
    }
    err = "field '" + name + "' cannot be written on the device";
    return SPH_ERR_ARG;
  }

  int cell_list(int sim, int32_t *pc, int32_t *pi, int32_t *cs, int32_t *ce,
                std::string &err) override {
    if (!have_cells_) {
      err = "no neighbours_search has run";
      return SPH_ERR_STATE;
    }
    if (verlet_dirty_) {
      // Steps with Verlet lists reorder only at rebuilds: the cell list of the
      // current positions needs a full neighbour search.
      int rc = settle(err);
      if (rc) return rc;
      HIPCHK(hipStreamSynchronize(stream_));
    }
    HIPCHK(hipStreamSynchronize(stream_));
    const SimInfo &si = sims_[sim];
    const int b = si.pbeg, n = si.n, nc = si.p.Ntotal_cells;
    if (pc) {
      HIPCHK(hipMemcpy(pc, sorted_cell_.p + b, n * sizeof(int), hipMemcpyDeviceToHost));
      for (int p = 0; p < n; ++p) pc[p] -= si.cell_off;
    }
    if (pi) {
      HIPCHK(hipMemcpy(pi, id_[cur_].p + b, n * sizeof(int), hipMemcpyDeviceToHost));
      for (int p = 0; p < n; ++p) pi[p] -= b;
    }
    if (cs || ce) {
      std::vector<int> tmp(nc + 1);
      HIPCHK(hipMemcpy(tmp.data(), cs_.p + si.cell_off, tmp.size() * sizeof(int),
                       hipMemcpyDeviceToHost));
      for (int c = 0; c < nc; ++c) {
        if (cs) cs[c] = tmp[c] - b;
        if (ce) ce[c] = tmp[c + 1] - 1 - b;
      }
    }
    return SPH_OK;
  }

  int counters(int sim, int64_t out[3], std::string &err) override {
    const SimInfo &si = sims_[sim];
    unsigned long long r = 0, sums[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&r, rebuilds_.p + sim, sizeof(r), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemsetAsync(count_out_.p, 0, 2 * sizeof(unsigned long long), stream_));
    k_count_pairs<<<std::min(1024, blocks(si.n)), kBlock, 0, stream_>>>(
        nl_count_.p, verlet_ok_ ? vl_count_.p : nullptr, si.pbeg, si.pbeg + si.n, count_out_.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sums, count_out_.p, sizeof(sums), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    out[0] = (int64_t)r;
    out[1] = (int64_t)sums[0];
    out[2] = (int64_t)sums[1];
    return SPH_OK;
  }

  void set_timing(bool on) override { timing_ = on; }

  int get_timing(double ms[8], int64_t calls[8], std::string &err) override {
    HIPCHK(hipStreamSynchronize(stream_));
    for (int k = 0; k < 8; ++k) { ms[k] = 0.0; calls[k] = 0; }
    for (auto &e : events_) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, e.a, e.b));
      ms[e.phase] += t;
      calls[e.phase] += 1;
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    events_.clear();
    return SPH_OK;
  }

 private:
  struct Ev { hipEvent_t a, b; int phase; };
  struct Mark {  // brackets one phase with events when timing is on
    DeviceImpl *d;
    int phase;
    hipEvent_t a = nullptr, b = nullptr;
    Mark(DeviceImpl *dd, int ph) : d(dd), phase(ph) {
      if (!d->timing_) return;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
      (void)hipEventRecord(a, d->stream_);
    }
    ~Mark() {
      if (!a) return;
      (void)hipEventRecord(b, d->stream_);
      d->events_.push_back({a, b, phase});
    }
  };

  static int blocks(int n) { return std::max(1, (n + kBlock - 1) / kBlock); }

  // Tiles of block_ particles that never straddle two simulations.
  int build_blocks(std::string &err) {
    std::vector<int> bs, bf;
    for (int s = 0; s < nsim(); ++s) {
      const SimInfo &si = sims_[s];
      const int nb = std::max(1, (si.n + block_ - 1) / block_);
      for (int k = 0; k < nb; ++k) {
        bs.push_back(s);
        bf.push_back(si.pbeg + k * block_);
      }
    }
    nblk_ = (int)bs.size();
    // A single 3D simulation with the XCD-aware launch: each XCD's slab of
    // tiles (a stretch of z cell layers, see xcd_tile) is walked in y bands -
    // every layer's tiles of the first quarter of the y rows, then of the
    // second, ... - instead of layer after layer.  The ~128 tiles an XCD holds
    // at a time then span ~4 layers x 1/4 of the rows instead of one whole
    // layer, and the neighbour data they read (their L2 working set) shrinks
    // by ~40 %.  The y row of a tile is estimated from its first particle's
    // place in the cell-major order.  Any order is correct: tiles are
    // independent.  (SPH_TILE_ORDER=k: k bands; 0: layer after layer.)
    int bands = 4;
    if (const char *e = std::getenv("SPH_TILE_ORDER")) bands = std::max(0, std::atoi(e));
    if (bands && nsim() == 1 && P_.dim == 3 && xcd_ && sims_[0].n > 0) {
      const long long nc = (long long)P_.Ncells[0] * P_.Ncells[1] * P_.Ncells[2];
      const int n = sims_[0].n, Nx = P_.Ncells[0], Ny = P_.Ncells[1];
      const int q = nblk_ / kXcds, r = nblk_ % kXcds;
      std::vector<int> order(nblk_);
      for (int x = 0; x < kXcds; ++x) {
        const int s0 = x * q + std::min(x, r), len = q + (x < r ? 1 : 0);
        std::vector<std::pair<int, int>> key(len);  // (band, tile)
        for (int k = 0; k < len; ++k) {
          const int t = s0 + k;
          const long long c = (long long)(bf[t] - sims_[0].pbeg) * nc / n;
          const int cy = (int)((c / Nx) % Ny);
          key[k] = {cy * bands / Ny, t};
        }
        std::stable_sort(key.begin(), key.end());
        for (int k = 0; k < len; ++k) order[s0 + k] = key[k].second;
      }
      std::vector<int> bs2(nblk_), bf2(nblk_);
      for (int j = 0; j < nblk_; ++j) {
        bs2[j] = bs[order[j]];
        bf2[j] = bf[order[j]];
      }
      bs.swap(bs2);
      bf.swap(bf2);
    }
    HIPCHK(blk_sim_.alloc(nblk_));
    HIPCHK(blk_first_.alloc(nblk_));
    HIPCHK(hipMemcpy(blk_sim_.p, bs.data(), nblk_ * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(blk_first_.p, bf.data(), nblk_ * sizeof(int), hipMemcpyHostToDevice));
    return SPH_OK;
  }

  void drop_graphs() {
    for (auto &e : exec_)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  }

  SimTable table() const {
    SimTable T;
    T.P = simP_.p;
    T.blk_sim = blk_sim_.p;
    T.blk_first = blk_first_.p;
    T.pbeg = pbeg_.p;
    T.cell_off = cell_off_.p;
    T.coll_off = coll_off_.p;
    T.coll_sim = coll_sim_.p;
    T.wrange = wrange_.p;
    T.halt = halt_.p;
    T.nsim = (int)sims_.size();
    T.nblk = nblk_;
    T.ntotal = N_;
    T.xcd = xcd_;
    return T;
  }

  template <class Fn>
  void dispatch3(Fn &&fn) {
    std::integral_constant<int, 2> d2;
    std::integral_constant<int, 3> d3;
    std::true_type w1;
    std::false_type w0;
    if (P_.dim == 3) {
      if (P_.wall) fn(d3, w1); else fn(d3, w0);
    } else {
      if (P_.wall) fn(d2, w1); else fn(d2, w0);
    }
  }

  Particles<R> parts(int b) {
    Particles<R> S;
    S.x = x_[b].p; S.y = y_[b].p; S.z = z_[b].p;
    for (int a = 0; a < 3; ++a) S.v[a] = v_[b][a].p;
    for (int a = 0; a < 6; ++a) S.c[a] = c_[b][a].p;
    S.mass = mass_[b].p; S.type = type_[b].p; S.id = id_[b].p;
    S.pd = pd_.p;
    S.pq = pq_.p;
    return S;
  }
  NList nlist() { return NList{nl_idx_.p, nl_count_.p, nl_queue_.p, nl_qlen_.p, kmax_}; }
  // The Verlet lists while a Verlet step runs; no lists otherwise.
  VList vlist() {
    VList v{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, skin_frac_};
    if (in_verlet_) {
      v.idx = vl_idx_.p;
      v.count = vl_count_.p;
      v.mask = vl_mask_.p;
      v.cap = vcap_;
      if (stage_) {
        v.slot = vl_slot_.p;
        v.win = vl_win_.p;
        v.wbase = wbase_.p;
        v.stage_cap = stage_cap_;
      }
    }
    return v;
  }
  VerletCtl vctl() {
    VerletCtl v{vrebuild_.p, voverflow_.p, ref4_.p, 0.49f * skin_frac_, in_verlet_ ? 1 : 0,
                any_.p, rebuilds_.p};
    return v;
  }

  Derived<R> derived() {
    Derived<R> D;
    D.dens = dens_.p; D.press = press_.p; D.dinv = dinv_.p;
    for (int a = 0; a < 3; ++a) D.f[a] = f_[a].p;
    for (int a = 0; a < 6; ++a) D.dcdt[a] = dcdt_[a].p;
    for (int a = 0; a < 9; ++a) D.gv[a] = gv_[a].p;
    return D;
  }
  Bodies bodies() {
    Bodies B;
    for (int k = 0; k < C_NUM; ++k) B.s[k] = coll_[k].p;
    B.clist = clist_.p; B.cstart = cstart_.p; B.cowner = cowner_.p;
    B.xc = xc_.p; B.yc = yc_.p; B.zc = zc_.p;
    B.wlist = wlist_.p; B.fwall = fwall_.p;
    return B;
  }

  R *real_ptr(int k) {
    switch (k) {
      case F_VX: return v_[cur_][0].p;
      case F_VY: return v_[cur_][1].p;
      case F_VZ: return v_[cur_][2].p;
      case F_MASS: return mass_[cur_].p;
      case F_DENS: return dens_.p;
      case F_PRESS: return press_.p;
      case F_FX: return f_[0].p;
      case F_FY: return f_[1].p;
      case F_FZ: return f_[2].p;
      case F_CXX: return c_[cur_][0].p;
      case F_CXY: return c_[cur_][1].p;
      case F_CYY: return c_[cur_][2].p;
      case F_CXZ: return c_[cur_][3].p;
      case F_CYZ: return c_[cur_][4].p;
      case F_CZZ: return c_[cur_][5].p;
      default:
        if (k >= F_DCDT_XX && k <= F_DCDT_ZZ) return dcdt_[k - F_DCDT_XX].p;
        if (k >= F_GXX && k <= F_GZZ) return gv_[k - F_GXX].p;
        return nullptr;
    }
  }

  // ID within its simulation of the particle at sorted position b + p, or -1
  // if that slot holds another simulation's particle (after a failed reorder).
  int local_id(int b, int n, int p) const {
    const int l = id_host_[b + p] - b;
    return l >= 0 && l < n ? l : -1;
  }

  int refresh_ids(std::string &err) {
    if (!id_dirty_) return SPH_OK;
    id_host_.resize(N_);
    HIPCHK(hipMemcpy(id_host_.data(), id_[cur_].p, N_ * sizeof(int), hipMemcpyDeviceToHost));
    id_dirty_ = false;
    return SPH_OK;
  }

  sph_params_t P_{};  // of simulation 0: dim and wall select the kernel variants
  std::vector<SimInfo> sims_;
  int N_ = 0, ncells_ = 0, ncoll_ = 0, nclist_ = 0, nwlist_ = 0;
  bool any_walls_ = false;
  bool coll_lds_ = true;  // every simulation's colloids fit the pair pass's LDS stage
  int nc_ = 3, cur_ = 0;
  int scan_sum_tiles_ = kScanSumTiles;
  // Threads per workgroup of the density pass: kBlock, or 1024 (four tiles
  // per workgroup; only with kBlock tiles).  SPH_DENSITY_BLOCK selects it.
  int block_ = kBlock, xcd_ = 1, nblk_ = 0, max_parts_ = 1;
  // Verlet lists (used by run_steps): enabled when skin > 0 and every periodic
  // axis of every simulation has at least 5 cells.
  bool verlet_ok_ = false, verlet_valid_ = false, verlet_dirty_ = false, in_verlet_ = false;
  float skin_frac_ = 0.04f;
  int vcap_ = 0;
  DevBuf<int4> vl_idx_;
  DevBuf<uint4> vl_slot_;  // VList::slot
  DevBuf<int> vl_win_;     // VList::win
  DevBuf<int> wbase_;      // VList::wbase
  bool stage_ = false;
  int stage_cap_ = 0;
  DevBuf<unsigned> vl_mask_;
  DevBuf<int> vl_count_, vrebuild_, voverflow_;
  DevBuf<int> any_;  // [1] batch rebuild flag (VerletCtl::any)
  DevBuf<unsigned long long> rebuilds_;   // [nsim] Verlet-list builds (VerletCtl::rebuilds)
  DevBuf<unsigned long long> count_out_;  // [2] k_count_pairs sums
  DevBuf<float4> ref4_;
  bool have_cells_ = false, id_dirty_ = false, timing_ = false;
  hipStream_t stream_ = nullptr;
  hipGraphExec_t exec_[2] = {nullptr, nullptr};
  int exec_next_[2] = {0, 0};  // the buffer parity each captured step leaves
  std::vector<Ev> events_;
  std::vector<int> id_host_;

  DevBuf<double> x_[2], y_[2], z_[2];
  DevBuf<R> v_[2][3], c_[2][6], mass_[2];
  DevBuf<int> type_[2], id_[2];
  DevBuf<double4> pd_;
  DevBuf<uint4> pq_;
  DevBuf<int> newpos_;  // old sorted position -> new, of the last stand-alone reorder
  DevBuf<R> ptmp_;
  DevBuf<R> dens_, press_, dinv_, f_[3], dcdt_[6], gv_[9];
  DevBuf<unsigned char> rec_;  // pair records, PairRec<R, dim> each
  template <int DIM>
  PairRec<R, DIM> *recs() { return reinterpret_cast<PairRec<R, DIM> *>(rec_.p); }
  bool rec_valid_ = false;
  DevBuf<int> cell_, slot_, rank_, stage_src_, stage_id_, sorted_cell_, count_, cs_, tsum_, flag_;
  DevBuf<int> halt_;  // [nsim] SimTable::halt
  DevBuf<int4> nl_idx_;
  DevBuf<int> nl_count_, nl_queue_, nl_qlen_;
  int kmax_ = 0;
  bool list_valid_ = false;
  bool qlen_clear_ = true;  // the overflow queue is empty for the next pair pass
  DevBuf<double> ke_part_, ke_out_, obs_part_, obs_out_;
  DevBuf<double> coll_[C_NUM], xc_, yc_, zc_, fwall_, mag_;
  DevBuf<double> wpart_;   // wall-force partial sums, 3 per chunk (k_walls)
  DevBuf<unsigned> wcnt_;  // [2 nsim] chunks of each wall summed so far (k_walls)
  int wmaxch_ = 1;         // chunks of the longest wall list
  DevBuf<int> clist_, cstart_, cowner_, wlist_, steps_;
  DevBuf<sph_params_t> simP_;
  DevBuf<int> blk_sim_, blk_first_, pbeg_, cell_off_, coll_off_, coll_sim_, wrange_;
};

}  // namespace

std::unique_ptr<Device> make_device(int precision) {
  if (precision == SPH_FP32) return std::unique_ptr<Device>(new DeviceImpl<float>());
  return std::unique_ptr<Device>(new DeviceImpl<double>());
}

}  // namespace sph

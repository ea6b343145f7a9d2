// sph_noalloc.hip - sph_neighbours_search_noalloc: the cell list of one
// simulation computed from device positions into device outputs with no
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  Every
// intermediate array lives in the file-scope `scratch` array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sph_cells.h"

#pragma clang fp contract(on)

namespace sph {
namespace {

constexpr int kBlock = kCellBlock;

// ---------------------------------------------------------------------------
// allocation-free neighbour search (sph_neighbours_search_noalloc)
// ---------------------------------------------------------------------------
// The cell list of one simulation from device positions, with every
// intermediate array in the file-scope `scratch` array below instead of
// allocated buffers.  The same stable (cell, ID) order as the handle's
// neighbour search: count per cell (k_na_hash), exclusive scan
// (k_scan_tiles, k_scan_add_sums), scatter by slot (k_na_scatter), rank within
// the cell by ID (k_na_rank); the reference's colloid cell grid in one
// workgroup (k_na_colloids).
constexpr int kNaMaxParticles = SPH_NOALLOC_MAX_PARTICLES;
constexpr int kNaMaxCells = SPH_NOALLOC_MAX_CELLS;
constexpr int kNaMaxColloids = SPH_NOALLOC_MAX_COLLOIDS;
constexpr int kNaTiles = (kNaMaxCells + kScanTile - 1) / kScanTile;
static_assert(kNaTiles <= kScanSumTiles, "k_scan_add_sums handles every tile count");
// int offsets of the parts of the scratch array
constexpr size_t kNaSlot = 0;                                // [particles] slot within its cell
constexpr size_t kNaStage = kNaSlot + kNaMaxParticles;       // [particles] ID by scatter position
constexpr size_t kNaCount = kNaStage + kNaMaxParticles;      // [cells] particles per cell
constexpr size_t kNaCs = kNaCount + kNaMaxCells;             // [cells + 1] exclusive scan
constexpr size_t kNaTsum = kNaCs + kNaMaxCells + 1;          // [tiles] scan tile totals
constexpr size_t kNaCcell = kNaTsum + kNaTiles;              // [colloids] colloid cells
constexpr size_t kNaInts = kNaCcell + kNaMaxColloids;

__global__ __launch_bounds__(kBlock) void k_na_clear(int *count, int ncells, int *status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncells) count[i] = 0;
  if (i == 0 && status) *status = SPH_OK;
}

// kernel_cell_list (src/kernel_cell_list.cu:15-63) + the per-cell count.
__global__ __launch_bounds__(kBlock) void k_na_hash(sph_params_t P, const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    const double *__restrict__ z, int *count,
                                                    int *slot, int *status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -2;  // lanes past the last particle
  if (i < P.N) {
    bool ok;
    c = cell_of(P, x[i], y[i], P.dim == 3 ? z[i] : 0.0, ok);
    if (!ok && status) *status = SPH_ERR_OUT_OF_GRID;
  }
  const int s = count_into_cell(c, count);  // (every lane of the wave takes part)
  if (c >= 0) slot[i] = s;
}

// Sorted cell of every position and the particle there before the in-cell
// rank (the cell of a position is the same whatever the order within it).
__global__ __launch_bounds__(kBlock) void k_na_scatter(sph_params_t P, const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       const double *__restrict__ z,
                                                       const int *__restrict__ cs,
                                                       const int *__restrict__ slot, int *stage,
                                                       int *particle_cell) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.N) return;
  bool ok;
  const int c = cell_of(P, x[i], y[i], P.dim == 3 ? z[i] : 0.0, ok);
  const int p = cs[c] + slot[i];
  stage[p] = i;
  particle_cell[p] = c;
}

// Final position = cell start + number of particles of the cell with a smaller
// ID: the stable order of the reference's sort_by_key
// (src/system_neighbours_search.cu:56-58), which starts from particle_index[i] = i.
__global__ __launch_bounds__(kBlock) void k_na_rank(int n, const int *__restrict__ cs,
                                                    const int *__restrict__ stage,
                                                    const int *__restrict__ particle_cell,
                                                    int *particle_index) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c = particle_cell[p];
  const int s0 = cs[c], s1 = cs[c + 1], id = stage[p];
  int r = 0;
  for (int q = s0; q < s1; ++q) r += stage[q] < id;
  particle_index[s0 + r] = id;
}

// kernel_cell_ranges (src/kernel_cell_ranges.cu:13-33) from the scan: every cell
// is written (an empty one gets start > end).
__global__ __launch_bounds__(kBlock) void k_na_ranges(int ncells, const int *__restrict__ cs,
                                                      int *cell_start, int *cell_end) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  cell_start[c] = cs[c];
  cell_end[c] = cs[c + 1] - 1;
}

// The colloid cell grid of src/system_neighbours_search.cu:70-106 in one
// workgroup: kernel_initialize_cell_colloids (empty = -1 / -2), the unclamped
// hash of kernel_cell_colloids_list (a centre outside the colloid grid is
// filed under the nearest cell and reported), the stable sort by (cell, index)
// and kernel_cell_colloids_ranges.
constexpr int kNaCollBlock = 1024;
__global__ __launch_bounds__(kNaCollBlock) void k_na_colloids(
    sph_params_t P, const double *__restrict__ cx, const double *__restrict__ cy,
    const double *__restrict__ cz, int *ccell, int *coll_cell, int *coll_index, int *cstart,
    int *cend, int *status) {
  const int n = P.N_colloids, ncc = P.Ntotal_cells_colloids;
  const int *Nc = P.Ncells_colloids;
  for (int c = threadIdx.x; c < ncc; c += kNaCollBlock) {
    cstart[c] = -1;
    cend[c] = -2;
  }
  for (int j = threadIdx.x; j < n; j += kNaCollBlock) {
    int a = (int)floor(cx[j] / P.cell_colloids_size[0]);
    int b = (int)floor(cy[j] / P.cell_colloids_size[1]);
    int e = P.dim == 3 ? (int)floor(cz[j] / P.cell_colloids_size[2]) : 0;
    const bool ok = a >= 0 && a < Nc[0] && b >= 0 && b < Nc[1] &&
                    (P.dim != 3 || (e >= 0 && e < Nc[2]));
    if (!ok) {
      a = min(max(a, 0), Nc[0] - 1);
      b = min(max(b, 0), Nc[1] - 1);
      e = P.dim == 3 ? min(max(e, 0), Nc[2] - 1) : 0;
      if (status) *status = SPH_ERR_OUT_OF_GRID;
    }
    ccell[j] = P.dim == 3 ? e * Nc[0] * Nc[1] + b * Nc[0] + a : b * Nc[0] + a;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += kNaCollBlock) {
    const int c = ccell[j];
    int r = 0;
    for (int k = 0; k < n; ++k) {
      const int ck = ccell[k];
      r += ck < c || (ck == c && k < j);
    }
    coll_cell[r] = c;
    coll_index[r] = j;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += kNaCollBlock) {
    const int c = coll_cell[r];
    if (r == 0 || coll_cell[r - 1] != c) cstart[c] = r;
    if (r == n - 1 || coll_cell[r + 1] != c) cend[c] = r;
  }
}

}  // namespace
}  // namespace sph

// Scratch of sph_neighbours_search_noalloc: slots, staged IDs, cell counts, the
// cell scan, scan tile totals and colloid cells for the largest simulation it
// accepts (SPH_NOALLOC_MAX_* in sph.h; 9.3 MB).  Each GPU the library runs on
// holds one copy at its own address until the process ends.  Host code never
// uses this name as an address: the kernels get the per-device address that
// hipGetSymbolAddress returns (noalloc_scratch).
static __device__ __align__(256) int scratch[sph::kNaInts];

namespace {

constexpr int kNaMaxDevices = 64;
int *g_noalloc_ptr[kNaMaxDevices] = {};
std::mutex g_noalloc_mu;

// The current device's copy of `scratch`, looked up once per device (with that
// device current, as the caller's first call on it makes it).
int noalloc_scratch(int **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return SPH_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return SPH_ERR_DEVICE;
    g_noalloc_ptr[dev] = static_cast<int *>(q);
  }
  *out = g_noalloc_ptr[dev];
  return SPH_OK;
}

}  // namespace

extern "C" int sph_neighbours_search_noalloc(
    const sph_params_t *p, const double *x, const double *y, const double *z,
    int32_t *particle_cell, int32_t *particle_index, int32_t *cell_start, int32_t *cell_end,
    const double *coll_x, const double *coll_y, const double *coll_z, int32_t *coll_cell,
    int32_t *coll_index, int32_t *coll_cell_start, int32_t *coll_cell_end, int32_t *status,
    void *stream) {
  using namespace sph;
  if (!p || !x || !y || (p->dim == 3 && !z) || !particle_cell || !particle_index ||
      !cell_start || !cell_end || (p->dim != 2 && p->dim != 3))
    return SPH_ERR_ARG;
  const sph_params_t P = *p;
  if (P.N < 1 || P.N > kNaMaxParticles || P.Ntotal_cells < 1 || P.Ntotal_cells > kNaMaxCells)
    return SPH_ERR_CAPACITY;
  const bool colloids = P.N_colloids >= 2;  // (src/system_neighbours_search.cu:70)
  if (colloids) {
    if (P.N_colloids > kNaMaxColloids || P.Ntotal_cells_colloids < 1) return SPH_ERR_CAPACITY;
    if (!coll_x || !coll_y || (P.dim == 3 && !coll_z) || !coll_cell || !coll_index ||
        !coll_cell_start || !coll_cell_end)
      return SPH_ERR_ARG;
  }
  int *base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  int *slot = base + kNaSlot, *stage = base + kNaStage, *count = base + kNaCount;
  int *cs = base + kNaCs, *tsum = base + kNaTsum, *ccell = base + kNaCcell;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = P.N, nc = P.Ntotal_cells;
  const int pb = (n + kBlock - 1) / kBlock, cb = (nc + kBlock - 1) / kBlock;
  const int ntiles = (nc + kScanTile - 1) / kScanTile;
  k_na_clear<<<cb, kBlock, 0, st>>>(count, nc, status);
  k_na_hash<<<pb, kBlock, 0, st>>>(P, x, y, z, count, slot, status);
  k_scan_tiles<<<ntiles, 1024, 0, st>>>(count, cs, tsum, nc, nullptr);
  if (ntiles > 1) k_scan_add_sums<<<cb, kBlock, 0, st>>>(cs, tsum, ntiles, nc, nullptr);
  k_na_scatter<<<pb, kBlock, 0, st>>>(P, x, y, z, cs, slot, stage, particle_cell);
  k_na_rank<<<pb, kBlock, 0, st>>>(n, cs, stage, particle_cell, particle_index);
  k_na_ranges<<<cb, kBlock, 0, st>>>(nc, cs, cell_start, cell_end);
  if (colloids)
    k_na_colloids<<<1, kNaCollBlock, 0, st>>>(P, coll_x, coll_y, coll_z, ccell, coll_cell,
                                              coll_index, coll_cell_start, coll_cell_end,
                                              status);
  return hipGetLastError() == hipSuccess ? SPH_OK : SPH_ERR_DEVICE;
}

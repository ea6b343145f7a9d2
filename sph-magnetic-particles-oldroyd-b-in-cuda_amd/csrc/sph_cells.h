// sph_cells.h - the cell formula, the per-cell count and the exclusive scan of
// the cell counts, shared by the handle's neighbour search (sph_device.hip) and
// the allocation-free one (sph_noalloc.hip).  Each including source file gets
// its own copy of these kernels (internal linkage).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/sph.h"

namespace sph {
namespace {

constexpr int kCellBlock = 256;  // threads per workgroup of the per-cell kernels

// Cell (within its simulation's grid) of a position by kernel_cell_list's
// formula: floor of the IEEE quotient, c == Ncells clamped to Ncells - 1 for x,
// z and for y without walls; with walls y is shifted by wall_width and not
// clamped.  A position outside the grid (NaN included: any int the conversion
// gives) sets ok = false and is filed under the nearest cell.
__device__ __forceinline__ int cell_of(const sph_params_t &P, double x, double y, double z,
                                       bool &ok) {
  int cx = (int)floor(x / P.cell_size[0]);
  if (cx == P.Ncells[0]) cx = cx - 1;
  int cy;
  if (P.wall == 0) {
    cy = (int)floor(y / P.cell_size[1]);
    if (cy == P.Ncells[1]) cy = cy - 1;
  } else {
    cy = (int)floor((y + P.wall_width) / P.cell_size[1]);
  }
  int cz = 0;
  if (P.dim == 3) {
    cz = (int)floor(z / P.cell_size[2]);
    if (cz == P.Ncells[2]) cz = cz - 1;
  }
  ok = cx >= 0 && cx < P.Ncells[0] && cy >= 0 && cy < P.Ncells[1] &&
       (P.dim != 3 || (cz >= 0 && cz < P.Ncells[2]));
  if (!ok) {
    cx = min(max(cx, 0), P.Ncells[0] - 1);
    cy = min(max(cy, 0), P.Ncells[1] - 1);
    cz = P.dim == 3 ? min(max(cz, 0), P.Ncells[2] - 1) : 0;
  }
  return (cz * P.Ncells[1] + cy) * P.Ncells[0] + cx;
}

// Counts a wavefront's particles into their cells (c < 0: lane without a
// particle) and returns each lane's slot among its cell's particles counted so
// far.  Particles arrive in the previous step's cell order, so neighbouring
// lanes mostly share a cell: one atomicAdd per run of equal cells in the
// wavefront, each lane's slot = run base + its position in the run.
__device__ __forceinline__ int count_into_cell(int c, int *count) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(c, 1, 64);
  const bool head = lane == 0 || prev != c;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const int leader = 63 - __clzll(heads & upto);
  const unsigned long long after = heads & ~upto;
  const int run_end = after ? __ffsll((long long)after) - 1 : 64;
  int base = 0;
  if (lane == leader && c >= 0) base = atomicAdd(&count[c], run_end - leader);
  base = __shfl(base, leader, 64);
  return base + (lane - leader);
}

// Exclusive scan of the per-cell counts into cs[0..ncells] in three launches:
// each workgroup scans one tile of kScanTile cells (zeroing the counts for the
// next hash), one workgroup scans the tile totals, and a last pass adds each
// tile's offset.  Every workgroup has 1024 threads owning 8 consecutive values.
constexpr int kScanTile = 8192;

// Exclusive scan of v[0..8) per thread across the workgroup; returns this
// thread's exclusive prefix and the workgroup total in *total.
__device__ __forceinline__ int block_scan8(const int (&v)[8], int *total) {
  __shared__ int wsum[16];
  __shared__ int tot;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  int local = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) local += v[k];
  int incl = local;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int n = __shfl_up(incl, off, 64);
    if (lane >= off) incl += n;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int w = 0; w < 16; ++w) {
      int s = wsum[w];
      wsum[w] = acc;
      acc += s;
    }
    tot = acc;
  }
  __syncthreads();
  *total = tot;
  return wsum[wid] + incl - local;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(int *count, int *cs, int *tsum, int ncells,
                                                     const int *any) {
  if (any && !*any) return;
  const int c0 = blockIdx.x * kScanTile + threadIdx.x * 8;
  // a thread's 8 cells as two 16-byte accesses (c0 is a multiple of 8 ints)
  // where all 8 exist, element by element at the end of the array
  const bool whole = c0 + 8 <= ncells;
  int v[8];
  if (whole) {
    const int4 a = *reinterpret_cast<const int4 *>(count + c0);
    const int4 b = *reinterpret_cast<const int4 *>(count + c0 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = c0 + k < ncells ? count[c0 + k] : 0;
  }
  int total;
  int run = block_scan8(v, &total);
  int o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (whole) {
    *reinterpret_cast<int4 *>(cs + c0) = make_int4(o[0], o[1], o[2], o[3]);
    *reinterpret_cast<int4 *>(cs + c0 + 4) = make_int4(o[4], o[5], o[6], o[7]);
    *reinterpret_cast<int4 *>(count + c0) = make_int4(0, 0, 0, 0);
    *reinterpret_cast<int4 *>(count + c0 + 4) = make_int4(0, 0, 0, 0);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (c0 + k < ncells) {
        cs[c0 + k] = o[k];
        count[c0 + k] = 0;
      }
    }
  }
  if (threadIdx.x == 0) {
    tsum[blockIdx.x] = total;
    if (gridDim.x == 1) cs[ncells] = total;  // one tile: the scan ends here
  }
}

// One workgroup: tile totals -> exclusive tile offsets (at most kScanTile
// tiles), and the grand total into cs[ncells].
__global__ __launch_bounds__(1024) void k_scan_tile_sums(int *tsum, int ntiles, int *cs,
                                                         int ncells, const int *any) {
  if (any && !*any) return;
  const int c0 = threadIdx.x * 8;
  int v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = c0 + k < ntiles ? tsum[c0 + k] : 0;
  int total;
  int run = block_scan8(v, &total);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (c0 + k < ntiles) tsum[c0 + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) cs[ncells] = total;
}

__global__ __launch_bounds__(kCellBlock) void k_scan_add(int *cs, const int *tsum, int ncells,
                                                     const int *any) {
  if (any && !*any) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells && c >= kScanTile) cs[c] += tsum[c / kScanTile];
}

// Up to kScanSumTiles tiles, k_scan_add_sums replaces k_scan_tile_sums +
// k_scan_add: each cell adds the totals of the tiles before its own, and the
// first thread writes the grand total into cs[ncells].
constexpr int kScanSumTiles = 64;
__global__ __launch_bounds__(kCellBlock) void k_scan_add_sums(int *cs, const int *tsum, int ntiles,
                                                          int ncells, const int *any) {
  if (any && !*any) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) {
    int t = 0;
    for (int k = 0; k < ntiles; ++k) t += tsum[k];
    cs[ncells] = t;
  }
  if (c < ncells && c >= kScanTile) {
    const int tile = c / kScanTile;
    int off = 0;
    for (int k = 0; k < tile; ++k) off += tsum[k];
    cs[c] += off;
  }
}

}  // namespace
}  // namespace sph

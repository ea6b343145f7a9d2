// sph_device.h - interface between the C-ABI layer and the gfx950 device code.
//
// One Device object owns every device buffer of a batch of simulations (one
// simulation is a batch of one): particle state in cell-sorted SoA order
// (double-buffered for the reorder), the cell list, the inverse permutation
// ID -> sorted position, colloid rigid-body state and reduction scratch.  The
// simulations' arrays are concatenated; every phase is one launch per kernel
// over the whole batch.  All launches go to one HIP stream owned by the object.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "sph_host.h"

namespace sph {

enum Phase { PH_VV1, PH_NS, PH_DENS, PH_GRADV, PH_FORCES, PH_BODIES, PH_VV2, PH_OTHER, PH_NUM };

class Device {
 public:
  virtual ~Device() = default;
  // Allocates and uploads the host states (ID order becomes the initial
  // order).  All simulations must share dim and wall.
  virtual int upload(const std::vector<const HostState *> &sims, std::string &err) = 0;
  virtual int nsim() const = 0;

  // One phase for every simulation of the batch.
  virtual int neighbours_search(std::string &err) = 0;
  virtual int density(std::string &err) = 0;
  virtual int pressure(std::string &err) = 0;
  virtual int vel_gradient(bool store_gradv, std::string &err) = 0;
  virtual int pair_forces(bool with_ext, std::string &err) = 0;
  virtual int external_forces(std::string &err) = 0;
  virtual int force_walls(std::string &err) = 0;
  virtual int force_colloids(std::string &err) = 0;
  // calculate_forces as one sequence (density, fused pair pass, walls, colloids).
  virtual int forces_all(bool store_gradv, std::string &err) = 0;
  virtual int vv1(std::string &err) = 0;
  virtual int vv2(std::string &err) = 0;
  // Sets every simulation's step counter (magnetic moment at t0 + dt*step).
  virtual int set_step(int step, std::string &err) = 0;
  // nsteps full steps (VV1, neighbour search, step + 1, forces, VV2) of every
  // simulation, queued without host synchronisation.  With `graph` the step is
  // captured once per buffer parity and replayed as a hipGraph.
  virtual int run_steps(int nsteps, int first_step, bool graph, std::string &err) = 0;

  // Per simulation: ke[s] (not checked for finiteness), and out[6 s + k] as below.
  virtual int kinetic_energy(double *ke, std::string &err) = 0;
  // [0] sum(cxx-cyy) [1] sum(cxy) over fluid, [2] fluid count, [3] sum sin 2phi,
  // [4] sum cos 2phi, [5] count of colloid pairs within the magnetic cut-off.
  virtual int observable_sums(double *out, std::string &err) = 0;
  // Wall force sums of every simulation: out[6 s + 2 a + w] is axis a of wall w
  // (0 bottom, 1 top), one transfer for the whole batch.
  virtual int wall_forces(double *out, std::string &err) = 0;
  virtual int synchronize(std::string &err) = 0;
  // Checks the device error flags (a particle outside the grid); resets them.
  virtual int check_flags(std::string &err) = 0;
  // The flags per simulation (bit 0: a particle was outside the grid); resets them.
  virtual int poll_flags(std::vector<int> &flags, std::string &err) = 0;

  // Copies a field of simulation `sim` into `hs` (ID order).  `name` as in sph.h.
  virtual int download_field(int sim, HostState &hs, const std::string &name,
                             std::string &err) = 0;
  virtual int upload_field(int sim, const HostState &hs, const std::string &name,
                           std::string &err) = 0;
  virtual int cell_list(int sim, int32_t *pc, int32_t *pi, int32_t *cs, int32_t *ce,
                        std::string &err) = 0;

  // Launch shape of the per-particle kernels: threads per block (64, 128 or
  // 256) and whether workgroups are remapped XCD by XCD.  Results do not
  // depend on it.
  virtual int set_launch(int block, int xcd, std::string &err) = 0;
  virtual void get_launch(int *block, int *xcd) const = 0;

  virtual void set_timing(bool on) = 0;
  virtual int get_timing(double ms[8], int64_t calls[8], std::string &err) = 0;
  // Work counters of simulation `sim`: [0] Verlet-list rebuilds since the
  // upload, [1] neighbour pairs within rcut of the last density pass (sum over
  // particles of their neighbour counts), [2] Verlet-list entries of the last
  // build (0 without Verlet lists).
  virtual int counters(int sim, int64_t out[3], std::string &err) = 0;
};

// precision: SPH_FP64 or SPH_FP32
std::unique_ptr<Device> make_device(int precision);

}  // namespace sph

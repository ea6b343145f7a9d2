/*
 * sph.h - C-ABI of the MI355X SPH Oldroyd-B solver (one GPU, gfx950).
 *
 * Every entry point replaces one method of the reference's host class
 * `class_system` (src/class_system.h:146-651) or one step of its driver
 * (src/main.cu:93-467).  The reference passes (numBlocks, threadsPerBlock)
 * plus 10-70 raw device pointers to every per-step method; here an opaque
 * handle owns all host and device state, launch geometry lives inside the
 * library, and every call returns an int status (SPH_OK == 0).
 *
 * Data order: the device keeps particles in cell-sorted order (structure of
 * arrays).  Every getter/setter and every output file uses particle-ID order,
 * the order of the reference's arrays, which are never permuted there.
 *
 * Threading: one handle is used by one host thread.  Calls that mirror a
 * reference method are synchronous (results are complete on return), like the
 * reference's cudaDeviceSynchronize after each launch.  sph_step() may leave
 * work queued; sph_synchronize() waits for it.
 */
#ifndef SPH_AMD_H
#define SPH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference: int 0/1 + printf, src/system_read_input.cpp:271-507,
 *      src/system_initialize.cpp:128-261, device assert src/kernel_cell_ranges.cu:20-23) */
enum {
  SPH_OK = 0,
  SPH_ERR_ARG = 1,          /* bad argument / unknown field name / wrong count      */
  SPH_ERR_IO = 2,           /* input or output file cannot be opened                */
  SPH_ERR_PARSE = 3,        /* missing key, dimension mismatch, bad option in input */
  SPH_ERR_GEOMETRY = 4,     /* cells too big, colloid overlap/outside, N > Nmax     */
  SPH_ERR_DEVICE = 5,       /* HIP runtime error or no GPU                          */
  SPH_ERR_OUT_OF_GRID = 6,  /* a particle left the cell grid (reference: assert)    */
  SPH_ERR_NAN = 7,          /* non-finite kinetic energy                            */
  SPH_ERR_STATE = 8,        /* call out of order (e.g. step before initialize)      */
  SPH_ERR_CAPACITY = 9      /* larger than sph_neighbours_search_noalloc accepts    */
};

/* ---- precision forms ---------------------------------------------------------
 * SPH_FP64: every field and all pair arithmetic in double (the reference built
 *           with `#define real double`, src/config.h:15-19).
 * SPH_FP32: positions, time, wall/colloid force-torque and kinetic-energy
 *           reductions in double; velocities, conformation tensor, density,
 *           pressure, forces, dc/dt and all pair arithmetic in float. */
enum { SPH_FP64 = 0, SPH_FP32 = 1 };

/* Physical and geometric parameters.  Replaces the ~52 __constant__ symbols of
 * src/kernel_declare_constants.cu:11-62 (uploaded at src/main.cu:185-236) and
 * the derived constants of src/system_initialize.cpp:47-262.  Filled by
 * sph_initialize(); read with sph_get_params(). */
typedef struct sph_params {
  int32_t dim, N, Nmax, wall, Nwalls;
  int32_t Nxyz[3];                 /* lattice sites per axis, y includes wall layers */
  int32_t Ncells[3], Ntotal_cells;
  int32_t N_colloids, Ncells_colloids[3], Ntotal_cells_colloids;
  int32_t Nlist_walls, Nlist_colloids;
  int32_t ext_force_type, coll_move, new_sim, Nsteps;
  int32_t freq_micro, freq_macro, freq_walls, freq_colloids;
  int32_t pad_;
  double L[3], cell_size[3], cell_colloids_size[3];
  double dt, t0, overlap, rho, c, csq, P0, rho0, eta, zeta, a, b;
  double np, kT, nkT, tau;
  double rcut, rcutsq, rcut_inv, cw, c_gradw, ky, ext_force[3];
  double wall_width, y_bottom, y_top, V_top, V_bottom, beta_max;
  double coll_R, coll_rho, coll_mass, coll_I;
  double coll_rep_cuton, coll_rep_cutoff, F0_rep, tau_rep;
  double rcuton_coll_rep, rcuton_coll_rep_sq, rcutoff_coll_rep, rcutoff_coll_rep_sq;
  double rcutoff_coll, rcutoff_coll_sq;
  double r0_magnet, r0_magnet_sq, F0_magnet, omega_magnet;
} sph_params_t;

typedef struct sph_system *sph_handle;

/* Lifecycle.  No reference counterpart: the reference uses a stack object
 * `class_system sys` (src/main.cu:88) plus loose device pointers freed at
 * src/main.cu:481-553.  sph_create touches no GPU. */
sph_handle sph_create(int precision);
void sph_destroy(sph_handle h);
const char *sph_last_error(sph_handle h);
int sph_precision(sph_handle h);

/* class_system::read_input (src/system_read_input.cpp:18-512): same key/value
 * grammar, every key mandatory, same validation.  Differences: the path is an
 * argument (reference: fixed "./input"), a missing file is an error
 * (reference returns 0, line 58-62), the 3D overlap check uses z for z
 * (reference line 485 uses y). */
int sph_read_input(sph_handle h, const char *path);

/* class_system::constructor + initialize + initialize_particles_new_sim/old_sim
 * (src/system_constructor.cpp:15-108, src/system_initialize.cpp:16-359,
 * src/system_initialize_particles_new_sim.cpp:16-253,
 * src/system_initialize_particles_old_sim.cpp:17-112).  Host only: device
 * buffers are allocated by the first device call.  `restart_path` names the
 * restart file read when new_sim == 1 (reference: fixed "input_particles.in");
 * NULL means "input_particles.in" next to the input file. */
int sph_initialize(sph_handle h, const char *restart_path);

int sph_get_params(sph_handle h, sph_params_t *out);

/* Field access in particle-ID order.  Real-valued fields are exchanged as
 * double (float fields of SPH_FP32 are widened/narrowed), integer fields as
 * int32.  `count` must equal the field's length.  Names: x y z vx vy vz mass
 * dens press fx fy fz cxx cxy cyy cxz cyz czz dcdt_xx dcdt_xy dcdt_yy dcdt_xz
 * dcdt_yz dcdt_zz gradvxx gradvxy gradvxz gradvyx gradvyy gradvyz gradvzx
 * gradvzy gradvzz type (int, length N); coll_x coll_y coll_z coll_vx coll_vy
 * coll_vz coll_omegax coll_omegay coll_omegaz coll_theta coll_quat0 coll_quatx
 * coll_quaty coll_quatz fx_colloid fy_colloid fz_colloid tx_colloid
 * ty_colloid tz_colloid (length N_colloids); fx_wall fy_wall fz_wall (length
 * Nwalls); walls_list colloids_list (int, Nlist_*); walls_start colloids_start
 * (int); x_center y_center z_center (length Nlist_colloids).
 * Replaces the D2H copies of src/system_print_particles.cpp:42-56 and friends. */
int sph_get_field(sph_handle h, const char *name, void *dst, int64_t count);
int sph_set_field(sph_handle h, const char *name, const void *src, int64_t count);

/* Cell list after the last neighbours_search, in the reference's layout
 * (src/system_neighbours_search.cu:40-67): particle_cell (sorted keys, length N),
 * particle_index (particle IDs in sorted order, length N), cell_start and
 * cell_end (length Ntotal_cells; an empty cell has start > end). */
int sph_get_cell_list(sph_handle h, int32_t *particle_cell, int32_t *particle_index,
                      int32_t *cell_start, int32_t *cell_end);

/* ---- per-step methods (each synchronous) ---------------------------------- */
/* class_system::neighbours_search (src/system_neighbours_search.cu:19-110):
 * cell hash, stable (cell, ID) sort, cell ranges, and the particle state
 * reordered into cell order on the device. */
int sph_neighbours_search(sph_handle h);
/* class_system::calculate_density (src/system_calculate_density.cu:15-49) */
int sph_calculate_density(sph_handle h);
/* class_system::calculate_vel_gradient (src/system_calculate_vel_gradient.cu:16-86) */
int sph_calculate_vel_gradient(sph_handle h);
/* class_system::calculate_pressure (src/system_calculate_pressure.cu:15-34) */
int sph_calculate_pressure(sph_handle h);
/* class_system::calculate_external_forces (src/system_calculate_external_forces.cu:15-40) */
int sph_calculate_external_forces(sph_handle h);
/* class_system::calculate_force_walls (src/system_calculate_force_walls.cu:15-38) */
int sph_calculate_force_walls(sph_handle h);
/* class_system::calculate_force_colloids (src/system_calculate_force_colloids.cu:15-97) */
int sph_calculate_force_colloids(sph_handle h);
/* class_system::calculate_forces (src/system_calculate_forces.cu:16-174): density,
 * velocity gradient, pressure, pair forces, external force, dc/dt, wall and
 * colloid force collection, colloid repulsion and magnetic forces. */
int sph_calculate_forces(sph_handle h);
/* class_system::V_Verlet_step1 (src/system_V_Verlet_step1.cu:15-122) */
int sph_V_Verlet_step1(sph_handle h);
/* class_system::V_Verlet_step2 (src/system_V_Verlet_step2.cu:15-94) */
int sph_V_Verlet_step2(sph_handle h);
/* Magnetic moment for time step `step`: m = (cos a, sin a, 0) with
 * a = omega_magnetic * (t0 + dt * step)  (src/main.cu:290-293, 395-400). */
int sph_set_step(sph_handle h, int step);

/* ---- drivers -------------------------------------------------------------- */
/* The initial neighbours_search + calculate_forces before the loop
 * (src/main.cu:298-335). */
int sph_prepare(sph_handle h);
/* `nsteps` iterations of the loop body of src/main.cu:358-459 without output:
 * VV1, neighbours_search, magnetic moment, calculate_forces, VV2.  Starts at
 * the handle's step counter + 1 and advances it.  Asynchronous.  (Inside a
 * step the VV2 kicks are applied by the force and colloid passes as they write
 * each force; the arithmetic is that of the method sequence.) */
int sph_step(sph_handle h, int nsteps);
int sph_synchronize(sph_handle h);
int sph_current_step(sph_handle h);
/* The full reference main(): prepare, output at step 0, Nsteps steps with
 * output at freq_* (src/main.cu:30-561).  Files go to `out_dir`. */
int sph_run(sph_handle h, const char *out_dir);

/* ---- reductions and output ---------------------------------------------- */
/* Total kinetic energy sum_i 0.5 m_i |v_i|^2 over all particles, FP64
 * deterministic reduction (src/kernel_calculate_macro_vars.cu:14-40). */
int sph_kinetic_energy(sph_handle h, double *ke);
/* Rheology observables of the current state (the sweep outputs; the reference
 * has no counterpart - it only writes the raw wall/colloid files):
 *   out[0] kinetic energy (as sph_kinetic_energy)
 *   out[1] first normal-stress difference of the polymer, nkT <cxx - cyy>_fluid
 *   out[2] polymer shear stress, nkT <cxy>_fluid
 *   out[3] bottom-wall shear stress Fx_wall[0] / A, out[4] top wall, with
 *          A = L0 (2D) or L0*L2 (3D); NaN without walls
 *   out[5] colloid chain angle: nematic mean 0.5 atan2(sum sin 2phi, sum cos 2phi)
 *          of the in-plane angle phi of colloid pairs closer than cutoff_magnetic
 *          (NaN when there is no such pair)
 *   out[6] field angle omega_magnetic (t0 + dt step), out[7] phase lag out[6] - out[5];
 *          all three angles reduced modulo pi into (-pi/2, pi/2]. */
int sph_observables(sph_handle h, double out[8]);
/* class_system::print_output (src/system_print_output.cpp:14-80): writes
 * micro-<step>.dat, walls-<step>.dat, colloids-<step>.dat and appends to
 * macro.dat in `out_dir` when step is a multiple of the matching freq_*. */
int sph_print_output(sph_handle h, int step, const char *out_dir);
/* Writes the restart file read by new_sim == 1 (column layout of
 * src/system_initialize_particles_old_sim.cpp:35-63). */
int sph_write_restart(sph_handle h, const char *path);
/* Exact binary checkpoint.  No reference counterpart: the reference's only
 * restart is the input_particles.in above, whose values it parses as float
 * and which leaves out forces, dc/dt and the colloids' velocities, spins and
 * orientations (src/system_initialize.cpp:185-201).  Writes the parameters,
 * the step counter and every particle, colloid and wall array at full
 * precision, in ID order; SPH_ERR_IO when the file cannot be written. */
int sph_checkpoint_write(sph_handle h, const char *path);
/* Replaces the handle's whole state with a checkpoint written by a handle of
 * the same precision form (SPH_ERR_ARG otherwise; SPH_ERR_IO for a missing,
 * foreign or damaged file, leaving the handle as it was).  No read_input or
 * initialize is needed; device state is dropped and rebuilt by the next
 * device call.  sph_step then continues the run from the saved state - its
 * first half kick uses the saved forces and dc/dt, as the uninterrupted run's
 * would - with neighbour lists rebuilt at that step, so the two runs agree to
 * rounding (neighbour sums in another order), not bit for bit. */
int sph_checkpoint_read(sph_handle h, const char *path);

/* ---- instrumentation ----------------------------------------------------- */
/* When enabled, each device phase is bracketed by HIP events on the handle's
 * stream.  Phases: 0 VV1, 1 neighbour search, 2 density, 3 velocity gradient
 * + dc/dt, 4 forces, 5 walls + colloids, 6 VV2 (in sph_step, VV2 is part of
 * phases 4 and 5).  `ms` receives the summed milliseconds per phase since the
 * last reset and `calls` the launch counts. */
int sph_set_timing(sph_handle h, int enable);
int sph_get_timing(sph_handle h, double ms[8], int64_t calls[8]);
/* Work counters (no reference counterpart; they size the work of a step for
 * the roofline): out[0] steps run, out[1] Verlet-list rebuilds since the first
 * device call, out[2] neighbour pairs within rcut found by the last density
 * pass (sum over particles of their neighbour counts), out[3] entries of the
 * Verlet lists of the last rebuild (0 when the step runs without them). */
int sph_get_counters(sph_handle h, int64_t out[4]);

/* ---- batches of independent simulations --------------------------------- */
/* A batch runs many independent simulations of uneven size - a parameter
 * sweep - in one device state: their arrays are concatenated and each phase of
 * the step is one launch over all of them, captured once as a hipGraph and
 * replayed every step.  The reference runs one simulation per process (each
 * one is the loop of src/main.cu:358-459); there is no counterpart API.
 * Every simulation of a batch gives, step by step, the same numbers as its own
 * sph_step run (the same kernels on the same data).  All simulations must share
 * dim and wall; they share the precision of the batch. */
typedef struct sph_batch *sph_batch_handle;
sph_batch_handle sph_batch_create(int precision);
void sph_batch_destroy(sph_batch_handle b);
const char *sph_batch_last_error(sph_batch_handle b);
/* Copies the current state of an initialized system into the batch (before
 * sph_batch_prepare); *index receives its position.  The system handle stays
 * independent of the batch. */
int sph_batch_add(sph_batch_handle b, sph_handle sim, int *index);
int sph_batch_size(sph_batch_handle b);
/* Uploads all simulations, then neighbours_search + calculate_forces at step 0
 * (as sph_prepare). */
int sph_batch_prepare(sph_batch_handle b);
/* `nsteps` steps of every simulation (as sph_step).  Asynchronous. */
int sph_batch_step(sph_batch_handle b, int nsteps);
/* Waits for queued steps.  A simulation that fails (a particle leaves the grid,
 * or its kinetic energy is not finite) is flagged and does not fail the call:
 * the other simulations continue unaffected.  flags[s] (sticky) receives
 * SPH_BATCH_OUT_OF_GRID and/or SPH_BATCH_NAN for simulation s. */
int sph_batch_synchronize(sph_batch_handle b);
enum { SPH_BATCH_OUT_OF_GRID = 1, SPH_BATCH_NAN = 2 };
int sph_batch_failed(sph_batch_handle b, int32_t *flags);
int sph_batch_current_step(sph_batch_handle b);
/* Field of simulation `sim`, ID order, names and counts as sph_get_field. */
int sph_batch_get_field(sph_batch_handle b, int sim, const char *name, void *dst,
                        int64_t count);
/* ke[s] for every simulation s (as sph_kinetic_energy; NaN flags the simulation). */
int sph_batch_kinetic_energy(sph_batch_handle b, double *ke);
/* out[8 s + k]: the sph_observables of simulation s. */
int sph_batch_observables(sph_batch_handle b, double *out);
/* Launch shape of the per-particle kernels: `block` threads per workgroup
 * (64, 128 or 256) and `xcd` = 1 to map consecutive workgroups XCD by XCD.
 * Results do not depend on it. */
int sph_batch_set_launch(sph_batch_handle b, int block, int xcd);
int sph_batch_get_launch(sph_batch_handle b, int *block, int *xcd);
/* Times `steps` steps under each launch shape and keeps the fastest.  The
 * timed steps are real steps of the batch (its step counter advances by
 * 6 * (steps + 2)); since results do not depend on the shape, the trajectory
 * is the one an untuned run would follow. */
int sph_batch_autotune(sph_batch_handle b, int steps, int *block, int *xcd,
                       double *ms_per_step);

/* ---- allocation-free neighbour search ------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory - code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the cell list of sph_neighbours_search (hash, stable
 * (cell, ID) sort, cell ranges) and, when N_colloids >= 2, the reference's
 * colloid cell grid, from positions in device memory into outputs in device
 * memory - the data of class_system::neighbours_search without the launch
 * geometry.  It allocates nothing: its intermediate arrays live in one static
 * device array per GPU (9.3 MB, held until the process ends) and it calls no
 * hipMalloc / hipFree.
 *   p        parameters of an initialized system (sph_get_params); N, the
 *            cell grid and the colloid grid must fit the limits below, else
 *            SPH_ERR_CAPACITY is returned and nothing is launched
 *   x y z    device positions in ID order (z may be NULL in 2D)
 *   particle_cell, particle_index [N], cell_start, cell_end [Ntotal_cells]:
 *            device outputs, the layout of sph_get_cell_list (an empty cell
 *            has start > end)
 *   coll_x coll_y coll_z [N_colloids] device colloid centres; coll_cell,
 *            coll_index [N_colloids], coll_cell_start, coll_cell_end
 *            [Ntotal_cells_colloids] device outputs (an empty cell -1 / -2);
 *            all ignored (may be NULL) when N_colloids < 2
 *   status   NULL or a device int32 that receives SPH_OK, or
 *            SPH_ERR_OUT_OF_GRID when a particle or colloid centre lay
 *            outside its grid (it is filed under the nearest cell)
 *   stream   the hipStream_t to queue on (NULL: the null stream).  The call
 *            only queues: outputs are complete once the stream has run it.
 * Calls that overlap on one GPU (from several threads or streams) share that
 * GPU's scratch array and race: serialise them.  Each call clears what it
 * needs cleared, so consecutive calls of different sizes are independent. */
#define SPH_NOALLOC_MAX_PARTICLES 1100000
#define SPH_NOALLOC_MAX_CELLS 65536
#define SPH_NOALLOC_MAX_COLLOIDS 1000
int sph_neighbours_search_noalloc(const sph_params_t *p, const double *x, const double *y,
                                  const double *z, int32_t *particle_cell,
                                  int32_t *particle_index, int32_t *cell_start,
                                  int32_t *cell_end, const double *coll_x, const double *coll_y,
                                  const double *coll_z, int32_t *coll_cell, int32_t *coll_index,
                                  int32_t *coll_cell_start, int32_t *coll_cell_end,
                                  int32_t *status, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SPH_AMD_H */

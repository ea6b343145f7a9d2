/*
 * sph_oracle.h - FP64 CPU oracle of the SPH Oldroyd-B step (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP path, never part of it: only tests/,
 * the smoke check and the bench's CPU-baseline leg load it.  It restates the
 * reference's per-step kernels (src/kernel_*.cu) as OpenMP host loops in
 * double, in the reference's layout: arrays in particle-ID order, neighbours
 * reached through particle_index after a stable sort of (cell, ID).
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>
#include "../include/sph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system orc_system;

/* Creates an oracle with parameters `p` (as filled by sph_initialize) and all
 * arrays zeroed; load the state with orc_set_field using sph.h field names. */
orc_system *orc_create(const sph_params_t *p);
void orc_destroy(orc_system *s);
int orc_set_field(orc_system *s, const char *name, const void *src, int64_t count);
int orc_get_field(orc_system *s, const char *name, void *dst, int64_t count);
int orc_get_cell_list(orc_system *s, int32_t *particle_cell, int32_t *particle_index,
                      int32_t *cell_start, int32_t *cell_end);

/* Colloid cell list of the last neighbours_search (N_colloids >= 2): sorted
 * colloid cells and colloid indices (length N_colloids), cell start / end
 * (length Ntotal_cells_colloids; an empty cell has -1 / -2,
 * src/kernel_initialize_cell_colloids.cu:15-25). */
int orc_get_colloid_cell_list(orc_system *s, int32_t *coll_cell, int32_t *coll_index,
                              int32_t *coll_cell_start, int32_t *coll_cell_end);
int orc_neighbours_search(orc_system *s);
int orc_calculate_density(orc_system *s);
int orc_calculate_vel_gradient(orc_system *s);
int orc_calculate_pressure(orc_system *s);
int orc_calculate_forces(orc_system *s);
/* The sub-steps of calculate_forces, on the current fields. */
int orc_calculate_pair_forces(orc_system *s);       /* zero + kernel_forces      */
int orc_calculate_external_forces(orc_system *s);   /* kernel_ext_force          */
int orc_calculate_dcdt(orc_system *s);              /* kernel_dcdt from gradv    */
int orc_calculate_force_walls(orc_system *s);       /* kernel_collect_force_walls */
int orc_calculate_force_colloids(orc_system *s);    /* zero, collect, repulsion, magnetic */
int orc_V_Verlet_step1(orc_system *s);
int orc_V_Verlet_step2(orc_system *s);
int orc_set_step(orc_system *s, int step);
/* prepare = first neighbours_search + calculate_forces (src/main.cu:298-335) */
int orc_prepare(orc_system *s);
/* nsteps loop bodies of src/main.cu:358-459 from the internal step counter */
int orc_step(orc_system *s, int nsteps);
double orc_kinetic_energy(orc_system *s);
/* The sweep observables, defined as for sph_observables in sph.h. */
int orc_observables(orc_system *s, double out[8]);
int orc_num_threads(void);

/* A planted mistake, for the tests that show a pin catches it (never set
 * otherwise): ORC_SLIP_EOS_NO_MASS drops the particle mass from the equation
 * of state, p = c^2 (d - rho0).  0 removes it. */
enum { ORC_SLIP_NONE = 0, ORC_SLIP_EOS_NO_MASS = 1 };
int orc_set_slip(orc_system *s, int slip);

/* Stand-alone pieces for closed-form checks. */
double orc_kernel_W(const sph_params_t *p, double r);
double orc_kernel_gradW(const sph_params_t *p, double r);
/* O(N^2) density over all pairs with minimum image (the commented-out loop of
 * src/kernel_densities.cu:167-191, with the periodic wrap of the cell loop). */
int orc_density_brute_force(orc_system *s, double *out);

#ifdef __cplusplus
}
#endif
#endif

/*
 * sph_surrogate.h - the sweep surrogate: an ensemble of MLPs on the matrix cores.
 *
 * No counterpart in the reference, which has no learned part.  An ensemble of
 * E members, each 4 inputs -> 3 hidden layers of 256 SiLU units -> 3 outputs,
 * with BF16 weights and activations on v_mfma_f32_16x16x32_bf16 and FP32
 * accumulation; FP32 master weights with Adam (beta1 0.9, beta2 0.999,
 * eps 1e-8).  It maps the standardised (Wi, Mn, phi, V) of a sweep point to its
 * standardised (wall stress, N1, phase lag).
 *
 * Parameters: one flat FP32 vector of E * P values, member e at e * P.  Within a
 * member, W1..W4 then b1..b4 (offsets and padded shapes from sph_mlp_layout):
 * W_l is row-major [rows x cols] with W1 64 columns wide (inputs padded with
 * zeros beyond column 4) and W4 64 rows tall (outputs beyond row 3 are zero
 * padding); b4 has 64 entries.  Padding entries must be zero and stay zero.
 *
 * Batches: `batch` rows per member, a positive multiple of 64.  x is
 * [E][batch][4], t and mask [E][batch][3], y [E][batch][3] (all FP32, row-major).
 * The loss of member e is sum(mask (y - t)^2) / max(1, sum(mask)).
 * Calls return the SPH_* status codes of sph.h.
 */
#ifndef SPH_SURROGATE_H
#define SPH_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sph_mlp *sph_mlp_handle;

/* E members (1..64), weights and biases uniform in +-1/sqrt(fan_in) from
 * `seed` (padding zero).  NULL on a bad argument or when no GPU is available. */
sph_mlp_handle sph_mlp_create(int members, uint64_t seed);
void sph_mlp_destroy(sph_mlp_handle h);
const char *sph_mlp_last_error(sph_mlp_handle h);
/* E * P */
int64_t sph_mlp_param_count(sph_mlp_handle h);
/* offsets[0..3] of W1..W4, offsets[4..7] of b1..b4 within a member;
 * shapes[2l], shapes[2l+1] = rows, cols of W_(l+1). */
int sph_mlp_layout(sph_mlp_handle h, int64_t offsets[8], int32_t shapes[8]);
/* Replaces the FP32 masters (and their BF16 copies); resets Adam. */
int sph_mlp_set_params(sph_mlp_handle h, const float *params, int64_t count);
int sph_mlp_get_params(sph_mlp_handle h, float *params, int64_t count);
/* y = f_e(x_e) for every member e. */
int sph_mlp_forward(sph_mlp_handle h, int batch, const float *x, float *y);
/* Loss per member (loss[E]) and its gradient w.r.t. every parameter
 * (grads[E * P]) on one batch, without updating. */
int sph_mlp_gradients(sph_mlp_handle h, int batch, const float *x, const float *t,
                      const float *mask, float *grads, float *loss);
/* One Adam step of every member on its own batch; loss[E] before the step. */
int sph_mlp_train_step(sph_mlp_handle h, int batch, const float *x, const float *t,
                       const float *mask, float lr, float *loss);
int sph_mlp_synchronize(sph_mlp_handle h);

#ifdef __cplusplus
}
#endif

#endif /* SPH_SURROGATE_H */

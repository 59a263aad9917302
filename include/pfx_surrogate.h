/*
 * pfx_surrogate.h -- learned warm start of the batched reaction step.
 *
 * No counterpart in the reference: RReact starts its Newton iteration from
 * the previous free-ion molalities (reaction.F90:3370), or from 1e-9 molal
 * after constraint speciation (reaction.F90:1481).  These entry points add
 * a multilayer perceptron that maps a cell's log transported totals, its
 * previous log free ions, temperature and mineral volume fractions to the
 * change of the log free-ion molalities over the step; the previous free
 * ions moved by that change are the initial iterate of pfx_react_batch with
 * PFX_GUESS_ARRAY_FALLBACK.  The FP64 Newton from RReact's own starting
 * point (opts->fallback_previous) or the default guess stays the fallback,
 * and every returned state passes the same FP64 convergence test.
 *
 * Matrix products run in BF16 on the matrix cores with FP32 accumulation;
 * master weights, gradients and Adam moments are FP32.  All array arguments
 * are DEVICE pointers unless stated; `stream` is a hipStream_t (NULL =
 * default stream).  Return 0 on success or a PFX_E_* code.
 *
 * Network layout: dims[0] inputs, dims[1..nlayers-1] hidden widths (GELU),
 * dims[nlayers] outputs (linear).  Weight matrix of layer l is
 * [dims[l+1]][dims[l]] row-major.
 *
 * Features of a cell (pfx_surrogate_features, pfx_surrogate_guess), N
 * primaries, K minerals, slots:
 *   [0, N)        slog(total[j])         slog(v) = sign(v) log10(1 + |v| 1e12)
 *   [N, 2N)       ln(pri_molal[j])       the previous free ions (molal;
 *                                        held within [1e-300, 1e300])
 *   2N            temperature (C)        (25 if the array is NULL)
 *   [2N+1, 2N+1+K) mineral volume fraction
 *   rest          0
 * each normalised as (x - in_shift) * in_scale.  The sorbed totals are not
 * an input: the previous free ions fix them.  Outputs are the normalised
 * change of ln molality over the step: the guess is
 * m_j = pri_molal_j * exp(y_j / out_scale_j + out_shift_j), the change
 * held within +-30.
 */
#ifndef PFX_SURROGATE_H
#define PFX_SURROGATE_H

#include <stdint.h>

#include "pfx_chem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfx_mlp pfx_mlp;   /* opaque, owns device weights and buffers */

/* Network with random (seeded) initial weights and activation buffers for
   up to batch_max training rows.  nlayers <= 8, every width <= 4096. */
int pfx_mlp_create(int32_t nlayers, const int32_t *dims, int32_t batch_max, uint64_t seed, pfx_mlp **out);
int pfx_mlp_destroy(pfx_mlp *mlp);

/* HOST <-> device copy of one parameter array: which = 0 weights of
   `layer`, 1 biases, 2 weight gradients, 3 bias gradients, 4 normalisation
   (in_shift[dims0], in_scale[dims0], out_shift[out], out_scale[out]
   concatenated, layer ignored).  to_device != 0 uploads. */
int pfx_mlp_params(pfx_mlp *mlp, int32_t which, int32_t layer, float *host, int32_t to_device);

/* Inference Y[B][out] = f(X[B][dims0]) with the BF16 product kernels. */
int pfx_mlp_forward(pfx_mlp *mlp, const float *X, int64_t B, float *Y, void *stream);

/* Gradients of the mean squared error mean((f(X) - T)^2) over B <=
   batch_max rows into the gradient arrays; *loss_out (HOST, may be NULL)
   receives the loss (the call then waits for the stream). */
int pfx_mlp_gradients(pfx_mlp *mlp, const float *X, const float *T, int32_t B, float *loss_out, void *stream);

/* One Adam step (bias-corrected) on the master weights from the current
   gradients. */
int pfx_mlp_adam(pfx_mlp *mlp, float lr, float beta1, float beta2, float eps, void *stream);

/* pfx_mlp_gradients followed by pfx_mlp_adam(lr, 0.9, 0.999, 1e-8). */
int pfx_mlp_train_step(pfx_mlp *mlp, const float *X, const float *T, int32_t B, float lr, float *loss_out,
                       void *stream);

/* Normalised features X[ncells][dims0] of a batch of cells (uses total,
   pri_molal, temperature, mnrl_volfrac). */
int pfx_surrogate_features(pfx_tables *tables, pfx_mlp *mlp, const pfx_cells *cells, float *X, void *stream);

/* Fused warm start: features, every layer and the guess in one launch,
   cells->guess[ncells*N] = pri_molal * exp(predicted change) in FP64.  Needs dims0 <= 256, hidden
   widths multiples of 64 up to 256, 1 <= outputs <= 16 and N <= outputs
   (else PFX_E_UNSUP; pfx_surrogate_features + pfx_mlp_forward +
   pfx_surrogate_guess_from cover other shapes).  Waits for nothing. */
int pfx_surrogate_guess(pfx_tables *tables, pfx_mlp *mlp, pfx_cells *cells, void *stream);

/* Guess from network outputs Y[ncells][out] and the previous free ions
   prev[ncells*N]: guess[c*N+j] = prev[c*N+j] *
   exp(Y[c][j] / out_scale_j + out_shift_j). */
int pfx_surrogate_guess_from(pfx_mlp *mlp, const float *Y, int64_t ncells, int32_t naqcomp, const double *prev,
                             double *guess, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PFX_SURROGATE_H */

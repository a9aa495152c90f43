/*
 * clode_mlp.h - C ABI of the feature-map surrogate (libclode_mlp.so).
 *
 * No counterpart in the reference.  A small MLP (2 normalised parameters ->
 * 16 Fourier features -> 4 x 256 SiLU -> 4 heads of (log-period, amplitude,
 * 16 class logits)) trained and evaluated with BF16 matrix-core kernels
 * accumulating in FP32.  It screens a parameter grid: cells where the heads
 * disagree or the predicted class changes between neighbours (and cells
 * within a dilation radius of those) are flagged for re-solving by the
 * integrator.  Callers: the Python clode_amd.surrogate module and any host
 * that wants the screening.
 *
 * Conventions: inputs X are [n][2] float in [0, 1]; regression targets T are
 * [n][2] float, already normalised ((log period - m0) / s0, (amplitude - m1)
 * / s1); classes are int in 0..15.  Every call returns CLODE_MLP_OK or an
 * error code and sets a thread-local message.
 */
#ifndef CLODE_MLP_H
#define CLODE_MLP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLODE_MLP_OK 0
#define CLODE_MLP_ERR_INVALID 1
#define CLODE_MLP_ERR_RUNTIME 2

typedef struct clode_mlp clode_mlp;

const char *clode_mlp_last_error(void);

/* network on device `device`, weights initialised from `seed` */
int clode_mlp_create(clode_mlp **out, int device, unsigned long long seed);
int clode_mlp_destroy(clode_mlp *h);

/* flat FP32 parameter vector: for each layer Wt[out][in] then bias[out];
 * layers 32->256, 256->256 x 3, 256->128 (head columns 72..127 are padding
 * and stay zero) */
int clode_mlp_n_params(clode_mlp *h, size_t *n);
int clode_mlp_get_params(clode_mlp *h, float *out, size_t n);
int clode_mlp_set_params(clode_mlp *h, const float *in, size_t n);
/* the fixed Fourier frequency matrix B[2][16]: feature f of x is
 * sin / cos(2 pi (B[0][f] x0 + B[1][f] x1)) */
int clode_mlp_get_fourier(clode_mlp *h, float *out, size_t n);
/* mean and scale used to map the regression heads back to feature units */
int clode_mlp_set_target_scaling(clode_mlp *h, const float *mean2, const float *std2);

/* raw head outputs [n][72] for n <= 262144 rows; encoded (optional) receives
 * the BF16 bits of the encoded inputs [n][32] */
int clode_mlp_forward(clode_mlp *h, const float *X, size_t n, float *out, unsigned short *encoded);
/* loss (mean over rows of the sum over heads of squared errors + class
 * cross-entropy, each head on its bootstrap subset) and its gradient with
 * respect to the flat parameter vector, for n <= 65536 rows */
int clode_mlp_loss_grad(clode_mlp *h, const float *X, const float *T, const int *cls, size_t n, float *loss,
                        float *grad);
/* Adam with cosine-decayed learning rate over `epochs` passes of shuffled
 * minibatches; epoch_loss (optional) receives the mean loss per epoch */
int clode_mlp_train(clode_mlp *h, const float *X, const float *T, const int *cls, size_t n, int epochs, int batch,
                    float lr, unsigned long long seed, float *epoch_loss);
/* Screen the nx x ny grid of normalised inputs (ix / (nx - 1), iy / (ny - 1)),
 * cell index ix * ny + iy: ensemble period, amplitude and class per cell, and
 * flags (bit 0 heads disagree by more than tol, bit 1 class differs from a
 * 4-neighbour, bit 2 within `dilate` cells of a flagged cell) */
int clode_mlp_screen_grid(clode_mlp *h, int nx, int ny, float tol, int dilate, float *period, float *amp, int *cls,
                          unsigned char *flags);
/* device time of the last train / screen call, milliseconds */
int clode_mlp_last_ms(clode_mlp *h, double *ms);

/* No counterpart in the reference (nor in the allocating call above): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same screening
 * as clode_mlp_screen_grid, with identical results, but the four outputs are
 * caller-owned DEVICE buffers of nx * ny elements on the handle's device
 * (d_period, d_amp: float; d_cls: int; d_flags: bytes), and the staging
 * buffer of the dilation lives in a static device array of the library
 * (CLODE_MLP_NOALLOC_MAX_CELLS bytes per device, held until the process
 * ends).  Grids above CLODE_MLP_NOALLOC_MAX_CELLS cells are rejected with
 * CLODE_MLP_ERR_INVALID.  The call allocates nothing and returns when the
 * results are complete.  All calls on one device share the static array:
 * calls that overlap in time on one device (several host threads, or several
 * handles on one device used concurrently) race and must be serialised by
 * the caller. */
#define CLODE_MLP_NOALLOC_MAX_CELLS 4194304 /* 2048 x 2048 */
int clode_mlp_screen_grid_noalloc(clode_mlp *h, int nx, int ny, float tol, int dilate, float *d_period, float *d_amp,
                                  int *d_cls, unsigned char *d_flags);

#ifdef __cplusplus
}
#endif

#endif /* CLODE_MLP_H */

/* Warm-start network of the steady solver: C-ABI of libfvs2d_hip.so.
 *
 * No reference counterpart: the reference (shirzadgit/fvs2d) starts every
 * steady solve from the free stream (src/initialize.f90:36-43).  This network
 * maps a cell record to an estimate of the converged conservative state, from
 * which the FP64 solver iterates to the same residual tolerance.
 *
 * Cell record (6 floats): x, y (cell centre, chord from x = 0 to 1), wall
 * distance d, Mach number, angle of attack (degrees), thickness ratio.
 * Feature map (68 entries), with r, th the polar coordinates about the
 * mid-chord point (0.5, 0) and s = ln(1 + 4 r) / ln(201):
 *   [0..31]   sin(k pi s), cos(k pi s)      k = 1..16 (pairs)
 *   [32..63]  sin(k th),   cos(k th)        k = 1..16 (pairs)
 *   [64]      ln(1 + 100 d) / ln(5001)
 *   [65..67]  (M - 0.675) / 0.175, (aoa - 1) / 3, (t - 0.125) / 0.025
 * Network: 68 -> depth hidden layers of width 256 with GELU (erf form) -> 4
 * outputs.  Weights and activations of the hidden layers are bfloat16 on the
 * matrix cores with float32 accumulation; master weights, Adam moments, the
 * output layer and the loss are float32.  Loss: mean over rows and outputs of
 * (y - target)^2.
 *
 * Flat parameter layout (float32): for each layer, the weights W[in][out] row
 * by row (input index slow), then the biases b[out].  depth = 5 gives
 * 281,860 parameters.
 *
 * Return codes are those of include/fvs2d.h; fvs_last_error() has the text.
 */
#ifndef FVS2D_MLP_H
#define FVS2D_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FVS_MLP_RAW 6
#define FVS_MLP_FEATURES 68
#define FVS_MLP_WIDTH 256
#define FVS_MLP_OUT 4

typedef struct fvs_mlp fvs_mlp;

typedef struct fvs_mlp_options {
  int depth;                     /* hidden layers, 1..8 (5) */
  int batch;                     /* rows per training step, multiple of 64 (65536) */
  double lr, beta1, beta2, eps;  /* Adam (1e-3, 0.9, 0.999, 1e-8) */
  unsigned long long seed;       /* weight initialisation and row sampling */
} fvs_mlp_options;

/* Glorot-uniform weights, zero biases; device work buffers for opt->batch rows. */
int fvs_mlp_create(fvs_mlp **m, const fvs_mlp_options *opt);
int fvs_mlp_destroy(fvs_mlp *m);
/* number of parameters (flat layout above) */
long long fvs_mlp_nparams(const fvs_mlp *m);
int fvs_mlp_get_params(fvs_mlp *m, float *params);
/* replaces the master weights and clears the Adam moments and step count */
int fvs_mlp_set_params(fvs_mlp *m, const float *params);
/* the feature map in float32 (n x 68), as the network sees it before the
 * bfloat16 conversion */
int fvs_mlp_features(fvs_mlp *m, int n, const float *raw, float *feat);
/* training set, kept on the device: raw[n][6], target[n][4] */
int fvs_mlp_set_data(fvs_mlp *m, long long n, const float *raw, const float *target);
/* steps Adam steps on batches drawn with replacement from the training set
 * (row = hash(seed, step, slot) mod n); loss[steps] (may be NULL) receives
 * each step's batch loss */
int fvs_mlp_train(fvs_mlp *m, int steps, double *loss);
/* loss and gradient (flat layout) over the n <= batch given rows, no update */
int fvs_mlp_grad(fvs_mlp *m, int n, const float *raw, const float *target, double *loss, float *grad);
/* one Adam step on the n <= batch given rows */
int fvs_mlp_step(fvs_mlp *m, int n, const float *raw, const float *target, double *loss);
/* network output out[n][4] for n rows (processed batch rows at a time) */
int fvs_mlp_predict(fvs_mlp *m, long long n, const float *raw, float *out);
/* Times steps training steps with events around every launch.  Per kernel
 * class [features, forward GEMM, output layer + loss, weight-gradient GEMM,
 * backward-data GEMM, Adam]: mean ms per launch, matrix-core flops per launch
 * and launch count. */
int fvs_mlp_kernel_times(fvs_mlp *m, int steps, double *ms6, double *flop6, int *launches6);
/* device time of the last fvs_mlp_train call (ms) */
double fvs_mlp_last_elapsed_ms(const fvs_mlp *m);

#ifdef __cplusplus
}
#endif

#endif

/*
 * Learned surrogate of the FDTD labeller (extension; the reference has no
 * learned part).  A 3D convolutional encoder over the scene's voxel grid
 * and an MLP conditioned on a per-query coordinate vector predict one
 * output vector per query (the real and imaginary parts of a probe's
 * transfer function at each frequency).  Training and inference run on
 * the GPU's matrix cores in BF16 with FP32 accumulation; master weights,
 * gradients and the Adam state are FP32.
 *
 * Network (S input voxels per axis, C0 input channels):
 *   4 x [3x3x3 convolution, stride 2, padding 1, bias, GELU] with
 *   channels[0..3] outputs: S^3 x C0 -> (S/16)^3 x channels[3] = F features;
 *   per query q of scene b:
 *   h1 = GELU(feat_b W1f + coord_q W1c + b1), h2 = GELU(h1 W2 + b2),
 *   y  = h2 W3 + b3 (outputs values);
 *   loss = mean over all outputs of (y - label)^2.
 * GELU is 0.5 z (1 + erf(z / sqrt 2)).
 *
 * Voxel input: the scenes are invariant along y, so a scene is given as
 * an (x, z) map [S][S][C0] (x slowest, channel fastest) that the GPU
 * repeats along y.  coords: [n][queries][coords]; labels and outputs:
 * [n][queries][outputs].
 *
 * Parameters, flat FP32 in this order: for each convolution layer l its
 * weights [27 C_in][C_out] (kernel offset (a, b, c) in {0,1,2}^3 along
 * the voxel axes (x, y, z) and input channel i give row (9a + 3b + c) C_in
 * + i; input voxel 2o + a - 1 feeds output voxel o) then its bias [C_out];
 * then W1f [F][width], W1c [coords][width], b1 [width], W2 [width][width],
 * b2 [width], W3 [width][outputs], b3 [outputs].  The feature index of
 * the encoder output is ((x s + y) s + z) channels[3] + c, s = S/16.
 */
#ifndef RFSYN_FDTD_SURROGATE_H
#define RFSYN_FDTD_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int grid;          /* S: voxels per axis, a multiple of 16 */
    int in_channels;   /* C0 */
    int channels[4];   /* convolution outputs */
    int coords;        /* coordinates per query */
    int width;         /* MLP width */
    int outputs;       /* outputs per query */
    int queries;       /* queries per scene */
    int max_batch;     /* scenes per call */
} fdtd_sur_config_t;

typedef struct fdtd_sur fdtd_sur_t;

/* Weights He-uniform from PCG32(seed) (engine/core/rng.c), biases 0.
 * NULL on invalid config or allocation failure. */
fdtd_sur_t *fdtd_sur_create(const fdtd_sur_config_t *cfg, uint64_t seed);
void fdtd_sur_destroy(fdtd_sur_t *m);

size_t fdtd_sur_n_params(const fdtd_sur_t *m);
int fdtd_sur_get_params(fdtd_sur_t *m, float *out);
int fdtd_sur_set_params(fdtd_sur_t *m, const float *in);   /* also resets the Adam state */

/* Forward pass of n <= max_batch scenes into out. */
int fdtd_sur_predict(fdtd_sur_t *m, const float *maps, const float *coords, int n, float *out);

/* Loss and the gradient of every parameter (flat order above) for n
 * scenes, without changing the weights.  grad may be NULL. */
int fdtd_sur_gradient(fdtd_sur_t *m, const float *maps, const float *coords, const float *labels, int n,
                      float *grad, double *loss);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on n scenes; *loss
 * (may be NULL) receives the loss before the step. */
int fdtd_sur_train_step(fdtd_sur_t *m, const float *maps, const float *coords, const float *labels, int n,
                        float lr, double *loss);

#ifdef __cplusplus
}
#endif

#endif

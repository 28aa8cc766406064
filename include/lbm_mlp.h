/*
 * lbm_mlp.h -- C ABI of the learned velocity surrogate in liblbm_hip.so: an
 * ensemble of small MLPs (5 -> 128 -> 128 -> 128 -> 6, tanh hidden layers,
 * 34,566 parameters each) trained and evaluated by hand-written HIP kernels.
 *
 * No counterpart in the reference: it stands in for the fluid solve of
 * LBM::step / advanceSquirmer (src/utils/lbm.cpp:128-263) by predicting the
 * squirmer's translational and angular velocity from its configuration; the
 * full lattice-Boltzmann step stays the fallback and the measure of its error.
 *
 * Parameters of one net, float, in this order: W1[128][5], b1[128],
 * W2[128][128], b2[128], W3[128][128], b3[128], W4[6][128], b4[6]
 * (row-major [out][in]).  Inputs and targets are the normalised values the
 * caller prepares.  Weights are kept in FP32 (master copy, Adam state) and,
 * in the BF16 mode, used as BF16 by matrix-core (MFMA) tiles that accumulate
 * in FP32; the FP32 mode multiplies in FP32 on the vector units.
 *
 * Every call returns 0 or a negative LBM_HIP_E* code (lbm_hip.h);
 * lbm_hip_last_error() holds the message.
 */
#ifndef LBM_MLP_H
#define LBM_MLP_H

#include "lbm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_MLP_IN 5
#define LBM_MLP_HIDDEN 128
#define LBM_MLP_OUT 6
#define LBM_MLP_PARAMS 34566

#define LBM_MLP_FP32 0 /* FP32 weights and products */
#define LBM_MLP_BF16 1 /* BF16 weights and activations into MFMA tiles, FP32 accumulation */

typedef struct lbm_mlp lbm_mlp;

/* `nets` independently initialised nets (seeded from `seed`), on `device`. */
int lbm_mlp_create(int nets, int mode, unsigned long long seed, int device, lbm_mlp** out);
void lbm_mlp_destroy(lbm_mlp* m);

/* Parameters of net k (LBM_MLP_PARAMS floats, layout above). */
int lbm_mlp_get_params(lbm_mlp* m, int k, float* params);
int lbm_mlp_set_params(lbm_mlp* m, int k, const float* params);

/* Outputs of every net for n inputs x[n][5] -> y[nets][n][6]. */
int lbm_mlp_forward(lbm_mlp* m, const float* x, int n, float* y);

/* Mean over the nets -> mean[n][6], and per input the spread: the largest
 * standard deviation over the nets among the outputs, divided by the
 * largest |mean| among them (relative disagreement of the ensemble). */
int lbm_mlp_predict(lbm_mlp* m, const float* x, int n, float* mean, float* spread);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): lbm_mlp_predict for at most
 * LBM_MLP_NOALLOC_MAX_SAMPLES inputs per call, with the same kernels and
 * results bit for bit, keeping its inputs, the nets' outputs, mean and
 * spread in one static device array of LBM_MLP_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends) instead of buffers it allocates.  It calls
 * no allocation or free; n above the maximum is rejected (LBM_HIP_EINVAL).
 * Every call on one GPU uses that GPU's array: calls that overlap on one GPU
 * (other handles, other threads or streams) race and must be serialised by
 * the caller. */
#define LBM_MLP_NOALLOC_MAX_SAMPLES 4096
#define LBM_MLP_NOALLOC_SCRATCH_BYTES 6488064 /* 4 * 4096 * (5 + 64 * 6 + 6 + 1) */
int lbm_mlp_predict_noalloc(lbm_mlp* m, const float* x, int n, float* mean, float* spread);

/* Loss L = mean over n samples and 6 outputs of (y_net - t)^2 for net k,
 * and its gradient dL/dparams (LBM_MLP_PARAMS floats); no update. */
int lbm_mlp_loss_grad(lbm_mlp* m, int k, const float* x, const float* t, int n, float* loss, float* grad);

/* Train every net on (x, t) for `epochs` passes of minibatches of `batch`
 * samples (each net draws its own sample order from the seed) with Adam
 * (lr, beta1 0.9, beta2 0.999, eps 1e-8); loss_out[k] = net k's loss over
 * all samples after training (NULL to skip). */
int lbm_mlp_train(lbm_mlp* m, const float* x, const float* t, int n, int epochs, int batch, float lr,
                  float* loss_out);

#ifdef __cplusplus
}
#endif
#endif /* LBM_MLP_H */

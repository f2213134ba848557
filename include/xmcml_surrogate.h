/*
 * xmcml_surrogate.h -- C-ABI of the learned surrogate of the photon walk.
 *
 * None of these entry points has a counterpart in the reference: xmcml has
 * no learned model.  They serve callers that ask many layered-tissue
 * questions (ring reflectance, time-of-flight histogram) and can accept a
 * model answer, falling back to xmc_batch_run() (xmcml.h) where the model is
 * out of its range or uncertain.
 *
 * The model is a multilayer perceptron: `inputs` features -> `depth` hidden
 * layers of `hidden` GELU units (tanh form) -> a linear head of 2*`outputs`
 * values, the first `outputs` being predicted means and the rest predicted
 * log-variances.  Operands are bfloat16 (round to nearest even), products
 * accumulate in fp32 on the matrix cores, master weights and the Adam moments
 * are fp32.  Training minimises the masked Gaussian negative log-likelihood
 *     L = scale * sum_{b,j} mask[b,j] * 0.5 * (exp(-s) * (y - mu)^2 + s),
 *     s = clamp(log-variance, -10, 10)   (no gradient where clamped).
 *
 * Parameter order of xmc_mlp_get_params/set_params/get_grads, fp32: for each
 * layer l = 0..depth (the head last), W_l as [out_l][in_l] row-major, then
 * b_l[out_l]; in_0 = inputs, out_depth = 2*outputs.
 *
 * Array arguments named *_dev are device pointers (fp32, row-major, batch
 * rows); all others are host pointers.  `stream` is a hipStream_t or NULL.
 * Calls on one xmc_mlp must not overlap: they share its activation buffers.
 */
#ifndef XMCML_AMD_XMCML_SURROGATE_H
#define XMCML_AMD_XMCML_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xmc_mlp xmc_mlp;

typedef struct xmc_mlp_desc {
    int inputs;    /* features per sample, 1..64 */
    int hidden;    /* hidden width, a multiple of 64 */
    int depth;     /* hidden layers, 1..8 */
    int outputs;   /* predicted quantities, a multiple of 32 (head width 2*outputs) */
    int maxBatch;  /* largest batch of one training step */
    int device;
} xmc_mlp_desc;

/* Allocates the model on `device` and draws initial weights from `seed`
 * (uniform in +-sqrt(6/(fan_in+fan_out)), zero biases). */
int xmc_mlp_create(const xmc_mlp_desc* desc, uint64_t seed, xmc_mlp** out);
int xmc_mlp_destroy(xmc_mlp* m);

/* Number of fp32 parameters in the order described above. */
int xmc_mlp_param_count(const xmc_mlp* m, int64_t* count);
int xmc_mlp_get_params(xmc_mlp* m, float* params, int64_t count);
/* Sets weights and resets the Adam state and step counter. */
int xmc_mlp_set_params(xmc_mlp* m, const float* params, int64_t count);

/* out_dev[b][0..2*outputs) = model(x_dev[b][0..inputs)) for b < batch; any
 * batch size.  Models up to 512 wide (hidden and head) run in one launch that
 * keeps each 128-row tile's activations on chip; wider ones, or an out_dev
 * that is not 16-byte aligned, take xmc_mlp_forward_layers.  Both give the
 * same bits. */
int xmc_mlp_forward(xmc_mlp* m, const float* x_dev, int batch, float* out_dev, void* stream);

/* The same forward as one matrix-core GEMM launch per layer, activations
 * through device memory, in chunks of maxBatch rows. */
int xmc_mlp_forward_layers(xmc_mlp* m, const float* x_dev, int batch, float* out_dev, void* stream);

/* Forward, loss and backward of one batch (batch <= maxBatch) into the
 * gradient buffers (overwritten, not accumulated).  y_dev and mask_dev are
 * [batch][outputs].  If loss != NULL the call synchronises and stores L. */
int xmc_mlp_backward(xmc_mlp* m, const float* x_dev, const float* y_dev, const float* mask_dev, int batch,
                     float scale, float* loss, void* stream);

/* Gradients of the last xmc_mlp_backward, in parameter order. */
int xmc_mlp_get_grads(xmc_mlp* m, float* grads, int64_t count);

/* One AdamW update from the current gradients:
 *   m = b1*m + (1-b1)*g;  v = b2*v + (1-b2)*g^2;
 *   w -= lr * (m/(1-b1^t)) / (sqrt(v/(1-b2^t)) + eps) + lr*weightDecay*w   (weights only). */
int xmc_mlp_adam(xmc_mlp* m, float lr, float beta1, float beta2, float eps, float weightDecay, void* stream);

/* One training step on `batch` rows drawn uniformly (with replacement) from
 * a resident data set X_dev [rows][inputs], Y_dev and M_dev [rows][outputs]:
 * row i of the batch is hash(seed, step, i) mod rows (SplitMix64), then
 * backward and AdamW as above.  If loss != NULL the call synchronises and
 * stores the batch loss before the update. */
int xmc_mlp_train_sampled(xmc_mlp* m, const float* X_dev, const float* Y_dev, const float* M_dev, int64_t rows,
                          int batch, uint64_t seed, int64_t step, float scale, float lr, float beta1, float beta2,
                          float eps, float weightDecay, float* loss, void* stream);

/* Feature rows of layered-tissue configurations, on the device.
 * layers_dev is [n][slots][5] layer tuples in xmcml's order (refractive
 * index, mua, mus, anisotropy g, thickness), count_dev [n] the layers used.
 * Row b of x_dev ([n][6 * slots]) holds, for each used slot, the values
 * (mua, mus (1 - g), g, n, thickness) mapped from [lo_k, hi_k] to [-1, 1]
 * (on log scale where logScale[k]), 0 for unused slots, then one 0/1
 * occupied flag per slot.  inRange_dev[b] = 1 when minLayers <= count <=
 * slots and every used value lies in [lo_k (1 - tol), hi_k (1 + tol)].
 * Double arithmetic, rounded to float at the end. */
typedef struct xmc_layer_box {
    int slots;
    int minLayers;
    double lo[5], hi[5];
    int logScale[5];
    double tol;
} xmc_layer_box;

int xmc_surrogate_encode(const double* layers_dev, const int32_t* count_dev, int64_t n, const xmc_layer_box* box,
                         float* x_dev, uint8_t* inRange_dev, int device, void* stream);

/* The model's answers for n feature rows on the device (any n; the forward
 * runs in chunks of 65,536 rows on a buffer the model keeps):
 *   values_dev[b][j] = exp(mu_j * scale[j] + shift[j]), j < outputs,
 *   uncertain_dev[b] = 1 when mean_{j < flagColumns} exp(s_j / 2) * scale[j] > sigmaMax,
 * mu and s the head's predicted means and clamped log-variances (host arrays
 * scale/shift of `outputs` floats: the targets' standardisation). */
int xmc_mlp_answer(xmc_mlp* m, const float* x_dev, int64_t n, const float* scale, const float* shift, int flagColumns,
                   float sigmaMax, float* values_dev, uint8_t* uncertain_dev, void* stream);

/* Device memory for the arrays above, on the model's kind of device (plain
 * hipMalloc/hipMemcpy; kind 1 = host to device, 2 = device to host,
 * 3 = device to device).  xmc_dev_sync waits for all work on `device`. */
int xmc_dev_alloc(int device, uint64_t bytes, void** ptr);
int xmc_dev_free(void* ptr);
int xmc_dev_copy(void* dst, const void* src, uint64_t bytes, int kind);
int xmc_dev_sync(int device);

/* Convenience: xmc_mlp_backward (no loss readback) then xmc_mlp_adam. */
int xmc_mlp_train_step(xmc_mlp* m, const float* x_dev, const float* y_dev, const float* mask_dev, int batch,
                       float scale, float lr, float beta1, float beta2, float eps, float weightDecay, void* stream);

/* C[M][N] = sum_k A[m][k] * B[n][k] with bf16 operands given as fp32 host
 * arrays (rounded to bf16 on upload) and fp32 accumulation on the matrix
 * cores; M, N and K multiples of 64.  tile = 128 forces the 128 x 128 tile
 * kernel (M, N multiples of 128), 0 the automatic choice.  A check of the
 * matrix kernels alone. */
int xmc_mlp_gemm_check(const float* a, const float* b, int M, int N, int K, float* c, int tile, int device);

#ifdef __cplusplus
}
#endif

#endif

/* C ABI of the warm-start network: a U-Net that maps a target density, box-downsampled to
 * S x S, to the dual weights of an S x S site grid, trained and run on the GPU with the
 * library's own BF16 matrix-core kernels (FP32 accumulation, FP32 master weights, Adam).
 *
 * No counterpart in the reference: the reference solver always starts L-BFGS from the caller's
 * weights (zeros at t = 0, ot_op.cuh:66-70 direct_ot's `weights` argument).  This network gives
 * that argument a learned value; the OT solve (caustic_ot.h) still decides convergence.
 *
 * Layout: images are float32 [B][S][S] (row-major, the site grid's row index first); parameters
 * are one float32 array, per layer (e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b out)
 * the weights [k*k*c_in][c_out] (row (ky*k + kx)*c_in + ci) followed by the bias [c_out].
 * All pointers are host pointers.  Functions return WSN_OK or an error code; wsn_last_error()
 * describes the last error of the calling thread. */
#ifndef CAUSTIC_WSN_H
#define CAUSTIC_WSN_H

#ifdef __cplusplus
extern "C" {
#endif

enum { WSN_OK = 0, WSN_EINVAL = 1, WSN_ENOMEM = 2, WSN_EHIP = 3 };

typedef struct wsn_net wsn_net;

const char* wsn_last_error(void);

/* S: image and site-grid side (a multiple of 64); max_batch: largest batch of any call. */
int wsn_create(int S, int max_batch, wsn_net** out);
int wsn_destroy(wsn_net* net);
long long wsn_param_count(const wsn_net* net);
int wsn_set_params(wsn_net* net, const float* params);  /* also resets the Adam state */
int wsn_get_params(wsn_net* net, float* params);

/* y[B][S][S] = network(x[B][S][S]) */
int wsn_forward(wsn_net* net, const float* x, int B, float* y);

/* The loss: with e = y - mean(y) - t per sample (the prediction's mean is a free gauge),
 * L = (sum e^2 + grad_weight * sum over neighbouring pixel pairs (e_a - e_b)^2) / (B*S*S).
 * grad_weight defaults to 0 (mean squared error). */
int wsn_set_loss(wsn_net* net, float grad_weight);

/* The loss and its gradient with respect to every parameter. */
int wsn_gradients(wsn_net* net, const float* x, const float* t, int B, float* grad, float* loss);

/* Keep `count` training pairs (x, t) on the device for wsn_train. */
int wsn_set_dataset(wsn_net* net, const float* x, const float* t, int count);

/* n_order / B Adam steps at learning rate lr over the resident pairs, in the given order;
 * losses (may be NULL) receives each step's loss. */
int wsn_train(wsn_net* net, const int* order, int n_order, int B, float lr, float* losses);

/* Time spent in wsn_train, samples processed, and the conv FLOPs of one training sample
 * (forward + backward). */
int wsn_train_stats(const wsn_net* net, double* seconds, long long* samples, double* flops);

#ifdef __cplusplus
}
#endif
#endif

/*
 * mbg_mlp.h - C ABI of libmbg_mlp.so: the sweep surrogate, a multilayer perceptron trained
 * and evaluated on an MI355X (gfx950) with hand-written BF16 MFMA kernels (FP32 accumulation).
 *
 * No reference counterpart: MaBoSSG has no learned component.  The surrogate maps a model
 * variant's parameters (log rate multipliers and initial probabilities) to the windowed
 * readout marginals the simulator produces; the simulator generates its training data and
 * judges its error.
 *
 * Network: n_in -> hidden -> ... -> hidden -> n_out, GELU (tanh form) after every hidden layer,
 * sigmoid on the output, mean squared error loss, Adam.  Inputs are padded to a multiple of
 * 64 columns with zeros; weights are stored [out][in].
 */
#ifndef MBG_MLP_H
#define MBG_MLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int mbg_status; /* same codes as mbg.h: 0 ok, 3 invalid argument, 4 HIP error */

typedef struct mbg_mlp mbg_mlp;

/* hidden_layers hidden layers of width hidden (a multiple of 64); weights drawn from the
 * uniform Glorot range with the given seed. */
mbg_status mbg_mlp_create(int device, int n_in, int hidden, int hidden_layers, int n_out, int max_batch, uint64_t seed,
                          mbg_mlp** out);
void mbg_mlp_destroy(mbg_mlp* m);
const char* mbg_mlp_last_error(const mbg_mlp* m);

/* Parameters as FP32, layer by layer: W [out][in_padded] then b [out]. */
size_t mbg_mlp_num_params(const mbg_mlp* m);
int mbg_mlp_padded_inputs(const mbg_mlp* m);
mbg_status mbg_mlp_get_params(mbg_mlp* m, float* out);
mbg_status mbg_mlp_set_params(mbg_mlp* m, const float* in);

/* Loss and gradients (same layout as the parameters) of one batch, without updating. */
mbg_status mbg_mlp_gradients(mbg_mlp* m, const float* x, const float* y, int batch, float* loss, float* grads);

/* Training set held on the device: x [n][n_in], y [n][n_out]. */
mbg_status mbg_mlp_load_dataset(mbg_mlp* m, const float* x, const float* y, size_t n);
/* epochs over the device training set in shuffled mini-batches of `batch` rows (a multiple of 64);
 * Adam with learning rate lr (beta1 0.9, beta2 0.999, eps 1e-8); mean loss of the last epoch. */
mbg_status mbg_mlp_train(mbg_mlp* m, int epochs, int batch, float lr, uint64_t shuffle_seed, float* last_loss);
/* Predictions y [n][n_out] for x [n][n_in] (host buffers); device time of the pass in ms. */
mbg_status mbg_mlp_predict(mbg_mlp* m, const float* x, size_t n, float* y, double* ms);
/* Ensemble prediction over `count` networks of equal input and output widths on one device:
 * mean [n][n_out] = the members' average, spread [n] (may be NULL) = per row, the largest
 * population standard deviation over members of any output.  compute_ms = device time from the
 * inputs resident in device memory to the reduced results, before the copy back. */
mbg_status mbg_mlp_ensemble_predict(mbg_mlp* const* members, int count, const float* x, size_t n, float* mean,
                                    float* spread, double* compute_ms);

/* The same ensemble prediction without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  x [n][n_in], mean [n][n_out] and spread [n] (may be NULL) are
 * DEVICE buffers of the members' device.  The members' outputs are staged in a fixed 64 MiB
 * device array of the library (16,777,216 floats) in chunks of rows, so members x n_out x 64
 * must not exceed 16,777,216 (MBG_ERR_ARG otherwise); n itself is unbounded.  Calls overlapping
 * on one device share that array and race: run them one at a time per device.  Synchronous on
 * the first member's stream; the results equal mbg_mlp_ensemble_predict's. */
mbg_status mbg_mlp_ensemble_predict_noalloc(mbg_mlp* const* members, int count, const float* x, size_t n, float* mean,
                                            float* spread);

#ifdef __cplusplus
}
#endif

#endif /* MBG_MLP_H */

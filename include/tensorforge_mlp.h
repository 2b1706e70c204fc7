/* Cost-model network of the autotuner (in libtensorforge_rt.so): a fully connected FP32 network
 * in -> hidden -> hidden -> hidden -> 1 with ReLU hidden layers, trained on measured kernel
 * times. New: the reference has no learned cost model (its fused-GEMM cost option,
 * tensorforge/interop.py:20-23, selects an analytic count).
 *
 * params: W1[hidden][in] b1[hidden] W2[hidden][hidden] b2 W3[hidden][hidden] b3 W4[1][hidden] b4,
 * tensorforge_mlp_param_count(in, hidden) floats. All pointers are device memory; calls are
 * asynchronous on `stream` and return the HIP status of the launches. */
#ifndef TENSORFORGE_MLP_H
#define TENSORFORGE_MLP_H
#ifdef __cplusplus
extern "C" {
#endif

long tensorforge_mlp_param_count(int in, int hidden);

/* out[b] for X[B][in]; work receives the three hidden activations (3 * B * hidden floats). */
int tensorforge_mlp_forward(const float* params, const float* X, float* work, float* out, int B,
                            int in, int hidden, void* stream);

/* Gradient of mean((out - target)^2) with respect to params, after a forward pass over the same
 * X into work/out; dwork holds 2 * B * hidden floats. */
int tensorforge_mlp_backward(const float* params, const float* X, const float* work,
                             const float* out, const float* target, float* grads, float* dwork,
                             int B, int in, int hidden, void* stream);

/* out[b] for X[B][in] like tensorforge_mlp_forward, without a workspace and without allocating:
 * the hidden activations live in a device array of the library (one per GPU, held until the
 * process ends). No counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, a tuning loop inside a serving process). B * hidden must not exceed
 * tensorforge_mlp_predict_max_activations() (262,144); larger batches return
 * hipErrorInvalidValue without launching. Calls that overlap on one GPU share the array and
 * race: serialise them. */
long tensorforge_mlp_predict_max_activations(void);
int tensorforge_mlp_predict(const float* params, const float* X, float* out, int B, int in,
                            int hidden, void* stream);

/* One Adam step (bias-corrected, step >= 1) over n parameters with moments m, v. */
int tensorforge_mlp_adam(float* params, const float* grads, float* m, float* v, long n, float lr,
                         float beta1, float beta2, float eps, int step, void* stream);

#ifdef __cplusplus
}
#endif
#endif

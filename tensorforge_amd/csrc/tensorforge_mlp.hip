// Cost-model network of the autotuner: a fully connected FP32 network
//   in -> H -> H -> H -> 1   (ReLU after each hidden layer)
// with forward, backward (mean squared error) and Adam kernels. Parameters are one flat array:
//   W1[H][in] b1[H] W2[H][H] b2[H] W3[H][H] b3[H] W4[1][H] b4[1]
// Activations of the hidden layers are kept in `work` (3 * B * H floats) for the backward pass.
// No counterpart in the reference (its fused-GEMM cost option is an analytic count,
// tensorforge/interop.py:20-23).
#include <hip/hip_runtime.h>
#include <cstddef>
#include <mutex>

extern "C" void tensorforge_record_error(int code, const char* where);

namespace {

constexpr int kThreads = 256;

// Y[b][o] = act(bias[o] + sum_i W[o][i] X[b][i]); one thread per (b, o), the weight row from L1
template <bool Relu>
__global__ void __launch_bounds__(kThreads)
dense_forward(const float* __restrict__ W, const float* __restrict__ bias,
              const float* __restrict__ X, float* __restrict__ Y, int B, int in, int out) {
  const long total = static_cast<long>(B) * out;
  for (long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long>(gridDim.x) * blockDim.x) {
    const int b = static_cast<int>(t / out);
    const int o = static_cast<int>(t % out);
    const float* w = W + static_cast<long>(o) * in;
    const float* x = X + static_cast<long>(b) * in;
    float acc = bias[o];
    for (int i = 0; i < in; ++i) acc = fmaf(w[i], x[i], acc);
    Y[t] = Relu ? fmaxf(acc, 0.0f) : acc;
  }
}

// dOut[b] = 2 (out[b] - target[b]) / B  (gradient of the mean squared error)
__global__ void __launch_bounds__(kThreads)
loss_grad(const float* __restrict__ out, const float* __restrict__ target,
          float* __restrict__ dout, int B) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x)
    dout[b] = 2.0f * (out[b] - target[b]) / static_cast<float>(B);
}

// dX[b][i] = (sum_o W[o][i] dZ[b][o]) * (Aprev[b][i] > 0) : gradient into the previous ReLU layer
__global__ void __launch_bounds__(kThreads)
dense_backward_input(const float* __restrict__ W, const float* __restrict__ dZ,
                     const float* __restrict__ Aprev, float* __restrict__ dX, int B, int in,
                     int out) {
  const long total = static_cast<long>(B) * in;
  for (long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; t < total;
       t += static_cast<long>(gridDim.x) * blockDim.x) {
    const int b = static_cast<int>(t / in);
    const int i = static_cast<int>(t % in);
    const float* dz = dZ + static_cast<long>(b) * out;
    float acc = 0.0f;
    for (int o = 0; o < out; ++o) acc = fmaf(W[static_cast<long>(o) * in + i], dz[o], acc);
    dX[t] = Aprev[t] > 0.0f ? acc : 0.0f;
  }
}

// dW[o][i] = sum_b dZ[b][o] X[b][i] (i < in); dbias[o] = sum_b dZ[b][o] (column i == in). One
// workgroup per 8 x 8 tile of (o, i); its 4 groups of 64 threads each sum every fourth batch row
// (b = q, q + 4, ...), and the four partial sums are added in a fixed order through LDS, so the
// result does not depend on scheduling.
__global__ void __launch_bounds__(kThreads)
dense_backward_weights(const float* __restrict__ dZ, const float* __restrict__ X,
                       float* __restrict__ dW, float* __restrict__ dbias, int B, int in, int out) {
  __shared__ float part[4][64];
  const int tiles_i = (in + 1 + 7) / 8;
  const int to = static_cast<int>(blockIdx.x) / tiles_i;
  const int ti = static_cast<int>(blockIdx.x) % tiles_i;
  const int q = static_cast<int>(threadIdx.x) / 64;
  const int l = static_cast<int>(threadIdx.x) % 64;
  const int o = to * 8 + l / 8;
  const int i = ti * 8 + l % 8;
  const bool valid = o < out && i <= in;
  float acc = 0.0f;
  if (valid) {
    for (int b = q; b < B; b += 4) {
      const float x = i < in ? X[static_cast<long>(b) * in + i] : 1.0f;
      acc = fmaf(dZ[static_cast<long>(b) * out + o], x, acc);
    }
  }
  part[q][l] = acc;
  __syncthreads();
  if (q == 0 && valid) {
    const float sum = ((part[0][l] + part[1][l]) + part[2][l]) + part[3][l];
    if (i < in)
      dW[static_cast<long>(o) * in + i] = sum;
    else
      dbias[o] = sum;
  }
}

unsigned weight_tiles(int in, int out) {
  return static_cast<unsigned>(((out + 7) / 8) * ((in + 1 + 7) / 8));
}

// Adam (Kingma & Ba 2015) with bias correction, step t >= 1
__global__ void __launch_bounds__(kThreads)
adam_step(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
          float* __restrict__ v, long n, float lr, float beta1, float beta2, float eps,
          float c1, float c2) {
  for (long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x; t < n;
       t += static_cast<long>(gridDim.x) * blockDim.x) {
    const float gt = g[t];
    const float mt = beta1 * m[t] + (1.0f - beta1) * gt;
    const float vt = beta2 * v[t] + (1.0f - beta2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  }
}

unsigned grid_for(long n) {
  long blocks = (n + kThreads - 1) / kThreads;
  if (blocks > 4096) blocks = 4096;
  return static_cast<unsigned>(blocks > 0 ? blocks : 1);
}

struct Layout {
  long w1, b1, w2, b2, w3, b3, w4, b4, total;
};

Layout layout(int in, int H) {
  Layout l;
  l.w1 = 0;
  l.b1 = l.w1 + static_cast<long>(H) * in;
  l.w2 = l.b1 + H;
  l.b2 = l.w2 + static_cast<long>(H) * H;
  l.w3 = l.b2 + H;
  l.b3 = l.w3 + static_cast<long>(H) * H;
  l.w4 = l.b3 + H;
  l.b4 = l.w4 + H;
  l.total = l.b4 + 1;
  return l;
}

int finish(const char* where) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), where);
  return static_cast<int>(err);
}

}  // namespace

extern "C" long tensorforge_mlp_param_count(int in, int hidden) {
  return layout(in, hidden).total;
}

extern "C" int tensorforge_mlp_forward(const float* params, const float* X, float* work,
                                       float* out, int B, int in, int hidden, void* stream) {
  if (B <= 0 || in <= 0 || hidden <= 0) return static_cast<int>(hipErrorInvalidValue);
  const Layout l = layout(in, hidden);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long bh = static_cast<long>(B) * hidden;
  float* a1 = work;
  float* a2 = work + bh;
  float* a3 = work + 2 * bh;
  hipLaunchKernelGGL(dense_forward<true>, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w1, params + l.b1, X, a1, B, in, hidden);
  hipLaunchKernelGGL(dense_forward<true>, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w2, params + l.b2, a1, a2, B, hidden, hidden);
  hipLaunchKernelGGL(dense_forward<true>, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w3, params + l.b3, a2, a3, B, hidden, hidden);
  hipLaunchKernelGGL(dense_forward<false>, dim3(grid_for(B)), dim3(kThreads), 0, s,
                     params + l.w4, params + l.b4, a3, out, B, hidden, 1);
  return finish("tensorforge_mlp_forward");
}

// Gradients of the mean squared error of `out` (from the forward pass over the same X and work)
// against `target`, into `grads` (same layout as params). `dwork` holds 2 * B * hidden floats.
extern "C" int tensorforge_mlp_backward(const float* params, const float* X, const float* work,
                                        const float* out, const float* target, float* grads,
                                        float* dwork, int B, int in, int hidden, void* stream) {
  if (B <= 0 || in <= 0 || hidden <= 0) return static_cast<int>(hipErrorInvalidValue);
  const Layout l = layout(in, hidden);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long bh = static_cast<long>(B) * hidden;
  const float* a1 = work;
  const float* a2 = work + bh;
  const float* a3 = work + 2 * bh;
  float* dA = dwork;            // B x hidden
  float* dB = dwork + bh;       // B x hidden
  float* dout = dwork;          // B (reused before dA is written)
  // the output gradient lives in dB's first B entries so dA can be written next
  dout = dB;
  hipLaunchKernelGGL(loss_grad, dim3(grid_for(B)), dim3(kThreads), 0, s, out, target, dout, B);
  // layer 4: dW4 = dout^T a3, db4; da3 = W4^T dout masked by a3 > 0
  hipLaunchKernelGGL(dense_backward_weights, dim3(weight_tiles(hidden, 1)), dim3(kThreads), 0, s,
                     dout, a3, grads + l.w4, grads + l.b4, B, hidden, 1);
  hipLaunchKernelGGL(dense_backward_input, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w4, dout, a3, dA, B, hidden, 1);
  // layer 3
  hipLaunchKernelGGL(dense_backward_weights, dim3(weight_tiles(hidden, hidden)), dim3(kThreads), 0,
                     s, dA, a2, grads + l.w3, grads + l.b3, B, hidden, hidden);
  hipLaunchKernelGGL(dense_backward_input, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w3, dA, a2, dB, B, hidden, hidden);
  // layer 2
  hipLaunchKernelGGL(dense_backward_weights, dim3(weight_tiles(hidden, hidden)), dim3(kThreads), 0,
                     s, dB, a1, grads + l.w2, grads + l.b2, B, hidden, hidden);
  hipLaunchKernelGGL(dense_backward_input, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w2, dB, a1, dA, B, hidden, hidden);
  // layer 1 (no gradient into the inputs)
  hipLaunchKernelGGL(dense_backward_weights, dim3(weight_tiles(in, hidden)), dim3(kThreads), 0, s,
                     dA, X, grads + l.w1, grads + l.b1, B, in, hidden);
  return finish("tensorforge_mlp_backward");
}

extern "C" int tensorforge_mlp_adam(float* params, const float* grads, float* m, float* v, long n,
                                    float lr, float beta1, float beta2, float eps, int step,
                                    void* stream) {
  if (n <= 0 || step < 1) return static_cast<int>(hipErrorInvalidValue);
  float c1 = 1.0f, c2 = 1.0f, p1 = 1.0f, p2 = 1.0f;
  for (int i = 0; i < step; ++i) {
    p1 *= beta1;
    p2 *= beta2;
  }
  c1 = 1.0f - p1;
  c2 = 1.0f - p2;
  hipLaunchKernelGGL(adam_step, dim3(grid_for(n)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), params, grads, m, v, n, lr, beta1, beta2,
                     eps, c1, c2);
  return finish("tensorforge_mlp_adam");
}

// ---------------------------------------------------------------------------------------------
// Inference without allocation: the hidden activations live in a file-scope device array instead
// of a caller workspace, for callers that cannot allocate (code under HIP graph capture, a tuning
// loop inside a serving process). Each GPU the library runs on holds its own copy of the array
// until the process ends. Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two threads) race and must be serialised by the caller. Nothing needs
// clearing between calls: each layer overwrites all B * hidden activations it later reads.
namespace {

constexpr long kPredictMaxActivations = 1L << 18;   // B * hidden <= 262,144 (3 MiB of floats)
constexpr int kMaxDevices = 64;

__device__ __align__(256) float scratch[3 * kPredictMaxActivations];

std::once_flag g_scratch_once[kMaxDevices];
float* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_status[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch`, looked up once per device with that device
// current (the array's name in host code is only a host-side stand-in, never a device address).
hipError_t scratch_on_current_device(float** ptr) {
  int dev = 0;
  hipError_t err = hipGetDevice(&dev);
  if (err != hipSuccess) return err;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::call_once(g_scratch_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  });
  if (g_scratch_status[dev] != hipSuccess) return g_scratch_status[dev];
  *ptr = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" long tensorforge_mlp_predict_max_activations(void) { return kPredictMaxActivations; }

extern "C" int tensorforge_mlp_predict(const float* params, const float* X, float* out, int B,
                                       int in, int hidden, void* stream) {
  if (B <= 0 || in <= 0 || hidden <= 0) return static_cast<int>(hipErrorInvalidValue);
  const long bh = static_cast<long>(B) * hidden;
  if (bh > kPredictMaxActivations) {
    tensorforge_record_error(static_cast<int>(hipErrorInvalidValue),
                             "tensorforge_mlp_predict: B * hidden above "
                             "tensorforge_mlp_predict_max_activations()");
    return static_cast<int>(hipErrorInvalidValue);
  }
  float* work = nullptr;
  const hipError_t err = scratch_on_current_device(&work);
  if (err != hipSuccess) {
    tensorforge_record_error(static_cast<int>(err), "tensorforge_mlp_predict: scratch address");
    return static_cast<int>(err);
  }
  const Layout l = layout(in, hidden);
  hipStream_t s = static_cast<hipStream_t>(stream);
  float* a1 = work;
  float* a2 = work + bh;
  float* a3 = work + 2 * bh;
  hipLaunchKernelGGL(dense_forward<true>, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w1, params + l.b1, X, a1, B, in, hidden);
  hipLaunchKernelGGL(dense_forward<true>, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w2, params + l.b2, a1, a2, B, hidden, hidden);
  hipLaunchKernelGGL(dense_forward<true>, dim3(grid_for(bh)), dim3(kThreads), 0, s,
                     params + l.w3, params + l.b3, a2, a3, B, hidden, hidden);
  hipLaunchKernelGGL(dense_forward<false>, dim3(grid_for(B)), dim3(kThreads), 0, s,
                     params + l.w4, params + l.b4, a3, out, B, hidden, 1);
  return finish("tensorforge_mlp_predict");
}

// Runtime support library: error status and synchronisation for generated launchers, plus the
// HBM copy probe used to measure the achievable bandwidth roof.
//
// Replaces the reference's native shim (tensorforge/include/tensorforge_aux.h:6-11,
// tensorforge_aux.cpp:9-36): instead of `throw;` on a launch error, launchers record the error
// here and callers read it back with tensorforge_last_error().
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstddef>
#include <cstdio>
#include <mutex>
#include <cstring>

namespace {
std::atomic<int> g_last_error{0};
std::mutex g_where_mutex;
char g_where[256] = {0};
}  // namespace

extern "C" {

void tensorforge_record_error(int code, const char* where) {
  if (code == 0) return;
  g_last_error.store(code);
  std::lock_guard<std::mutex> lock(g_where_mutex);
  std::snprintf(g_where, sizeof(g_where), "%s", where ? where : "");
}

int tensorforge_last_error(void) { return g_last_error.load(); }

const char* tensorforge_last_error_where(void) { return g_where; }

void tensorforge_clear_error(void) {
  g_last_error.store(0);
  std::lock_guard<std::mutex> lock(g_where_mutex);
  g_where[0] = 0;
}

const char* tensorforge_error_string(int code) {
  return hipGetErrorString(static_cast<hipError_t>(code));
}

int tensorforge_synch_device(void* stream) {
  hipError_t err = stream ? hipStreamSynchronize(static_cast<hipStream_t>(stream))
                          : hipDeviceSynchronize();
  if (err == hipSuccess) err = hipGetLastError();
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), "tensorforge_synch_device");
  return static_cast<int>(err);
}

}  // extern "C"

// 16-byte vector copy: each lane moves one 16-B chunk per iteration of a grid-stride loop.
typedef double tf_double2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) tensorforge_copy_kernel(const tf_double2* __restrict__ src,
                                                               tf_double2* __restrict__ dst,
                                                               size_t n2) {
  size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  const size_t step = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n2; i += step) dst[i] = __builtin_nontemporal_load(&src[i]);
}

extern "C" int tensorforge_copy(double* dst, const double* src, size_t n, void* stream) {
  // restates the bandwidth probe of benchmarks/bandwidth/global.cpp:30-39 (To[i] = From[i]);
  // n must be even (16-byte chunks)
  if (n == 0) return 0;
  if (n % 2 != 0) return static_cast<int>(hipErrorInvalidValue);
  const size_t n2 = n / 2;
  size_t blocks = (n2 + 255) / 256;
  if (blocks > 256 * 64) blocks = 256 * 64;
  hipLaunchKernelGGL(tensorforge_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const tf_double2*>(src),
                     reinterpret_cast<tf_double2*>(dst), n2);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), "tensorforge_copy");
  return static_cast<int>(err);
}

// Time-step sequences as HIP graphs: the caller brackets a sequence of launcher calls on one
// stream with begin/end, then replays the instantiated graph. Launchers that query the device
// on first use (occupancy of persistent kernels) must have been called once before capture.
extern "C" int tensorforge_graph_begin(void* stream) {
  if (stream == nullptr) return static_cast<int>(hipErrorInvalidValue);
  hipError_t err = hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                         hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), "tensorforge_graph_begin");
  return static_cast<int>(err);
}

extern "C" int tensorforge_graph_end(void* stream, void** graph_exec) {
  if (stream == nullptr || graph_exec == nullptr) return static_cast<int>(hipErrorInvalidValue);
  *graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  hipError_t err = hipStreamEndCapture(static_cast<hipStream_t>(stream), &graph);
  if (err == hipSuccess) {
    hipGraphExec_t exec = nullptr;
    err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (err == hipSuccess) *graph_exec = exec;
  }
  if (graph != nullptr) (void)hipGraphDestroy(graph);
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), "tensorforge_graph_end");
  return static_cast<int>(err);
}

extern "C" int tensorforge_graph_launch(void* graph_exec, void* stream) {
  if (graph_exec == nullptr) return static_cast<int>(hipErrorInvalidValue);
  hipError_t err = hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                                  static_cast<hipStream_t>(stream));
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), "tensorforge_graph_launch");
  return static_cast<int>(err);
}

extern "C" int tensorforge_graph_destroy(void* graph_exec) {
  if (graph_exec == nullptr) return 0;
  return static_cast<int>(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
}

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
#include <string>

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
  // HIP's own last error too: a stale one would be read back by the next launcher
  (void)hipGetLastError();
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

// The reference's C++ shim (tensorforge/include/tensorforge_aux.h:6-11,
// tensorforge_aux.cpp:9-36), exported with its signatures so objects compiled against that header
// link here. checkErr reports the error a launcher recorded or HIP's last error (and clears both,
// so one error is reported once) instead of terminating with `throw;`; the caller reads the code
// with tensorforge_last_error() before calling it, or checks the return of the C functions.
namespace tensorforge {

namespace {
std::string g_prev_file;
int g_prev_line = 0;
std::mutex g_check_mutex;
}  // namespace

void checkErr(const std::string& file, int line) {
  int code = tensorforge_last_error();
  std::string where = tensorforge_last_error_where();
  const hipError_t hip = hipGetLastError();
  if (code == 0 && hip != hipSuccess) {
    code = static_cast<int>(hip);
    where = "HIP";
  }
  std::lock_guard<std::mutex> lock(g_check_mutex);
  if (code != 0) {
    std::fprintf(stderr, "\n%s, line %d: %s (%d) in %s\n", file.c_str(), line,
                 hipGetErrorString(static_cast<hipError_t>(code)), code, where.c_str());
    if (g_prev_line > 0)
      std::fprintf(stderr, "Previous check:\n%s, line %d\n", g_prev_file.c_str(), g_prev_line);
    g_last_error.store(0);
    std::lock_guard<std::mutex> wl(g_where_mutex);
    g_where[0] = 0;
    return;
  }
  g_prev_file = file;
  g_prev_line = line;
}

// Whole-device synchronisation, as the reference's synchDevice (the stream argument is unused
// there too); then the same check.
void synchDevice(void* stream) {
  (void)stream;
  const hipError_t err = hipDeviceSynchronize();
  if (err != hipSuccess) tensorforge_record_error(static_cast<int>(err), "tensorforge::synchDevice");
  checkErr(__FILE__, __LINE__);
}

}  // namespace tensorforge

// HBM copy probe: each workgroup of 256 lanes copies TF_COPY_CHUNKS x 256 consecutive 16-byte
// chunks; every lane has TF_COPY_CHUNKS independent 16-byte loads in flight before its first
// store, each wavefront access covers 1 KB contiguous, and loads and stores are non-temporal
// (the data is touched once). One pass, no grid-stride loop: the grid has one workgroup per
// TF_COPY_CHUNKS x 256 chunks.
typedef double tf_double2 __attribute__((ext_vector_type(2)));
#define TF_COPY_CHUNKS 4

__global__ void __launch_bounds__(256) tensorforge_copy_kernel(const tf_double2* __restrict__ src,
                                                               tf_double2* __restrict__ dst,
                                                               size_t n2) {
  const size_t base = static_cast<size_t>(blockIdx.x) * (256 * TF_COPY_CHUNKS) + threadIdx.x;
  tf_double2 v[TF_COPY_CHUNKS];
  if (base + (TF_COPY_CHUNKS - 1) * 256 < n2) {
#pragma unroll
    for (int u = 0; u < TF_COPY_CHUNKS; ++u) v[u] = __builtin_nontemporal_load(&src[base + u * 256]);
#pragma unroll
    for (int u = 0; u < TF_COPY_CHUNKS; ++u) __builtin_nontemporal_store(v[u], &dst[base + u * 256]);
  } else {
    for (int u = 0; u < TF_COPY_CHUNKS; ++u) {
      const size_t i = base + u * 256;
      if (i < n2) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
    }
  }
}

extern "C" int tensorforge_copy(double* dst, const double* src, size_t n, void* stream) {
  // restates the bandwidth probe of benchmarks/bandwidth/global.cpp:30-39 (To[i] = From[i]);
  // n must be even (16-byte chunks)
  if (n == 0) return 0;
  if (n % 2 != 0) return static_cast<int>(hipErrorInvalidValue);
  const size_t n2 = n / 2;
  const size_t per_block = 256 * TF_COPY_CHUNKS;
  const size_t blocks = (n2 + per_block - 1) / per_block;
  if (blocks > 0x7fffffffu) return static_cast<int>(hipErrorInvalidValue);
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

/* C ABI of the gfx950 TensorForge runtime library (libtensorforge_rt.so).
 *
 * Generated kernels are exported from their own shared objects as
 *   void launcher_kernel_<md5>(<operands...>, size_t numElements,
 *                              unsigned* flags, void* streamPtr);
 * with the parameter list of the reference's generated launchers
 * (tensorforge/generators/generator.py:384-436); the prototypes of the prebuilt ones are in
 * tensorforge_catalog.h. The functions below are the runtime those launchers report to.
 */
#ifndef TENSORFORGE_AMD_H
#define TENSORFORGE_AMD_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Record a HIP error code raised in a launcher. Replaces the reference's CHECK_ERR path
 * (tensorforge/include/tensorforge_aux.cpp:11-29), which terminates with `throw;`. */
void tensorforge_record_error(int code, const char* where);

/* Last recorded HIP error code (0 when none). New: the reference has no status query. */
int tensorforge_last_error(void);

/* Name of the entry point that recorded the last error. */
const char* tensorforge_last_error_where(void);

/* Reset the recorded error. */
void tensorforge_clear_error(void);

/* hipGetErrorString of a code. */
const char* tensorforge_error_string(int code);

/* Wait for `stream` (a hipStream_t; NULL: the whole device) and return the HIP status.
 * Replaces tensorforge::synchDevice (tensorforge/include/tensorforge_aux.cpp:32-36). */
int tensorforge_synch_device(void* stream);

/* Copy n doubles (n even) device-to-device on `stream`: the HBM bandwidth probe, restating
 * benchmarks/bandwidth/global.cpp:30-39 and benchmarks/bandwidth/hip_kernel.cpp:4-13. */
int tensorforge_copy(double* dst, const double* src, size_t n, void* stream);

/* HIP-graph capture of launcher sequences (time-step loops). New: the reference has no graph
 * support. begin/end bracket launcher calls issued on `stream` (a non-null hipStream_t; call each
 * launcher once before capturing it, so its one-time device queries are done); end returns an
 * instantiated graph in *graph_exec; launch replays it on `stream`; destroy frees it. Each returns
 * the HIP status. */
int tensorforge_graph_begin(void* stream);
int tensorforge_graph_end(void* stream, void** graph_exec);
int tensorforge_graph_launch(void* graph_exec, void* stream);
int tensorforge_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif
#endif

/* The reference's C++ runtime shim (tensorforge/include/tensorforge_aux.h:6-11), with its
 * declarations unchanged: tensorforge::checkErr and tensorforge::synchDevice are exported
 * out of line by libtensorforge_rt.so, so code generated against the reference's header links
 * against this library as it is.
 *
 * Differences in behaviour (tensorforge/include/tensorforge_aux.cpp:9-36): checkErr reports
 * an error recorded by a generated launcher as well as HIP's last error, clears both after
 * reporting them, and returns instead of terminating with `throw;`; the status functions of
 * tensorforge_amd.h (tensorforge_last_error, tensorforge_synch_device) give the code to callers
 * that act on it. synchDevice synchronises the whole device, as in the reference. */
#ifndef TENSORFORGE_AUX_H
#define TENSORFORGE_AUX_H
#include <string>
#include "tensorforge_amd.h"

#define CHECK_ERR tensorforge::checkErr(__FILE__, __LINE__)

namespace tensorforge {

void checkErr(const std::string &file, int line);
void synchDevice(void *stream = nullptr);

}  // namespace tensorforge

#endif

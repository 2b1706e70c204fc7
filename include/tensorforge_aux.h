/* C++ compatibility header for code generated against the reference's runtime shim
 * (tensorforge/include/tensorforge_aux.h:6-11): CHECK_ERR and tensorforge::synchDevice keep
 * their names and now report through the status functions of tensorforge_amd.h. */
#ifndef TENSORFORGE_AUX_H
#define TENSORFORGE_AUX_H
#include <cstdio>
#include <string>
#include "tensorforge_amd.h"

namespace tensorforge {
inline int checkErr(const std::string& file, int line) {
  const int code = tensorforge_last_error();
  if (code != 0) {
    std::fprintf(stderr, "%s:%d: tensorforge error %d (%s) in %s\n", file.c_str(), line, code,
                 tensorforge_error_string(code), tensorforge_last_error_where());
  }
  return code;
}
inline int synchDevice(void* stream = nullptr) { return tensorforge_synch_device(stream); }
}  // namespace tensorforge

#define CHECK_ERR tensorforge::checkErr(__FILE__, __LINE__)
#endif

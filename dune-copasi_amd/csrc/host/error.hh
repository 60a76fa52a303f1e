// Error types crossing the C ABI as negative status codes plus a thread-local
// message, the way the reference reports failures through format_exception
// (dune/copasi/common/exceptions.hh:21-33).
#pragma once
#include <stdexcept>
#include <string>

#include <hip/hip_runtime.h>

namespace dcopasi {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct IOError : Error {
  using Error::Error;
};
struct NotImplemented : Error {
  using Error::Error;
};
struct DeviceError : Error {
  using Error::Error;
};

#define DCG_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::dcopasi::DeviceError(std::string(#call) + " failed: " + hipGetErrorString(e_) +      \
                                   " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");        \
  } while (0)

}  // namespace dcopasi

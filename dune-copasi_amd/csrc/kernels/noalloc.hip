// The library's fixed device work space for entry points that must not
// allocate (dcg_linear_solve_noalloc): one static array per GPU, held until
// the process ends.  Host code never passes the array's name to a kernel or
// copy; it asks for the array's device address once per device and uses that.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels.hh"

namespace dcopasi::k {

namespace {
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range for the work space");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (rc != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress of the work space failed: ") + hipGetErrorString(rc));
    g_ptr[device] = static_cast<double*>(p);
  }
  return g_ptr[device];
}

}  // namespace dcopasi::k

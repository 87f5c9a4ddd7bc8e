// fl_noalloc.hip -- the device scratch of fl_generate_f64_noalloc.
//
// A file-scope device array of FL_NOALLOC_SCRATCH_BYTES holds every device
// buffer of an allocation-free generate() call (the terrain's arrays and the
// visit order's scratch).  Each GPU that runs the library holds its own copy
// until the process ends.  In host code the array's name only stands in for
// it: its device address is taken once per device, with that device
// current, when a context is created on it (fl_noalloc_scratch_init), and
// only that address is passed to kernels and copies.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "fl_internal.h"

static __device__ __align__(256) unsigned char scratch[FL_NOALLOC_SCRATCH_BYTES];

namespace {

struct ScratchRegistry {
  std::mutex mu;
  std::vector<uint8_t*> ptr;  // per device: the address of `scratch`, null until taken
  std::vector<std::mutex*> call_mu;  // per device: serialises the calls that use it
};

ScratchRegistry& registry() {
  static ScratchRegistry* r = new ScratchRegistry();  // never destroyed: used until the process ends
  return *r;
}

}  // namespace

int fl_noalloc_scratch_init(int device) {
  ScratchRegistry& r = registry();
  std::lock_guard<std::mutex> g(r.mu);
  if ((int)r.ptr.size() <= device) {
    r.ptr.resize(device + 1, nullptr);
    r.call_mu.resize(device + 1, nullptr);
  }
  if (r.ptr[device]) return FL_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the calling thread's current device
  if (e != hipSuccess || !p) {
    std::string msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    fl_internal_set_error(msg.c_str());
    return FL_ERR_HIP;
  }
  r.ptr[device] = (uint8_t*)p;
  r.call_mu[device] = new std::mutex();
  return FL_OK;
}

int fl_noalloc_scratch(int device, uint8_t** ptr, uint64_t* bytes, std::mutex** call_mu) {
  ScratchRegistry& r = registry();
  std::lock_guard<std::mutex> g(r.mu);
  if ((int)r.ptr.size() <= device || !r.ptr[device]) {
    fl_internal_set_error("the device's static scratch has not been set up (no context was created on it)");
    return FL_ERR_INVALID;
  }
  *ptr = r.ptr[device];
  *bytes = FL_NOALLOC_SCRATCH_BYTES;
  *call_mu = r.call_mu[device];
  return FL_OK;
}

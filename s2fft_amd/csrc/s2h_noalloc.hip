// s2h_forward_noalloc: the FP64 forward transform with its workspace in a
// file-scope device array instead of a caller-provided buffer, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The array lives once per device for the life of the process.  Host code
// never passes its name to a kernel (on the host that name is only a
// stand-in): the device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and that address is what the
// forward's kernels receive.  Every call on one device shares the array and
// leaves the last call's contents in it; the forward overwrites every
// workspace element it reads, so nothing needs clearing between calls.  Calls
// that overlap on one device (two streams, two threads) race on the array.
#include <mutex>

#include "s2h_plan.h"

namespace {

__device__ __align__(256) unsigned char scratch[S2H_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int scratch_of_current_device(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return S2H_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return S2H_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return S2H_ERR_HIP;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return S2H_OK;
}

}  // namespace

extern "C" size_t s2h_noalloc_scratch_bytes(void) { return S2H_NOALLOC_SCRATCH_BYTES; }

extern "C" int s2h_forward_noalloc(const s2h_plan* p, const void* f, int f_is_complex, void* flm,
                                   int batch, s2h_stream stream) {
  if (!p || !f || !flm || batch < 0) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  if (s2h_workspace_bytes(p, batch) > (size_t)S2H_NOALLOC_SCRATCH_BYTES) return S2H_ERR_SCRATCH;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return S2H_ERR_HIP;
  if (dev != p->device) return S2H_ERR_INVALID;
  void* work = nullptr;
  const int st = scratch_of_current_device(&work);
  if (st) return st;
  return s2h_forward_ex(p, f, f_is_complex, flm, work, batch, S2H_PREC_FP64, 0, stream);
}

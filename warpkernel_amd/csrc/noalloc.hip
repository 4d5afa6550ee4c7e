// noalloc.hip -- wk_engine_set_colinds_noalloc: the device form of
// wk_engine_set_colinds without any device allocation, for callers that cannot
// allocate (tight serving loops that swap an engine's columns, harnesses that
// forbid hipMalloc).  The column checks write their flags into a file-scope
// device array instead of a buffer allocated per call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "wk_internal.hpp"

// The check flags: one per check block.  The check kernels run at most
// kColFlags blocks (grid-stride loops) whatever the number of slots, so no
// input needs more.  Every device the library runs on holds its own copy until
// the process ends, and every call on one device shares it: calls that overlap
// on one device race.  Host code never passes this name to a kernel (on the
// host it is a stand-in, not a device address); it uses the per-device address
// from hipGetSymbolAddress.
static __device__ __align__(256) int32_t scratch[wk::kColFlags];

namespace wk {
namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
int32_t *g_scratch[kMaxDevices];  // per device: the address of `scratch` there

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int32_t *scratch_on_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) fail(WK_ERR_LIMIT, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = static_cast<int32_t *>(p);
  }
  return g_scratch[dev];
}

}  // namespace
}  // namespace wk

using namespace wk;

extern "C" int wk_engine_set_colinds_noalloc(wk_engine *E, const int32_t *pcols_dev, void *stream) {
  return guarded([&] {
    if (!E) fail(WK_ERR_ARG, "engine is NULL");
    if (!pcols_dev && E->fwd.allocate_nz) fail(WK_ERR_ARG, "pcols_dev is NULL");
    // the engine's device is current for the lookup and the launches
    int cur = 0;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != E->device) hip_check(hipSetDevice(E->device), "hipSetDevice");
    struct Restore {
      int dev, was;
      ~Restore() {
        if (dev != was) (void)hipSetDevice(was);
      }
    } restore{E->device, cur};
    swap_colinds(E, pcols_dev, 1, static_cast<hipStream_t>(stream), scratch_on_current_device());
  });
}

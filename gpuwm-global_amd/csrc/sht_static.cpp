// The point sampler for callers that cannot allocate device memory: code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  arwen_sht_sample_static runs arwen_sht_sample's kernels with
// the workspace in a library-owned device array instead of a caller buffer.
//
// `scratch` is one array per GPU the library runs on, held until the process
// ends.  Its address is resolved once per device, with that device current,
// when the device's first plan is created (sht_static_setup, called from
// arwen_sht_plan_create); host code only ever passes that resolved address
// to the kernels.  Every call on one GPU shares the array and leaves the
// last call's contents: nothing in it needs to start at zero (the pack
// kernel writes every packed coefficient and the sampler every per-part sum
// before they are read), and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "arwen_sht.h"
#include "sht_static.h"

static __device__ __align__(256) unsigned char scratch[ARWEN_SHT_SAMPLE_STATIC_BYTES];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch;  // per device: the resolved address of `scratch`
}  // namespace

hipError_t sht_static_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0) return hipErrorInvalidDevice;
  if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
  if (g_scratch[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = ptr;
  return hipSuccess;
}

void* sht_static_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < (int)g_scratch.size() ? g_scratch[device] : nullptr;
}

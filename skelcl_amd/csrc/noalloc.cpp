// Reduce without allocation: skl_reduce_run_noalloc runs the Reduce kernel
// of skl_reduce_run with its scratch (completion counters and one partial
// per work-group) in a static device array instead of memory the skeleton
// allocates, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// The array is `kScratchBytes` per GPU and lives until the process ends.
// Every call on a GPU shares it: each call clears the counters with a fill
// kernel on its own stream before the Reduce kernel runs (no memset node,
// so the pair can be captured), and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include "common.h"
#include "skeleton.h"
#include "skelcl_skeletons.h"

using namespace skl;

namespace {
constexpr size_t kScratchBytes = 64 * 1024;
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device

// one work-group: the counters are a few KB
__global__ __launch_bounds__(256) void noalloc_clear(unsigned long long* p, unsigned words)
{
  for (unsigned i = threadIdx.x; i < words; i += 256) p[i] = 0ull;
}
}  // namespace

// Host code must not use this name as a device address: it stands for the
// symbol only; skl::noalloc_setup looks the address up per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int skl::noalloc_setup(int device)
{
  if (device < 0 || device >= kMaxDevices)
    return fail(SKL_ERR_INVALID, "device %d outside the no-allocation table", device);
  if (g_scratch[device]) return SKL_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SKL_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
  g_scratch[device] = p;
  return SKL_OK;
}

extern "C" int skl_reduce_run_noalloc(skl_skeleton s, void* out, const void* in, size_t n,
                                      const skl_arg* args, int nargs, skl_stream stream)
{
  SKL_TRY(ensure_initialized());
  const int dev = device_id();
  void* base = dev >= 0 && dev < kMaxDevices ? g_scratch[dev] : nullptr;
  if (!base) return fail(SKL_ERR_NOT_INIT, "no-allocation scratch of device %d is not set up", dev);
  if (n == 0) return fail(SKL_ERR_INVALID, "Reduce of an empty input");
  size_t counters = 0, need = 0;
  SKL_TRY(reduce_scratch_need(s, n, &counters, &need));
  if (need > kScratchBytes)
    return fail(SKL_ERR_INVALID,
                "Reduce needs %zu bytes of scratch for this launch shape; the no-allocation "
                "scratch holds %zu",
                need, kScratchBytes);
  hipStream_t st = resolve_stream(stream);
  hipLaunchKernelGGL(noalloc_clear, dim3(1), dim3(256), 0, st,
                     static_cast<unsigned long long*>(base), (unsigned)(counters / 8));
  SKL_HIP_TRY(hipGetLastError());
  return reduce_run_fixed_scratch(s, out, in, n, args, nargs, stream, base, kScratchBytes);
}

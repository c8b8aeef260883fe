// Work tickets for persistent work-groups.
//
// Loop shape used by every persistent kernel in this directory:
//
//   if (threadIdx.x == 0) slot = ticket_draw(counter);
//   __syncthreads();
//   for (;;) {
//     const long long t = ticket_read(slot);   // wave-uniform (scalar) value
//     __syncthreads();                          // everyone has read slot
//     if (t >= total) break;
//     ... item t ...
//     if (threadIdx.x == 0) { tail work of item t; slot = ticket_draw(counter); }
//     __syncthreads();
//   }
//
// Exactly one thread-0 region per iteration, closed by a barrier.  A
// thread-0 region at the end of the body next to another one at the top of
// the loop (the draw) lets the optimiser thread the two branches across the
// back-edge; the structuriser then turns that into a divergent loop exit in
// which the other 63 lanes of thread 0's wave keep executing the loop's
// barriers with a stale ticket while thread 0 waits to reconverge, and the
// work-group never leaves the loop.  Reading the ticket with readfirstlane
// makes the exit test a scalar branch.
#pragma once

#include <hip/hip_runtime.h>

__device__ __forceinline__ long long ticket_draw(unsigned long long* counter)
{
  return (long long)__hip_atomic_fetch_add(counter, 1ull, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ long long ticket_read(const long long& slot)
{
  const unsigned long long v = (unsigned long long)slot;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Device helpers shared by the pair kernel and the adjoint kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "nbtypes.h"

namespace mi355nb {

// r^2 <= c2 exactly; dx, dy, dz are exact (multiples of 2^-17 A below 128 A),
// so their FP64 squares and sum are exact.
__device__ __forceinline__ bool within(float dx, float dy, float dz, float r2, const CutTest& t) {
  if (r2 < t.lo) return true;
  if (r2 > t.hi) return false;
  const double x = dx, y = dy, z = dz;
  return x * x + y * y + z * z <= t.c2;
}

// Orders LDS accesses between the lanes of one wave: a compiler barrier plus
// wavefront-scope fences (a wave's LDS operations complete in order).
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace mi355nb

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

// Loads a record that every lane of the wave reads at the same address and
// moves it to scalar registers (v_readfirstlane per 32-bit word), so
// wave-uniform data does not occupy vector registers.
template <typename T>
__device__ __forceinline__ T waveUniformLoad(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  constexpr int W = sizeof(T) / 4;
  const int* src = reinterpret_cast<const int*>(p);
  T out;
  int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
  for (int k = 0; k < W; ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
  return out;
}

}  // namespace mi355nb

// PME exclusion correction in FP64 for gfx950.  The reciprocal sum contains
// every pair, the fully excluded ones included; NAMD's GPU path removes their
// smooth part in the bonded kernel (exclusionForce,
// src/ComputeBondedCUDAKernel.cu:633-772): for each excluded pair with
// r^2 < cutoff^2 the slow energy gains -qq erf(beta r)/r and the slow forces
// its gradient.  NAMD interpolates that function from a table in FP32; here
// it is evaluated in FP64 (erf and exp of the device library), with the pair
// displacement taken in FP64 from patch centre + patch-centred position.
//
// Work layout: the excluded pairs of a batch are stored by system in slot
// space; a block of at most EXCL_THREADS pairs of one system is one
// workgroup.  Forces go to the float4 slow-force output with float atomics;
// energy and virial (sum f_i (x) r_ij, component a*3+b = f_a r_b) are summed
// per workgroup and then per system in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>

#include "devutil.h"
#include "nb_kernels.h"
#include "nbtypes.h"

namespace mi355nb {

constexpr int EXCL_THREADS = 256;
constexpr int EXCL_VALUES = 10;  // energy, virial[9]

__device__ __forceinline__ double blockSum256(double v, double* sh) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < EXCL_THREADS / 64; ++k) s += sh[k];
  return s;
}

__global__ void __launch_bounds__(EXCL_THREADS)
exclusionCorrectionKernel(const int4* __restrict__ pairs, const int4* __restrict__ blocks,
                          const F4* __restrict__ xyzq, const int* __restrict__ slotPatch,
                          const double* __restrict__ centre, const double* __restrict__ sysBox, double cutoff2,
                          double beta, F4* __restrict__ forceSlow, int doEnergyVirial,
                          double* __restrict__ partials) {
  __shared__ double sh[EXCL_THREADS / 64];
  const int4 bk = blocks[blockIdx.x];   // {sys, begin, end, -}
  const int k = bk.y + threadIdx.x;
  double v[EXCL_VALUES];
#pragma unroll
  for (int u = 0; u < EXCL_VALUES; ++u) v[u] = 0.0;
  if (k < bk.z) {
    const int4 pr = pairs[k];           // {slot i, slot j, -, -}
    const F4 a = xyzq[pr.x], b = xyzq[pr.y];
    const int pa = slotPatch[pr.x], pb = slotPatch[pr.y];
    const double* L = sysBox + 3 * bk.x;
    double d[3] = {(centre[3 * pa] + (double)a.x) - (centre[3 * pb] + (double)b.x),
                   (centre[3 * pa + 1] + (double)a.y) - (centre[3 * pb + 1] + (double)b.y),
                   (centre[3 * pa + 2] + (double)a.z) - (centre[3 * pb + 2] + (double)b.z)};
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] -= L[c] * rint(d[c] / L[c]);
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 < cutoff2 && r2 > 0.0) {
      const double qq = (double)a.w * (double)b.w;
      const double r = sqrt(r2);
      const double rinv = 1.0 / r;
      const double e = erf(beta * r) * rinv;
      // E = -qq erf(beta r)/r;  f_i = -dE/dr * d/r
      const double dedr = -qq * (1.1283791670955126 * beta * exp(-beta * beta * r2) * rinv - e * rinv);
      const double fr = -dedr * rinv;
      const double f[3] = {fr * d[0], fr * d[1], fr * d[2]};
      atomicAdd(&forceSlow[pr.x].x, (float)f[0]);
      atomicAdd(&forceSlow[pr.x].y, (float)f[1]);
      atomicAdd(&forceSlow[pr.x].z, (float)f[2]);
      atomicAdd(&forceSlow[pr.y].x, (float)-f[0]);
      atomicAdd(&forceSlow[pr.y].y, (float)-f[1]);
      atomicAdd(&forceSlow[pr.y].z, (float)-f[2]);
      v[0] = -qq * e;
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) v[1 + 3 * p + q] = f[p] * d[q];
    }
  }
  if (doEnergyVirial) {
    for (int u = 0; u < EXCL_VALUES; ++u) {
      const double t = blockSum256(v[u], sh);
      if (threadIdx.x == 0) partials[(size_t)blockIdx.x * EXCL_VALUES + u] = t;
    }
  }
}

__global__ void __launch_bounds__(EXCL_THREADS)
exclusionFinalKernel(const int* __restrict__ sysBlockStart, const double* __restrict__ partials,
                     double* __restrict__ out) {
  __shared__ double sh[EXCL_THREADS / 64];
  const int sys = blockIdx.x;
  const int b0 = sysBlockStart[sys], b1 = sysBlockStart[sys + 1];
  for (int u = 0; u < EXCL_VALUES; ++u) {
    double v = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += EXCL_THREADS) v += partials[(size_t)b * EXCL_VALUES + u];
    const double t = blockSum256(v, sh);
    if (threadIdx.x == 0) out[(size_t)sys * EXCL_VALUES + u] = t;
  }
}

void buildExclusionBlocks(const std::vector<int>& pairSys, int numSystems, std::vector<int4>& blocks,
                          std::vector<int>& sysBlockStart) {
  blocks.clear();
  sysBlockStart.assign(numSystems + 1, 0);
  size_t k = 0;
  for (int s = 0; s < numSystems; ++s) {
    sysBlockStart[s] = (int)blocks.size();
    size_t e = k;
    while (e < pairSys.size() && pairSys[e] == s) ++e;
    for (size_t b = k; b < e; b += EXCL_THREADS)
      blocks.push_back(make_int4(s, (int)b, (int)std::min(b + EXCL_THREADS, e), 0));
    k = e;
  }
  sysBlockStart[numSystems] = (int)blocks.size();
}

hipError_t launchExclusionCorrection(int numBlocks, int numSystems, const int4* pairs, const int4* blocks,
                                     const int* sysBlockStart, const F4* xyzq, const int* slotPatch,
                                     const double* centre, const double* sysBox, double cutoff, double beta,
                                     F4* forceSlow, double* partials, double* out, hipStream_t s) {
  if (numBlocks > 0) {
    hipLaunchKernelGGL(exclusionCorrectionKernel, dim3(numBlocks), dim3(EXCL_THREADS), 0, s, pairs, blocks, xyzq,
                       slotPatch, centre, sysBox, cutoff * cutoff, beta, forceSlow, out ? 1 : 0, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (out) {
    hipLaunchKernelGGL(exclusionFinalKernel, dim3(numSystems), dim3(EXCL_THREADS), 0, s, sysBlockStart, partials,
                       out);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace mi355nb

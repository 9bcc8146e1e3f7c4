// Force-matching adjoint for gfx950: gradient of
//     L = sum_atoms |F_i - T_i|^2,   F = fast + slow nonbonded force,
// with respect to the LJ table entries A_ab, B_ab (ordered type bins) and the
// scaled charges, by recomputing every pair of the refined tile lists (no
// pair terms are stored).  No counterpart in the reference.
//
// For a pair with d = x_j - x_i and force coefficient f (F_i += f d,
// F_j -= f d) the loss depends on f through s = (g_i - g_j) . d with
// g = dL/dF = 2 (F - T), and
//     f = A cA(r) + B cB(r) + q_i q_j fe(r)
//     cA = r^-12 dS_r - 12 S r^-14,   cB = 6 S r^-8 - r^-6 dS_r
//     fe = beta^3 H(beta^2 r^2) - r^-3   (fast + slow; the 1/rc^3 terms cancel)
// where S is the LJ energy switch and dS_r = 2 dS/d(r^2) (the pair kernel's
// ds_r), and H is the rational factor of the slow force.  So
//     dL/dA_{t_i t_j} += s cA,  dL/dB_{t_i t_j} += s cB,
//     dL/dq_i += s q_j fe,      dL/dq_j += s q_i fe.
// Pair terms are FP32; sums are FP64: per-lane FP32 bins over one tile list
// flushed into per-workgroup FP64 bins (LDS), FP64 partials per workgroup
// summed in a fixed order; charge gradients go to FP64 atomics per slot.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nb_device.h"
#include "nb_kernels.h"
#include "slowforce.h"

namespace mi355nb {

constexpr int ADJ_WAVES = 4;
constexpr int ADJ_THREADS = 64 * ADJ_WAVES;
constexpr int ADJ_RED = 256;

// residual: g = 2 (F + Fs - T) per real slot, FP64 loss partials per block
__global__ void __launch_bounds__(ADJ_RED)
fmResidualKernel(int n, const F4* __restrict__ f, const F4* __restrict__ fs, const double* __restrict__ tgt,
                 F4* __restrict__ g, double* __restrict__ partials) {
  __shared__ double sh[ADJ_RED / 64];
  const int s = blockIdx.x * ADJ_RED + threadIdx.x;
  double l = 0.0;
  if (s < n) {
    F4 out{0.f, 0.f, 0.f, 0.f};
    if (tgt[4 * (size_t)s + 3] != 0.0) {
      const F4 a = f[s], b = fs[s];
      const double dx = ((double)a.x + (double)b.x) - tgt[4 * (size_t)s];
      const double dy = ((double)a.y + (double)b.y) - tgt[4 * (size_t)s + 1];
      const double dz = ((double)a.z + (double)b.z) - tgt[4 * (size_t)s + 2];
      l = dx * dx + dy * dy + dz * dz;
      out = F4{(float)(2.0 * dx), (float)(2.0 * dy), (float)(2.0 * dz), 0.f};
    }
    g[s] = out;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < ADJ_RED / 64; ++k) t += sh[k];
    partials[blockIdx.x] = t;
  }
}

// sums n doubles in a fixed order (one block)
__global__ void __launch_bounds__(ADJ_RED) sumDoublesKernel(int n, const double* __restrict__ in,
                                                              double* __restrict__ out) {
  __shared__ double sh[ADJ_RED / 64];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += ADJ_RED) v += in[k];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < ADJ_RED / 64; ++k) t += sh[k];
    *out = t;
  }
}

struct AdjArgs {
  const int* numLists;
  const TileList* tileLists;
  const int* tileJatomStart;
  const TileExcl* tileExcls;
  const F4* xyzq;
  const int* vdwTypes;
  int ntypes;
  NbConstants c;
  CutTest cut;
  const F4* g;            // per slot dL/dF (x, y, z)
  double* gq;             // per slot dL/d(scaled q), FP64 atomics
  double* binPartials;    // per workgroup: 2 * ntypes^2 (A bins, then B bins)
};

// Persistent workgroups of ADJ_WAVES waves; the lane / half-wave / rotation
// layout is that of the pair kernel, so every pair is visited once.
__global__ void __launch_bounds__(ADJ_THREADS) adjointKernel(const AdjArgs args) {
  extern __shared__ double s_dyn[];
  __shared__ F4 s_xyzq[ADJ_WAVES][TILE];
  __shared__ F4 s_g[ADJ_WAVES][TILE];
  __shared__ int s_type[ADJ_WAVES][TILE];
  __shared__ float s_gqj[ADJ_WAVES][2][TILE];

  const int nt = args.ntypes;
  const int nbins = 2 * nt * nt;
  double* s_wg = s_dyn;                                  // nbins FP64
  float* s_bin = reinterpret_cast<float*>(s_dyn + nbins);  // [wave][2 nt][64] FP32
  for (int k = threadIdx.x; k < nbins; k += ADJ_THREADS) s_wg[k] = 0.0;
  for (int k = threadIdx.x; k < ADJ_WAVES * 2 * nt * 64; k += ADJ_THREADS) s_bin[k] = 0.f;
  __syncthreads();

  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int i = lane & 31;
  const int h = lane >> 5;
  const float beta3 = args.c.ewald_2 * args.c.ewald_0;
  float* const myBins = s_bin + (size_t)wv * 2 * nt * 64 + lane;   // bin b at myBins[b * 64]
  const int numLists = *args.numLists;

  for (int list = blockIdx.x * ADJ_WAVES + wv; list < numLists; list += gridDim.x * ADJ_WAVES) {
    const TileList tl = args.tileLists[list];
    const int iatom = tl.iatomStart + i;
    F4 xi = args.xyzq[iatom];
    xi.x += tl.offx;
    xi.y += tl.offy;
    xi.z += tl.offz;
    const int ti = args.vdwTypes[iatom];
    const F4 gi = args.g[iatom];
    float gqi = 0.f;

    for (int jt = tl.jtileStart; jt <= tl.jtileEnd; ++jt) {
      const int ja = args.tileJatomStart[jt];
      const unsigned mask = args.tileExcls[jt].excl[i];
      if (h == 0) {
        s_xyzq[wv][i] = args.xyzq[ja + i];
        s_g[wv][i] = args.g[ja + i];
        s_type[wv][i] = args.vdwTypes[ja + i];
      }
      s_gqj[wv][h][i] = 0.f;
      waveSync();
      for (int t = 0; t < 16; ++t) {
        const int j = (i + t + 16 * h) & 31;
        const F4 xj = s_xyzq[wv][j];
        const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        if (((mask >> j) & 1u) && within(dx, dy, dz, r2, args.cut)) {
          const F4 gj = s_g[wv][j];
          const float sp = (gi.x - gj.x) * dx + (gi.y - gj.y) * dy + (gi.z - gj.z) * dz;
          const float rinv = __builtin_amdgcn_rsqf(r2);
          const float rinv2 = rinv * rinv;
          const float rinv3 = rinv * rinv2;
          const float rinv6 = rinv3 * rinv3;
          const bool sw = r2 > args.c.lj_5;
          const float delta_r = args.c.lj_4 - r2;
          const float S = sw ? delta_r * delta_r * (args.c.lj_0 + args.c.lj_1 * r2) : 1.0f;
          const float ds_r = sw ? delta_r * (args.c.lj_3 + args.c.lj_2 * r2) : 0.0f;
          const float cA = rinv6 * rinv6 * (ds_r - 12.0f * S * rinv2);
          const float cB = rinv6 * (6.0f * S * rinv2 - ds_r);
          const float u = r2 * args.c.ewald_2;
          const float P = HP0 + u * (HP1 + u * (HP2 + u * (HP3 + u * (HP4 + u * (HP5 + u * HP6)))));
          const float Q = 1.0f + u * (HQ1 + u * (HQ2 + u * (HQ3 + u * (HQ4 + u * (HQ5 + u * HQ6)))));
          const float fe = beta3 * (P * __builtin_amdgcn_rcpf(Q)) - rinv3;
          const int tj = s_type[wv][j];
          myBins[tj * 64] += sp * cA;
          myBins[(nt + tj) * 64] += sp * cB;
          const float se = sp * fe;
          gqi += se * xj.w;
          s_gqj[wv][h][j] += se * xi.w;
        }
      }
      waveSync();
      if (h == 0) {
        const float v = s_gqj[wv][0][i] + s_gqj[wv][1][i];
        if (v != 0.f) atomicAdd(&args.gq[ja + i], (double)v);
      }
      waveSync();
    }
    gqi += __shfl_down(gqi, 32, 64);
    if (h == 0 && gqi != 0.f) atomicAdd(&args.gq[iatom], (double)gqi);
    // flush this lane's FP32 bins (row t_i) into the workgroup's FP64 bins
    for (int b = 0; b < 2 * nt; ++b) {
      const float v = myBins[b * 64];
      if (v != 0.f) {
        const int bin = (b < nt) ? ti * nt + b : nt * nt + ti * nt + (b - nt);
        atomicAdd(&s_wg[bin], (double)v);
        myBins[b * 64] = 0.f;
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += ADJ_THREADS)
    args.binPartials[(size_t)blockIdx.x * nbins + k] = s_wg[k];
}

// one thread per bin, workgroups summed in order
__global__ void adjointBinsKernel(int nblocks, int nbins, const double* __restrict__ partials,
                                  double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbins) return;
  double v = 0.0;
  for (int b = 0; b < nblocks; ++b) v += partials[(size_t)b * nbins + k];
  out[k] = v;
}

int adjointGridBlocks() {
  static int numCU = 0;
  if (numCU == 0) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    numCU = prop.multiProcessorCount;
  }
  return numCU * 4;
}

size_t adjointSharedBytes(int ntypes) {
  return sizeof(double) * 2 * (size_t)ntypes * ntypes + sizeof(float) * ADJ_WAVES * 2 * (size_t)ntypes * 64;
}

int fmResidualBlocks(int n) { return (n + ADJ_RED - 1) / ADJ_RED; }

hipError_t launchForceMatchingAdjoint(const AdjointArgsHost& h, hipStream_t s) {
  const int nres = fmResidualBlocks(h.nslots);
  hipLaunchKernelGGL(fmResidualKernel, dim3(nres), dim3(ADJ_RED), 0, s, h.nslots, h.force, h.forceSlow, h.target,
                     h.g, h.lossPartials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sumDoublesKernel, dim3(1), dim3(ADJ_RED), 0, s, nres, h.lossPartials, h.loss);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  AdjArgs a;
  a.numLists = h.numLists;
  a.tileLists = h.tileLists;
  a.tileJatomStart = h.tileJatomStart;
  a.tileExcls = h.tileExcls;
  a.xyzq = h.xyzq;
  a.vdwTypes = h.vdwTypes;
  a.ntypes = h.ntypes;
  a.c = h.c;
  a.cut = h.cut;
  a.g = h.g;
  a.gq = h.gq;
  a.binPartials = h.binPartials;
  const int nb = h.gridBlocks;
  hipLaunchKernelGGL(adjointKernel, dim3(nb), dim3(ADJ_THREADS), adjointSharedBytes(h.ntypes), s, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int nbins = 2 * h.ntypes * h.ntypes;
  hipLaunchKernelGGL(adjointBinsKernel, dim3((nbins + 255) / 256), dim3(256), 0, s, nb, nbins, h.binPartials,
                     h.bins);
  return hipGetLastError();
}

}  // namespace mi355nb

// Global observables of the mesh solution: the reference's reset / map / merge
// cycle (Toolkit/exahype/toolkit/templates/solvers/
// AbstractGenericADERDGSolverImplementation.template:412-453) with the Euler
// example's observables (ApplicationExamples/Euler/Euler/
// EulerSolver_ADERDG.cpp:607-628: dgcells = 1 and fvcells = 0 per cell, merged
// by +=) and, per quantity k of nData, the integral sum_cells sum_nodes
// W_node h^D Q_k (merged by +=) and the nodal minimum and maximum (merged by
// min / max).
//
// Two passes with a fixed reduction order, so the result does not depend on
// scheduling: pass 1 gives each of nb workgroups a contiguous range of cells
// (thread-strided over its (cell, node) pairs, tree-reduced in LDS) and writes
// [sum | min | max] per quantity; pass 2 merges the nb partials in block order.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "engine.h"

namespace exa {
namespace {

constexpr int kObsThreads = 256;

template <int D>
__global__ __launch_bounds__(kObsThreads) void observables_partial_kernel(long ncells, int nb, int B, int ND,
                                                                          double hD, const double* __restrict__ luh,
                                                                          const DevTables* __restrict__ T,
                                                                          double* __restrict__ part) {
  __shared__ double ssum[kObsThreads], smin[kObsThreads], smax[kObsThreads];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int NS = (D == 3) ? B * B * B : B * B;
  const long c0 = ncells * b / nb, c1 = ncells * (b + 1) / nb;
  const long items = (c1 - c0) * NS;
  for (int k = 0; k < ND; k++) {
    double s = 0.0, mn = DBL_MAX, mx = -DBL_MAX;
    for (long i = tid; i < items; i += kObsThreads) {
      const long cell = c0 + i / NS;
      const int node = static_cast<int>(i % NS);
      const int ix = node % B, iy = (node / B) % B, iz = node / (B * B);
      const double W = (D == 3) ? T->w[iz] * T->w[iy] * T->w[ix] : T->w[iy] * T->w[ix];
      const double q = luh[(cell * NS + node) * ND + k];
      s += W * q;
      mn = fmin(mn, q);
      mx = fmax(mx, q);
    }
    ssum[tid] = s * hD;
    smin[tid] = mn;
    smax[tid] = mx;
    __syncthreads();
    for (int w = kObsThreads / 2; w > 0; w >>= 1) {
      if (tid < w) {
        ssum[tid] += ssum[tid + w];
        smin[tid] = fmin(smin[tid], smin[tid + w]);
        smax[tid] = fmax(smax[tid], smax[tid + w]);
      }
      __syncthreads();
    }
    if (tid == 0) {
      part[(long(b) * 3 + 0) * ND + k] = ssum[0];
      part[(long(b) * 3 + 1) * ND + k] = smin[0];
      part[(long(b) * 3 + 2) * ND + k] = smax[0];
    }
    __syncthreads();  // the LDS arrays are rewritten for the next quantity
  }
}

// out = [dgcells, fvcells, integral[ND], min[ND], max[ND]]
__global__ void observables_merge_kernel(long ncells, int nb, int ND, const double* __restrict__ part,
                                         double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k == 0) {
    out[0] = static_cast<double>(ncells);
    out[1] = 0.0;
  }
  if (k >= ND) return;
  double s = 0.0, mn = DBL_MAX, mx = -DBL_MAX;
  for (int b = 0; b < nb; b++) {
    s += part[(long(b) * 3 + 0) * ND + k];
    mn = fmin(mn, part[(long(b) * 3 + 1) * ND + k]);
    mx = fmax(mx, part[(long(b) * 3 + 2) * ND + k]);
  }
  out[2 + k] = s;
  out[2 + ND + k] = mn;
  out[2 + 2 * ND + k] = mx;
}

}  // namespace

int observables_blocks(long ncells) { return static_cast<int>(ncells < 1024 ? (ncells > 0 ? ncells : 1) : 1024); }

int launch_observables(const KernelCfg& k, hipStream_t s, long ncells, double h, const double* luh, double* part,
                       double* out) {
  const int B = k.order + 1, ND = k.nvar + k.npar;
  if (ND > 64) return -1;
  const int nb = observables_blocks(ncells);
  const double hD = (k.dim == 3) ? h * h * h : h * h;
  if (k.dim == 3)
    hipLaunchKernelGGL(observables_partial_kernel<3>, dim3(nb), dim3(kObsThreads), 0, s, ncells, nb, B, ND, hD, luh,
                       k.tab, part);
  else
    hipLaunchKernelGGL(observables_partial_kernel<2>, dim3(nb), dim3(kObsThreads), 0, s, ncells, nb, B, ND, hD, luh,
                       k.tab, part);
  hipLaunchKernelGGL(observables_merge_kernel, dim3(1), dim3(64), 0, s, ncells, nb, ND, part, out);
  return static_cast<int>(hipGetLastError());
}

}  // namespace exa

// Point probes of the mesh solution: the reference's probe plotter
// (ExaHyPE/exahype/plotters/ADERDG2ProbeAscii.cpp:127-180) takes the leaf cell
// with offset <= x < offset + size and evaluates kernels::legendre::interpolate
// (ExaHyPE/kernels/GaussLegendreBasis.cpp:21-49): the tensor-product Lagrange
// basis at xRef = (x - offset) / size times the nodal values, x fastest.
// Here one thread per (point, unknown) evaluates the basis in product form
// phi_i(xi) = prod_{m != i} (xi - x_m) / (x_i - x_m) (the reference expands the
// polynomials; values agree to rounding).
#include <hip/hip_runtime.h>

#include "engine.h"

namespace exa {
namespace {

__device__ __forceinline__ double lagrange(const double* __restrict__ xn, int B, int i, double xi) {
  double r = 1.0;
  for (int m = 0; m < B; m++)
    if (m != i) r *= (xi - xn[m]) / (xn[i] - xn[m]);
  return r;
}

template <int D>
__global__ void probe_kernel(int n, double h, int B, int ND, const double* __restrict__ luh, int npts,
                             const double* __restrict__ pts, const DevTables* __restrict__ T,
                             double* __restrict__ out) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= long(npts) * ND) return;
  const int p = static_cast<int>(gid / ND), k = static_cast<int>(gid - long(p) * ND);
  long c[3] = {0, 0, 0};
  double xr[3] = {0, 0, 0};
  for (int d = 0; d < D; d++) {
    const double x = pts[p * D + d];
    long i = static_cast<long>(floor(x / h));
    if (i < 0) i = 0;
    if (i > n - 1) i = n - 1;
    // containment offset <= x < offset + size, as the plotter tests it
    if (x < i * h && i > 0) i--;
    if (x >= (i + 1) * h && i < n - 1) i++;
    c[d] = i;
    xr[d] = (x - i * h) / h;
  }
  const long cell = (D == 3) ? (c[2] * n + c[1]) * n + c[0] : c[1] * n + c[0];
  const int NS = (D == 3) ? B * B * B : B * B;
  const double* u = luh + cell * NS * ND;
  const double* xn = T->x;
  double r = 0.0;
  for (int s = 0; s < NS; s++) {
    const int ix = s % B, iy = (s / B) % B, iz = s / (B * B);
    double b = lagrange(xn, B, ix, xr[0]) * lagrange(xn, B, iy, xr[1]);
    if (D == 3) b *= lagrange(xn, B, iz, xr[2]);
    r += b * u[s * ND + k];
  }
  out[gid] = r;
}

}  // namespace

int launch_probe(const KernelCfg& k, hipStream_t s, int n, double h, const double* luh, int npts, const double* pts,
                 double* out) {
  if (npts <= 0) return 0;
  const int B = k.order + 1, ND = k.nvar + k.npar;
  const long total = long(npts) * ND;
  const dim3 grid(static_cast<unsigned>((total + 255) / 256)), block(256);
  if (k.dim == 3)
    hipLaunchKernelGGL(probe_kernel<3>, grid, block, 0, s, n, h, B, ND, luh, npts, pts, k.tab, out);
  else
    hipLaunchKernelGGL(probe_kernel<2>, grid, block, 0, s, n, h, B, ND, luh, npts, pts, k.tab, out);
  return static_cast<int>(hipGetLastError());
}

}  // namespace exa

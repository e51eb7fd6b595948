// Initial conditions evaluated on the device at the Gauss-Legendre nodes of
// every cell, the way the reference's solutionAdjustment kernel calls the
// user's adjustPointSolution at t = 0
// (ExaHyPE/kernels/aderdg/generic/c/3d/solutionAdjustment.cpph:18-63:
//  x = centre + dx * (xi - 1/2)).
//
// EXA_IC_DENSITY_WAVE (Euler, gamma = 1.4), params:
//   [0] a0   amplitude of the base wave  rho += a0 sin(2 pi (x + y [+ z]))
//   [1..3] u0, v0, w0  base velocity
//   [4] p0   pressure
//   [5] M    number of Fourier modes (per field), [6] amp  amplitude per mode
//   then M x (kx, ky, kz, phase) for rho, then M x (kx, ky, kz, phase) for u:
//   rho += amp cos(2 pi k.x + phase), u += amp cos(2 pi k.x + phase)
#include <hip/hip_runtime.h>

#include "engine.h"
#include "pde.h"

namespace exa {
namespace {

__global__ void density_wave_kernel(long ncells, int n, int D, int B, int ND, double h, const double* __restrict__ xi,
                                    const double* __restrict__ prm, double* __restrict__ luh) {
  const long NS = (D == 3) ? long(B) * B * B : long(B) * B;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ncells * NS) return;
  const long cell = gid / NS;
  const int s = static_cast<int>(gid - cell * NS);
  const int jx = s % B, jy = (s / B) % B, jz = (D == 3) ? s / (B * B) : 0;
  const long ix = cell % n, iy = (cell / n) % n, iz = (D == 3) ? cell / (long(n) * n) : 0;
  const double x = (ix + 0.5) * h + h * (xi[jx] - 0.5);
  const double y = (iy + 0.5) * h + h * (xi[jy] - 0.5);
  const double z = (D == 3) ? (iz + 0.5) * h + h * (xi[jz] - 0.5) : 0.0;
  const double two_pi = 6.283185307179586476925286766559;
  const double a0 = prm[0], u0 = prm[1], v0 = prm[2], w0 = prm[3], p0 = prm[4];
  const int M = static_cast<int>(prm[5]);
  const double amp = prm[6];
  double rho = 1.0 + a0 * sin(two_pi * ((D == 3) ? (x + y + z) : (x + y)));
  double u = u0;
  const double* mr = prm + 7;
  const double* mu = prm + 7 + 4 * M;
  for (int m = 0; m < M; m++) {
    rho += amp * cos(two_pi * (mr[4 * m] * x + mr[4 * m + 1] * y + mr[4 * m + 2] * z) + mr[4 * m + 3]);
    u += amp * cos(two_pi * (mu[4 * m] * x + mu[4 * m + 1] * y + mu[4 * m + 2] * z) + mu[4 * m + 3]);
  }
  const double v = v0, w = (D == 3) ? w0 : 0.0;
  double* Q = luh + gid * ND;
  Q[0] = rho;
  Q[1] = rho * u;
  Q[2] = rho * v;
  Q[3] = rho * w;
  Q[4] = p0 / (1.4 - 1) + 0.5 * rho * (u * u + v * v + w * w);
  for (int k = 5; k < ND; k++) Q[k] = 0.0;
}

// EXA_IC_ENTROPY_WAVE: the exact entropy wave at time t at every node
template <int D>
__global__ void entropy_wave_kernel(long ncells, int n, int B, int ND, double h, const double* __restrict__ xi,
                                    double t, double* __restrict__ luh) {
  const long NS = (D == 3) ? long(B) * B * B : long(B) * B;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ncells * NS) return;
  const long cell = gid / NS;
  const int s = static_cast<int>(gid - cell * NS);
  const int jx = s % B, jy = (s / B) % B, jz = (D == 3) ? s / (B * B) : 0;
  const long ix = cell % n, iy = (cell / n) % n, iz = (D == 3) ? cell / (long(n) * n) : 0;
  double x[3];
  x[0] = (ix + 0.5) * h + h * (xi[jx] - 0.5);
  x[1] = (iy + 0.5) * h + h * (xi[jy] - 0.5);
  x[2] = (D == 3) ? (iz + 0.5) * h + h * (xi[jz] - 0.5) : 0.5;
  double* Q = luh + gid * ND;
  euler_entropy_wave<D>(x, t, Q);
  for (int k = 5; k < ND; k++) Q[k] = 0.0;
}

}  // namespace

int launch_entropy_wave(hipStream_t s, long ncells, int n, int D, int B, int ND, double h, const double* d_xi,
                        double t, double* luh) {
  const long total = ncells * ((D == 3) ? long(B) * B * B : long(B) * B);
  if (total <= 0) return 0;
  if (D == 3)
    hipLaunchKernelGGL(entropy_wave_kernel<3>, dim3((total + 255) / 256), dim3(256), 0, s, ncells, n, B, ND, h, d_xi,
                       t, luh);
  else
    hipLaunchKernelGGL(entropy_wave_kernel<2>, dim3((total + 255) / 256), dim3(256), 0, s, ncells, n, B, ND, h, d_xi,
                       t, luh);
  return static_cast<int>(hipGetLastError());
}

int launch_density_wave(hipStream_t s, long ncells, int n, int D, int B, int ND, double h, const double* d_xi,
                        const double* d_prm, double* luh) {
  const long total = ncells * ((D == 3) ? long(B) * B * B : long(B) * B);
  if (total <= 0) return 0;
  hipLaunchKernelGGL(density_wave_kernel, dim3((total + 255) / 256), dim3(256), 0, s, ncells, n, D, B, ND, h, d_xi,
                     d_prm, luh);
  return static_cast<int>(hipGetLastError());
}

}  // namespace exa

// Boundary conditions of the nonlinear (Euler) path: the outside state and
// flux of every boundary face node, written into a ghost-face buffer that the
// Riemann kernels then read in place of a neighbour's extrapolated face.
//
// Reference:
//   kernel  ExaHyPE/kernels/aderdg/generic/c/3d/boundaryConditions.cpph:29-122
//           (2D: generic/c/2d/boundaryConditions.cpph:28-100): the hook at each
//           face node x[normal] = centre + (orientation - 1/2) dx; of the two
//           free directions, d2 (slow face index) and d1 (fast) are
//           normal 0 -> (z, y), normal 1 -> (z, x), normal 2 -> (y, x);
//   glue    AbstractGenericADERDGSolverImplementation.template:307-350: the
//           outside data goes on the left of the Riemann problem on a left
//           face and on the right on a right face;
//   hooks   ApplicationExamples/Euler/Euler/EulerSolver_ADERDG.cpp:525-554
//           (reflecting wall; Dirichlet = Gauss time average of the entropy
//           wave at the predictor's time nodes) and :266-286 (entropyWave).
// The ghost buffer replaces the reference's per-call stack arrays
// (stateOut/fluxOut); the Riemann solve itself is the unchanged Rusanov
// kernel, so interior and boundary faces share one code path.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "pde.h"

namespace exa {
namespace {

constexpr int kV = 5;  // Euler

// outside state (ND entries) and normal flux (5) of one face node
template <int D>
__device__ __forceinline__ void boundary_node(int kind, const double* x, double t, double dt, int d, int ND, int B,
                                              const double* __restrict__ qin, double* __restrict__ qout,
                                              double* __restrict__ fout, const DevTables* __restrict__ T,
                                              const PdeParams& P) {
  for (int k = kV; k < ND; k++) qout[k] = qin[k];
  double F[D][kV];
  if (kind == kBcWall) {
    double Q[kV];
#pragma unroll
    for (int v = 0; v < kV; v++) Q[v] = qin[v];
#pragma unroll
    for (int dd = 0; dd < D; dd++)
      if (dd == d) Q[1 + dd] = -Q[1 + dd];
    pde_flux<D, double>(P, Q, F);
#pragma unroll
    for (int v = 0; v < kV; v++) qout[v] = Q[v];
#pragma unroll
    for (int dd = 0; dd < D; dd++)
      if (dd == d)
#pragma unroll
        for (int v = 0; v < kV; v++) fout[v] = F[dd][v];
    return;
  }
  double qa[kV] = {0, 0, 0, 0, 0}, fa[kV] = {0, 0, 0, 0, 0};
  for (int i = 0; i < B; i++) {
    double Q[kV];
    euler_entropy_wave<D>(x, t + dt * T->x[i], Q);
    pde_flux<D, double>(P, Q, F);
    const double wi = T->w[i];
#pragma unroll
    for (int dd = 0; dd < D; dd++)
      if (dd == d)
#pragma unroll
        for (int v = 0; v < kV; v++) {
          qa[v] += Q[v] * wi;
          fa[v] += F[dd][v] * wi;
        }
  }
#pragma unroll
  for (int v = 0; v < kV; v++) {
    qout[v] = qa[v];
    fout[v] = fa[v];
  }
}

// coordinates of face node fn of face (d, o) of the cell with this centre
template <int D>
__device__ __forceinline__ void face_node_x(const double* centre, const double* dx, int d, int o, int fn, int B,
                                            const double* __restrict__ xi, double* x) {
  x[d] = centre[d] + (-0.5 + o) * dx[d];
  if (D == 3) {
    const int d1 = (d == 0) ? 1 : 0, d2 = (d == 2) ? 1 : 2;
    const int i = fn / B, j = fn - i * B;
    x[d2] = centre[d2] + dx[d2] * (xi[i] - 0.5);
    x[d1] = centre[d1] + dx[d1] * (xi[j] - 0.5);
  } else {
    const int d1 = (d == 0) ? 1 : 0;
    x[d1] = centre[d1] + dx[d1] * (xi[fn] - 0.5);
  }
}

// Uniform n^D mesh on [0, n h]^D, cells x fastest.  Ghost face g = f * n^(D-1)
// + (transverse cell index) for domain face f = 2d + o; transverse index
// (iz, iy) for d = 0, (iz, ix) for d = 1, (iy, ix) for d = 2 (2D: iy, ix).
// One thread per ghost face node; faces of periodic directions are skipped.
// t and dt are the step's, read from the device state {t, dt, ...}.
template <int D>
__global__ __launch_bounds__(256) void boundary_mesh_kernel(int n, int B, int ND, double h, BcKinds bk,
                                                            const double* __restrict__ qb, double* __restrict__ gq,
                                                            double* __restrict__ gf, const DevTables* __restrict__ T,
                                                            PdeParams P, const double* __restrict__ state) {
  const int NF = (D == 3) ? B * B : B;
  const long per = (D == 3) ? long(n) * n : long(n);
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= 2 * D * per * NF) return;
  const long g = gid / NF;
  const int fn = static_cast<int>(gid - g * NF);
  const int f = static_cast<int>(g / per);
  const long tr = g - f * per;
  const int kind = bk.k[f];
  if (kind == kBcPeriodic || kind == kBcExchange) return;  // exchange faces: ghost data come from a halo message
  const int d = f >> 1, o = f & 1;
  const long id = o ? n - 1 : 0;
  long ix = 0, iy = 0, iz = 0;
  if (D == 3) {
    const long a = tr / n, b = tr - (tr / n) * n;
    if (d == 0) {
      iz = a, iy = b, ix = id;
    } else if (d == 1) {
      iz = a, ix = b, iy = id;
    } else {
      iy = a, ix = b, iz = id;
    }
  } else {
    if (d == 0) {
      iy = tr, ix = id;
    } else {
      ix = tr, iy = id;
    }
  }
  const long cell = (D == 3) ? (iz * n + iy) * n + ix : iy * n + ix;
  const double centre[3] = {(ix + 0.5) * h, (iy + 0.5) * h, (iz + 0.5) * h};
  const double dx[3] = {h, h, h};
  double x[3] = {0.5, 0.5, 0.5};
  face_node_x<D>(centre, dx, d, o, fn, B, T->x, x);
  boundary_node<D>(kind, x, state[0], state[1], d, ND, B, qb + ((cell * 2 * D + f) * NF + fn) * ND,
                   gq + (g * NF + fn) * ND, gf + (g * NF + fn) * kV, T, P);
}

// cell on domain face f = 2d + o with transverse index tr (ghost-face order)
template <int D>
__device__ __forceinline__ long boundary_cell(int n, int f, long tr) {
  const int d = f >> 1, o = f & 1;
  const long id = o ? n - 1 : 0;
  if (D == 3) {
    const long a = tr / n, b = tr - (tr / n) * n;
    if (d == 0) return (a * n + b) * n + id;
    if (d == 1) return (a * n + id) * n + b;
    return (id * n + a) * n + b;
  }
  return (d == 0) ? tr * n + id : id * n + tr;
}

template <int D>
__global__ __launch_bounds__(256) void halo_pack_kernel(int n, int f, int NF, int ND, int V,
                                                        const double* __restrict__ qb, const double* __restrict__ fb,
                                                        double* __restrict__ msg_q, double* __restrict__ msg_f) {
  const long per = (D == 3) ? long(n) * n : long(n);
  const long nq = per * NF * ND, nf = per * NF * V;
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid < nq) {
    const long tr = gid / (long(NF) * ND), r = gid - tr * NF * ND;
    msg_q[gid] = qb[(boundary_cell<D>(n, f, tr) * 2 * D + f) * NF * ND + r];
  } else if (gid < nq + nf) {
    const long i = gid - nq;
    const long tr = i / (long(NF) * V), r = i - tr * NF * V;
    msg_f[i] = fb[(boundary_cell<D>(n, f, tr) * 2 * D + f) * NF * V + r];
  }
}

struct Geometry {
  double centre[3], dx[3];
};

// one face of one cell (the compatibility entry point): qin [NF][ND]
template <int D>
__global__ __launch_bounds__(64) void boundary_face_kernel(int kind, Geometry geo, double t, double dt, int d, int o,
                                                           int B, int ND, const double* __restrict__ qin,
                                                           double* __restrict__ gq, double* __restrict__ gf,
                                                           const DevTables* __restrict__ T, PdeParams P) {
  const int NF = (D == 3) ? B * B : B;
  for (int fn = threadIdx.x; fn < NF; fn += blockDim.x) {
    double x[3] = {0.5, 0.5, 0.5};
    face_node_x<D>(geo.centre, geo.dx, d, o, fn, B, T->x, x);
    boundary_node<D>(kind, x, t, dt, d, ND, B, qin + fn * ND, gq + fn * ND, gf + fn * kV, T, P);
  }
}

}  // namespace

bool boundary_supported(const KernelCfg& k) { return k.pde.id == kPdeEuler && k.nvar == kV; }

int launch_boundary_mesh(const KernelCfg& k, hipStream_t s, int n, double h, const BcKinds& bk, const double* qb,
                         double* gq, double* gf, const double* state) {
  if (!boundary_supported(k)) return -1;
  const int D = k.dim, B = k.order + 1, ND = k.nvar + k.npar;
  const long NF = (D == 3) ? long(B) * B : long(B);
  const long total = 2L * D * ((D == 3) ? long(n) * n : long(n)) * NF;
  const dim3 grid(static_cast<unsigned>((total + 255) / 256)), block(256);
  if (D == 3)
    hipLaunchKernelGGL(boundary_mesh_kernel<3>, grid, block, 0, s, n, B, ND, h, bk, qb, gq, gf, k.tab, k.pde, state);
  else
    hipLaunchKernelGGL(boundary_mesh_kernel<2>, grid, block, 0, s, n, B, ND, h, bk, qb, gq, gf, k.tab, k.pde, state);
  return static_cast<int>(hipGetLastError());
}

int launch_boundary_face(const KernelCfg& k, hipStream_t s, int kind, const double* centre, const double* dx,
                         double t, double dt, int d, int o, const double* qin, double* gq, double* gf) {
  if (!boundary_supported(k)) return -1;
  const int D = k.dim, B = k.order + 1, ND = k.nvar + k.npar;
  Geometry geo{};
  for (int i = 0; i < 3; i++) {
    geo.centre[i] = i < D ? centre[i] : 0.5;
    geo.dx[i] = i < D ? dx[i] : 1.0;
  }
  if (D == 3)
    hipLaunchKernelGGL(boundary_face_kernel<3>, dim3(1), dim3(64), 0, s, kind, geo, t, dt, d, o, B, ND, qin, gq, gf,
                       k.tab, k.pde);
  else
    hipLaunchKernelGGL(boundary_face_kernel<2>, dim3(1), dim3(64), 0, s, kind, geo, t, dt, d, o, B, ND, qin, gq, gf,
                       k.tab, k.pde);
  return static_cast<int>(hipGetLastError());
}

int launch_halo_pack(const KernelCfg& k, hipStream_t s, int n, int f, const double* qb, const double* fb,
                     double* msg_q, double* msg_f) {
  const int D = k.dim, B = k.order + 1, ND = k.nvar + k.npar, V = k.nvar;
  const int NF = (D == 3) ? B * B : B;
  const long per = (D == 3) ? long(n) * n : long(n);
  const long total = per * NF * (ND + V);
  const dim3 grid(static_cast<unsigned>((total + 255) / 256)), block(256);
  if (D == 3)
    hipLaunchKernelGGL(halo_pack_kernel<3>, grid, block, 0, s, n, f, NF, ND, V, qb, fb, msg_q, msg_f);
  else
    hipLaunchKernelGGL(halo_pack_kernel<2>, grid, block, 0, s, n, f, NF, ND, V, qb, fb, msg_q, msg_f);
  return static_cast<int>(hipGetLastError());
}

}  // namespace exa

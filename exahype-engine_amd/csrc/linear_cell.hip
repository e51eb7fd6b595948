// Per-cell kernels of the linear (Cauchy-Kowalevski) ADER-DG path for the
// acoustic PDE, used by the compatibility entry points (reference per-cell
// layouts, per-node parameters allowed).  One workgroup per call.
//   lin_cell_predictor_kernel  spaceTimePredictorLinear.cpph:30-647 (point-source
//                              terms :111-182, :482-503) + volumeIntegralLinear.cpph:21-70
//   lin_cell_riemann_kernel    riemannSolverLinear.cpph:23-167
//   lin_cell_face_kernel       faceIntegralLinear.cpph:8-101
//   lin_cell_update_kernel     solutionUpdate.cpph:28-67
//   lin_cell_dt_kernel         stableTimeStepSize.cpph:14-62 with eigenvalues (c, -c, 0, 0)
//   lin_cell_split_kernel      the optimised kernels' split Cauchy-Kowalevski predictor +
//                              volume integral (KernelGenerator/kernelgenerator/templates/
//                              aderdg/fusedSPTVI_linear_split_ck_cpp.template)
// Hooks: ApplicationExamples/Linear/LinearWave3D/MyLinearWaveSolver.cpp:108-209.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "engine.h"

namespace exa {
namespace {

constexpr int LT = 256;

template <int B>
__device__ __forceinline__ int line_node3(int d, int x, int y, int z, int m) {
  return d == 0 ? (z * B + y) * B + m : (d == 1 ? (z * B + m) * B + x : (m * B + y) * B + x);
}

// srow (nullable): {sigma0 = dt sum_t w_t f_t, s_1 .. s_{B-1}, sum_t w_t f_t}; sg[N]: spatial factor
template <int B>
__global__ __launch_bounds__(LT) void lin_cell_predictor_kernel(const double* __restrict__ luh, double dx0,
                                                                double dx1, double dx2, double dt,
                                                                const double* __restrict__ sg,
                                                                const double* __restrict__ srow,
                                                                double* __restrict__ lduh, double* __restrict__ lQb,
                                                                double* __restrict__ lFb,
                                                                const DevTables* __restrict__ T) {
  constexpr int N = B * B * B, F = B * B;
  __shared__ double sL[4][N];
  __shared__ double sQh[6][N];
  __shared__ double sFh[3][4][N];
  const int nd = threadIdx.x;
  const bool on = nd < N;
  const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
  const double dxv[3] = {dx0, dx1, dx2};
  double q[4] = {0, 0, 0, 0}, rho = 1.0, c = 1.0;
  if (on) {
    for (int v = 0; v < 4; v++) q[v] = luh[nd * 6 + v];
    rho = luh[nd * 6 + 4];
    c = luh[nd * 6 + 5];
  }
  const double lam = rho * c * c, irho = 1.0 / rho;
  const double g = (srow && on) ? sg[nd] : 0.0;
  if (srow) q[0] += srow[0] * g;
  double a[B];
  a[0] = 1.0;
  {
    double f = 0.5 * dt;
    for (int i = 1; i < B; i++) {
      a[i] = f;
      f *= dt / (i + 2);
    }
  }
  double qh[4] = {q[0], q[1], q[2], q[3]};
  double fh[3][4];
  for (int d = 0; d < 3; d++)
    for (int v = 0; v < 4; v++) fh[d][v] = 0.0;
#pragma unroll
  for (int i = 0; i < B; i++) {
    __syncthreads();
    if (on)
      for (int v = 0; v < 4; v++) sL[v][nd] = q[v];
    __syncthreads();
    if (on) {
      double Dp[3], Du[3];
      const int c3[3] = {x, y, z};
      for (int d = 0; d < 3; d++) {
        double sp = 0.0, su = 0.0;
        for (int m = 0; m < B; m++) {
          const double dd = T->dudx[c3[d] * B + m] / dxv[d];
          const int nn = line_node3<B>(d, x, y, z, m);
          sp += sL[0][nn] * dd;
          su += sL[1 + d][nn] * dd;
        }
        Dp[d] = sp;
        Du[d] = su;
      }
      // lFi[d][0] = lambda * (-D_d u_d), lFi[d][1+d] = (1/rho) * (-D_d p)
      double nq[4] = {0, 0, 0, 0};
      for (int d = 0; d < 3; d++) {
        const double f0 = lam * (-Du[d]);
        const double f1 = 1.0 / rho * (-Dp[d]);
        fh[d][0] += a[i] * f0;
        fh[d][1 + d] += a[i] * f1;
        nq[0] -= f0;
        nq[1 + d] -= f1;
      }
      for (int v = 0; v < 4; v++) q[v] = nq[v];
      if (i + 1 < B) {
        if (srow) q[0] += srow[i + 1] * g;
        for (int v = 0; v < 4; v++) qh[v] += a[i + 1] * q[v];
      }
    }
  }
  (void)irho;
  if (on && srow)
    for (int d = 0; d < 3; d++) fh[d][0] += srow[B] * g / 3;
  // volume integral: lduh = -W sum_d lFhi[d]
  if (on) {
    const double W = T->w[z] * T->w[y] * T->w[x];
    for (int v = 0; v < 4; v++) {
      double r = 0.0;
      for (int d = 0; d < 3; d++) r -= W * fh[d][v];
      lduh[nd * 4 + v] = r;
    }
    for (int v = 0; v < 4; v++) sQh[v][nd] = qh[v];
    sQh[4][nd] = rho;
    sQh[5][nd] = c;
    for (int d = 0; d < 3; d++)
      for (int v = 0; v < 4; v++) sFh[d][v][nd] = fh[d][v];
  }
  __syncthreads();
  // extrapolation of lQhi (with parameters) and of lFhi[d] to the faces
  for (int task = nd; task < 6 * F; task += LT) {
    const int f = task / F, fn = task - f * F, d = f >> 1, o = f & 1;
    const int fa = fn / B, fb = fn - fa * B;
    double sq[6] = {0, 0, 0, 0, 0, 0}, sf[4] = {0, 0, 0, 0};
    for (int l = 0; l < B; l++) {
      const int node = d == 0 ? (fa * B + fb) * B + l : (d == 1 ? (fa * B + l) * B + fb : (l * B + fa) * B + fb);
      const double fc = o ? T->F1[l] : T->F0[l];
      for (int k = 0; k < 6; k++) sq[k] += sQh[k][node] * fc;
      for (int v = 0; v < 4; v++) sf[v] += sFh[d][v][node] * fc;
    }
    for (int k = 0; k < 6; k++) lQb[(f * F + fn) * 6 + k] = sq[k];
    for (int v = 0; v < 4; v++) lFb[(f * F + fn) * 4 + v] = sf[v];
  }
}

template <int B>
__global__ __launch_bounds__(64) void lin_cell_riemann_kernel(double* __restrict__ FL, double* __restrict__ FR,
                                                              const double* __restrict__ QL,
                                                              const double* __restrict__ QR, int d,
                                                              const DevTables* __restrict__ T) {
  constexpr int F = B * B;
  __shared__ double sav[2][6];
  if (threadIdx.x == 0) {
    // averageRiemannInputs: sum over (i, j) of w_i w_j Q
    double aL[6] = {0, 0, 0, 0, 0, 0}, aR[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double wij = T->w[i] * T->w[j];
        for (int k = 0; k < 6; k++) {
          aL[k] += wij * QL[(i * B + j) * 6 + k];
          aR[k] += wij * QR[(i * B + j) * 6 + k];
        }
      }
    for (int k = 0; k < 6; k++) {
      sav[0][k] = aL[k];
      sav[1][k] = aR[k];
    }
  }
  __syncthreads();
  const double sL = fabs(sav[0][5]), sR = fabs(sav[1][5]);
  const double smax = sL > sR ? sL : sR;
  const double rhoa = 0.5 * (sav[1][4] + sav[0][4]);
  const double ca = 0.5 * (sav[1][5] + sav[0][5]);
  const double lama = rhoa * ca * ca;
  for (int fn = threadIdx.x; fn < F; fn += 64) {
    double qd[4], acc[4] = {0, 0, 0, 0};
    for (int v = 0; v < 4; v++) qd[v] = 0.5 * (QR[fn * 6 + v] - QL[fn * 6 + v]);
    acc[1 + d] += 1.0 / rhoa * (-qd[0]);
    acc[0] += lama * (-qd[1 + d]);
    for (int v = 0; v < 4; v++) {
      FL[fn * 4 + v] = acc[v] - smax * qd[v];
      FR[fn * 4 + v] = acc[v] + smax * qd[v];
    }
  }
}

template <int B>
__global__ __launch_bounds__(LT) void lin_cell_face_kernel(double* __restrict__ lduh, const double* __restrict__ Ff,
                                                           int d, int o, double dxd, const DevTables* __restrict__ T) {
  constexpr int N = B * B * B;
  const int nd = threadIdx.x;
  if (nd >= N) return;
  const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
  const int c3[3] = {x, y, z};
  double sw = 1.0 / dxd;
  for (int e = 0; e < 3; e++)
    if (e != d) sw *= T->w[c3[e]];
  const int fn = d == 0 ? z * B + y : (d == 1 ? z * B + x : y * B + x);
  const double cf = sw * (o ? T->F1[c3[d]] : T->F0[c3[d]]);
  for (int v = 0; v < 4; v++) lduh[nd * 4 + v] -= cf * Ff[fn * 4 + v];
}

template <int B>
__global__ __launch_bounds__(LT) void lin_cell_update_kernel(double* __restrict__ luh, const double* __restrict__ lduh,
                                                             double dt, const DevTables* __restrict__ T) {
  constexpr int N = B * B * B;
  const int nd = threadIdx.x;
  if (nd >= N) return;
  const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
  const double upd = dt / (T->w[z] * T->w[y] * T->w[x]);
  for (int v = 0; v < 4; v++) luh[nd * 6 + v] += lduh[nd * 4 + v] * upd;
}

template <int B>
__global__ __launch_bounds__(LT) void lin_cell_dt_kernel(const double* __restrict__ luh, double dx0, double dx1,
                                                         double dx2, double cflpnpm, double* __restrict__ out) {
  constexpr int N = B * B * B;
  __shared__ double sm[LT];
  const int nd = threadIdx.x;
  double r = DBL_MAX;
  if (nd < N) {
    const double c = fabs(luh[nd * 6 + 5]);
    const double den = c / dx0 + c / dx1 + c / dx2;
    r = cflpnpm / den;
  }
  sm[nd] = r;
  __syncthreads();
  if (nd == 0) {
    double m = DBL_MAX;
    for (int k = 0; k < N; k++) m = fmin(m, sm[k]);
    out[0] = m;
  }
}

// Split Cauchy-Kowalevski predictor (fusedSPTVI_linear_split_ck_cpp.template;
// "tpl" below).  One derivative level L is kept; with the acoustic hooks
//   L_{k+1}[p] = lambda sum_d D_d u_d + s_k g,   L_{k+1}[u_d] = D_d p
// (tpl:201-331: the material matrix multiplies the flux part only, the NCP
// -D_d p enters unscaled), k = 0..B-1, and qh = L_0 + sum_k f_k L_{k+1} with
// the template's factors f_0 = dt/2, f_{k+1} = f_k dt/(k+2) (tpl:196, :346-353).
// Then per direction lFhi_d[p] = g sum_t w_t f_t / 3 - lambda D_d qh_{u_d},
// lFhi_d[u_d] = -D_d qh_p (tpl:474-640), lduh = -W sum_d lFhi_d, and qh (with
// the parameters) and lFhi_d extrapolated to the faces.
// srow (nullable): {sigma0 = dt sum_t w_t f_t, s_0 .. s_{B-1}, sum_t w_t f_t},
// f_t the source at the Gauss time nodes, s_k the k-th derivative at 0 of its
// interpolant on the unit interval; sg[N]: spatial factor.
template <int B>
__global__ __launch_bounds__(LT) void lin_cell_split_kernel(const double* __restrict__ luh, double dx0, double dx1,
                                                            double dx2, double dt, const double* __restrict__ sg,
                                                            const double* __restrict__ srow,
                                                            double* __restrict__ lduh, double* __restrict__ lQb,
                                                            double* __restrict__ lFb,
                                                            const DevTables* __restrict__ T) {
  constexpr int N = B * B * B, F = B * B;
  __shared__ double sL[4][N];
  __shared__ double sQh[6][N];
  __shared__ double sFh[3][4][N];
  const int nd = threadIdx.x;
  const bool on = nd < N;
  const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
  const int c3[3] = {x, y, z};
  const double idx[3] = {1.0 / dx0, 1.0 / dx1, 1.0 / dx2};
  double q[4] = {0, 0, 0, 0}, rho = 1.0, c = 1.0;
  if (on) {
    for (int v = 0; v < 4; v++) q[v] = luh[nd * 6 + v];
    rho = luh[nd * 6 + 4];
    c = luh[nd * 6 + 5];
  }
  const double lam = rho * c * c;
  const double g = (srow && on) ? sg[nd] : 0.0;
  if (srow) q[0] += srow[0] * g;  // tpl:181-188
  double qh[4] = {q[0], q[1], q[2], q[3]};
  double fac = 0.5 * dt;
  // D_d of p and of u_d at this node from the level in sL
  auto derivatives = [&](double (&Dp)[3], double (&Du)[3]) {
    for (int d = 0; d < 3; d++) {
      double sp = 0.0, su = 0.0;
      for (int m = 0; m < B; m++) {
        const double dd = idx[d] * T->dudx[c3[d] * B + m];
        const int nn = line_node3<B>(d, x, y, z, m);
        sp += sL[0][nn] * dd;
        su += sL[1 + d][nn] * dd;
      }
      Dp[d] = sp;
      Du[d] = su;
    }
  };
#pragma unroll 1
  for (int k = 0; k < B; k++) {
    __syncthreads();
    if (on)
      for (int v = 0; v < 4; v++) sL[v][nd] = q[v];
    __syncthreads();
    if (on) {
      double Dp[3], Du[3];
      derivatives(Dp, Du);
      q[0] = lam * (Du[0] + Du[1] + Du[2]);
      for (int d = 0; d < 3; d++) q[1 + d] = Dp[d];
      if (srow) q[0] += srow[1 + k] * g;
      for (int v = 0; v < 4; v++) qh[v] += fac * q[v];
      fac *= dt / (k + 2);
    }
  }
  // lFhi_d from qh
  __syncthreads();
  if (on)
    for (int v = 0; v < 4; v++) sL[v][nd] = qh[v];
  __syncthreads();
  if (on) {
    double Dp[3], Du[3];
    derivatives(Dp, Du);
    const double ps = srow ? srow[B + 1] * g / 3 : 0.0;
    const double W = T->w[z] * T->w[y] * T->w[x];
    double fh[3][4];
    for (int d = 0; d < 3; d++) {
      for (int v = 0; v < 4; v++) fh[d][v] = 0.0;
      fh[d][0] = ps - lam * Du[d];
      fh[d][1 + d] = -Dp[d];
    }
    for (int v = 0; v < 4; v++) {
      double r = 0.0;
      for (int d = 0; d < 3; d++) r -= fh[d][v] * W;
      lduh[nd * 4 + v] = r;
    }
    for (int v = 0; v < 4; v++) sQh[v][nd] = qh[v];
    sQh[4][nd] = rho;
    sQh[5][nd] = c;
    for (int d = 0; d < 3; d++)
      for (int v = 0; v < 4; v++) sFh[d][v][nd] = fh[d][v];
  }
  __syncthreads();
  for (int task = nd; task < 6 * F; task += LT) {
    const int f = task / F, fn = task - f * F, d = f >> 1, o = f & 1;
    const int fa = fn / B, fb = fn - fa * B;
    double sq[6] = {0, 0, 0, 0, 0, 0}, sf[4] = {0, 0, 0, 0};
    for (int l = 0; l < B; l++) {
      const int node = d == 0 ? (fa * B + fb) * B + l : (d == 1 ? (fa * B + l) * B + fb : (l * B + fa) * B + fb);
      const double fc = o ? T->F1[l] : T->F0[l];
      for (int k = 0; k < 6; k++) sq[k] += sQh[k][node] * fc;
      for (int v = 0; v < 4; v++) sf[v] += sFh[d][v][node] * fc;
    }
    for (int k = 0; k < 6; k++) lQb[(f * F + fn) * 6 + k] = sq[k];
    for (int v = 0; v < 4; v++) lFb[(f * F + fn) * 4 + v] = sf[v];
  }
}

template <int B>
int launch_b(const LinearCellArgs& a, hipStream_t s) {
  const DevTables* T = a.tab;
  switch (a.op) {
    case kLinPredictor:
      lin_cell_predictor_kernel<B><<<1, LT, 0, s>>>(a.luh, a.dx[0], a.dx[1], a.dx[2], a.dt, a.sg, a.srow, a.lduh,
                                                     a.lQb, a.lFb, T);
      break;
    case kLinPredictorSplit:
      lin_cell_split_kernel<B><<<1, LT, 0, s>>>(a.luh, a.dx[0], a.dx[1], a.dx[2], a.dt, a.sg, a.srow, a.lduh, a.lQb,
                                                 a.lFb, T);
      break;
    case kLinRiemann:
      lin_cell_riemann_kernel<B><<<1, 64, 0, s>>>(a.FL, a.FR, a.QL, a.QR, a.direction, T);
      break;
    case kLinFace:
      lin_cell_face_kernel<B><<<1, LT, 0, s>>>(a.lduh, a.Fface, a.direction, a.orientation, a.dx[a.direction], T);
      break;
    case kLinUpdate:
      lin_cell_update_kernel<B><<<1, LT, 0, s>>>(a.luh_out, a.lduh, a.dt, T);
      break;
    case kLinDt:
      lin_cell_dt_kernel<B><<<1, LT, 0, s>>>(a.luh, a.dx[0], a.dx[1], a.dx[2], a.cflpnpm, a.out);
      break;
    default:
      return -1;
  }
  return static_cast<int>(hipGetLastError());
}

}  // namespace

int launch_linear_cell(int order, const LinearCellArgs& a, hipStream_t s) {
  switch (order) {
    case 1: return launch_b<2>(a, s);
    case 2: return launch_b<3>(a, s);
    case 3: return launch_b<4>(a, s);
    case 4: return launch_b<5>(a, s);
    case 5: return launch_b<6>(a, s);
    default: return -1;
  }
}

}  // namespace exa

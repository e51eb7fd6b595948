// Linear acoustics (3D; variables p, u, v, w; per-cell material rho, c) on a
// device-resident uniform mesh: forward ADER-DG step and its discrete adjoint.
//
// Forward step (reference semantics, generic linear kernels of
// ExaHyPE/kernels/aderdg/generic/c/3d with the hooks of
// ApplicationExamples/Linear/LinearWave3D/MyLinearWaveSolver.cpp):
//   ac_predict_kernel    spaceTimePredictorLinear.cpph:30-647 (Cauchy-Kowalevski
//                        levels, point-source terms :111-182, Taylor averages
//                        :395-503, extrapolation :507-640) fused with
//                        volumeIntegralLinear.cpph:21-70 and the volume part of
//                        the solution update (solutionUpdate.cpph).  With the
//                        acoustic hooks one level is
//                          L_{i+1}[p]   = lambda * sum_d D_d u_d   (flux -u_d, times lambda = rho c^2)
//                          L_{i+1}[u_d] = (1/rho) * D_d p          (NCP -p_d, times 1/rho)
//                        and -sum_d lFhi[d] = sum_i a_i (L_{i+1} - source_{i+1}), so
//                        the volume update dt/W * lduh needs no lFhi array.
//   ac_corrector_kernel  riemannSolverLinear.cpph:23-167 for the six faces of a
//                        cell (each face solved by both of its cells, each
//                        keeping its own fluctuation) + faceIntegralLinear.cpph:8-101
//                        + the face part of the update.
// Layout in HBM: u[cell][node][4] (node = (z*B+y)*B+x), mat[cell][2] = (rho, c),
// qb[cell][face][B*B][4] (face 2d+o, face node index as the reference's
// lQhbnd: x-face (z,y), y-face (z,x), z-face (y,x)).
//
// Adjoint (this project's own; the reference has none): the transpose of the
// step above for fixed (rho, c) plus d/dc through lambda = rho c^2 in the
// predictor and the Riemann material average, and through the Rusanov speed
// max(c_L, c_R).  Kernels ac_fbar_kernel, ac_riemann_adj_kernel,
// ac_predict_adj_kernel.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace exa {
namespace {

constexpr int kAcThreads = 256;

template <int B>
struct AcShape {
  static constexpr int N = B * B * B;
  static constexpr int F = B * B;
  static constexpr int CPB = kAcThreads / N;  // cells per block
};

__device__ __forceinline__ long nb_cell(long cell, int n, int d, int dirsign) {
  long ix = cell % n, iy = (cell / n) % n, iz = cell / (long(n) * n);
  if (d == 0) ix = (ix + dirsign + n) % n;
  if (d == 1) iy = (iy + dirsign + n) % n;
  if (d == 2) iz = (iz + dirsign + n) % n;
  return (iz * n + iy) * n + ix;
}

// volume node on the line through face node (a, b) of a face normal to d, at
// coordinate l along d
template <int B>
__device__ __forceinline__ int face_line_node(int d, int a, int b, int l) {
  return d == 0 ? (a * B + b) * B + l : (d == 1 ? (a * B + l) * B + b : (l * B + a) * B + b);
}

template <int B>
__device__ __forceinline__ int line_node(int d, int x, int y, int z, int m) {
  return d == 0 ? (z * B + y) * B + m : (d == 1 ? (z * B + m) * B + x : (m * B + y) * B + x);
}

template <int B>
__device__ __forceinline__ void taylor(double dt, double (&a)[B]) {
  a[0] = 1.0;
  double f = 0.5 * dt;
  for (int i = 1; i < B; i++) {
    a[i] = f;
    f *= dt / (i + 2);
  }
}

// ---------------------------------------------------------------------------
// forward predictor (+ volume update when UPDATE)
// ---------------------------------------------------------------------------
template <int B, bool UPDATE>
__global__ void __launch_bounds__(kAcThreads) ac_predict_kernel(AcousticArgs a) {
  using S = AcShape<B>;
  constexpr int N = S::N, F = S::F, CPB = S::CPB;
  __shared__ double sD[B * B];  // dudx / h
  __shared__ double sFc[2][B];
  __shared__ double sL[CPB][4][N];
  const int t = threadIdx.x;
  const int lc = t / N, nd = t - lc * N;
  const long cell = long(blockIdx.x) * CPB + lc;
  const bool slot = lc < CPB;
  const bool on = slot && cell < a.ncells;
  if (t < B * B) sD[t] = a.tab->dudx[t] / a.h;
  if (t < B) {
    sFc[0][t] = a.tab->F0[t];
    sFc[1][t] = a.tab->F1[t];
  }
  const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
  double u0[4] = {0, 0, 0, 0};
  double rho = 1.0, c = 1.0;
  if (on) {
    const double4 q = reinterpret_cast<const double4*>(a.u)[cell * N + nd];
    u0[0] = q.x;
    u0[1] = q.y;
    u0[2] = q.z;
    u0[3] = q.w;
    rho = a.mat[2 * cell];
    c = a.mat[2 * cell + 1];
  }
  const double lam = rho * c * c, irho = 1.0 / rho;
  const int step = a.step_override >= 0 ? a.step_override : *a.step_ptr;
  const bool src = on && cell == a.src_cell;
  double g = 0.0;
  double st[B];
  for (int i = 0; i < B; i++) st[i] = 0.0;
  if (src) {
    g = a.src_g[nd];
    for (int i = 0; i < B; i++) st[i] = a.src_tab[long(step) * B + i];
  }
  double at[B];
  taylor<B>(a.dt, at);
  double q[4] = {u0[0] + st[0] * g, u0[1], u0[2], u0[3]};  // L_0
  double qh[4] = {q[0], q[1], q[2], q[3]};
  double acc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < B; i++) {
    __syncthreads();
    if (slot)
      for (int v = 0; v < 4; v++) sL[lc][v][nd] = q[v];
    __syncthreads();
    if (on) {
      double dp0 = 0, dp1 = 0, dp2 = 0, du = 0, dv = 0, dw = 0;
#pragma unroll
      for (int m = 0; m < B; m++) {
        const double cx = sD[x * B + m], cy = sD[y * B + m], cz = sD[z * B + m];
        const int nx = (z * B + y) * B + m, ny = (z * B + m) * B + x, nz = (m * B + y) * B + x;
        dp0 += cx * sL[lc][0][nx];
        dp1 += cy * sL[lc][0][ny];
        dp2 += cz * sL[lc][0][nz];
        du += cx * sL[lc][1][nx];
        dv += cy * sL[lc][2][ny];
        dw += cz * sL[lc][3][nz];
      }
      q[0] = lam * (du + dv + dw);
      q[1] = irho * dp0;
      q[2] = irho * dp1;
      q[3] = irho * dp2;
      for (int v = 0; v < 4; v++) acc[v] += at[i] * q[v];
      if (i + 1 < B) {
        q[0] += st[i + 1] * g;
        for (int v = 0; v < 4; v++) qh[v] += at[i + 1] * q[v];
      }
    }
  }
  // extrapolate the time-averaged variables to the six faces
  __syncthreads();
  if (slot)
    for (int v = 0; v < 4; v++) sL[lc][v][nd] = qh[v];
  __syncthreads();
  if (on) {
    for (int task = nd; task < 6 * F; task += N) {
      const int f = task / F, fn = task - f * F, d = f >> 1, o = f & 1;
      const int fa = fn / B, fb = fn - fa * B;
      double s[4] = {0, 0, 0, 0};
#pragma unroll
      for (int l = 0; l < B; l++) {
        const int node = face_line_node<B>(d, fa, fb, l);
        const double cf = sFc[o][l];
        for (int v = 0; v < 4; v++) s[v] += cf * sL[lc][v][node];
      }
      reinterpret_cast<double4*>(a.qb)[(cell * 6 + f) * F + fn] = make_double4(s[0], s[1], s[2], s[3]);
    }
    if (UPDATE) {
      double4 r;
      r.x = u0[0] + a.dt * acc[0] - st[0] * g;
      r.y = u0[1] + a.dt * acc[1];
      r.z = u0[2] + a.dt * acc[2];
      r.w = u0[3] + a.dt * acc[3];
      reinterpret_cast<double4*>(a.u)[cell * N + nd] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// forward predictor, x-line form: one thread per x-line (z, y) of a cell.
// x-derivatives come from the thread's registers; only the fields needed
// along y and z (p, v for y; p, w for z) go through LDS, double-buffered so
// each level needs one barrier.  Same mathematics as ac_predict_kernel; the
// derivative sums are scaled by 1/h after summation (rounding differs from
// the node form at the 1e-16 level).
// ---------------------------------------------------------------------------
template <int B>
struct AcLineShape {
  static constexpr int N = B * B * B;
  static constexpr int F = B * B;    // lines per cell
  static constexpr int BUF = 3 * N;  // p, v, w of one level
  static constexpr int CPB_T = kAcThreads / F;
  static constexpr int CPB_L = (60 * 1024) / (2 * BUF * 8);  // two level buffers per cell
  static constexpr int CPB = CPB_T < CPB_L ? CPB_T : CPB_L;    // cells per block
};

// Derivative pass of the x-line form: from the thread's x-line (P, U in
// registers) and the cell's p, v, w in LDS (sb), the divergence of u and the
// gradient of p at the B nodes of the line.
template <int B>
__device__ __forceinline__ void line_derivatives(const double* __restrict__ Dm, double invh, int z, int y,
                                                 const double (&P)[B], const double (&U)[B],
                                                 const double* sb, double (&div)[B], double (&px)[B],
                                                 double (&py)[B], double (&pz)[B]) {
  constexpr int N = B * B * B;
  double qv[B], qw[B];
  // x: from the line's own registers
#pragma unroll
  for (int x = 0; x < B; x++) {
    double sp = 0.0, su = 0.0;
#pragma unroll
    for (int m = 0; m < B; m++) {
      sp += Dm[x * B + m] * P[m];
      su += Dm[x * B + m] * U[m];
    }
    px[x] = sp;
    div[x] = su;
    py[x] = pz[x] = qv[x] = qw[x] = 0.0;
  }
  // y and z: one row of p, v (y) and p, w (z) per step m
#pragma unroll 1
  for (int m = 0; m < B; m++) {
    const double cy = Dm[y * B + m], cz = Dm[z * B + m];
    const double* ry = sb + (z * B + m) * B;  // (z, m, x)
    const double* rz = sb + (m * B + y) * B;  // (m, y, x)
#pragma unroll
    for (int x = 0; x < B; x++) {
      py[x] += cy * ry[x];
      qv[x] += cy * ry[N + x];
      pz[x] += cz * rz[x];
      qw[x] += cz * rz[2 * N + x];
    }
  }
#pragma unroll
  for (int x = 0; x < B; x++) {
    px[x] *= invh;
    py[x] *= invh;
    pz[x] *= invh;
    div[x] = div[x] * invh + qv[x] * invh + qw[x] * invh;
  }
}

// Levels 1..B-1 give the time average qh = sum_i a_i L_i; by linearity the
// volume update is dt * A(qh) - sigma0 g (A: one Cauchy-Kowalevski level,
// i.e. the reference's -sum_d lFhi[d] of the linear flux), so a last
// derivative pass on qh replaces per-level update accumulators.
// SPLIT: the optimised kernels' split Cauchy-Kowalevski predictor
// (KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_linear_split_ck_cpp.template):
// levels 1..B, the NCP term D_d p without 1/rho (the template's material
// matrix multiplies the flux part only, :201-331), the template's averaging
// factors f_0 = dt/2, f_{k+1} = f_k dt/(k+2) (:196, :346-353), source terms
// s_k of the per-step row {sigma0, s_0..s_{B-1}}; the volume update is
// dt * A(qh) - sigma0 g with that A (:474-640).
template <int B, bool UPDATE, bool SPLIT>
__global__ void __launch_bounds__(kAcThreads) ac_predict_lines_kernel(AcousticArgs a) {
  using S = AcLineShape<B>;
  constexpr int N = S::N, F = S::F, CPB = S::CPB;
  __shared__ double sbuf[2][CPB][S::BUF];
  const int t = threadIdx.x;
  const int lc = t / F, ln = t - lc * F;
  const int z = ln / B, y = ln - z * B;
  const long cell = long(blockIdx.x) * CPB + lc;
  const bool slot = lc < CPB;
  const int cs = slot ? lc : 0;
  const bool on = slot && cell < a.ncells;
  const double invh = 1.0 / a.h;
  const double* Dm = a.tab->dudx;  // dudx[i][m]
  double P[B], U[B], Vv[B], Wv[B];
  double rho = 1.0, c = 1.0;
  if (on) {
    const double4* src = reinterpret_cast<const double4*>(a.u) + cell * N + ln * B;
#pragma unroll
    for (int x = 0; x < B; x++) {
      const double4 q = src[x];
      P[x] = q.x;
      U[x] = q.y;
      Vv[x] = q.z;
      Wv[x] = q.w;
    }
    rho = a.mat[2 * cell];
    c = a.mat[2 * cell + 1];
  } else {
#pragma unroll
    for (int x = 0; x < B; x++) P[x] = U[x] = Vv[x] = Wv[x] = 0.0;
  }
  const double lam = rho * c * c, irho = SPLIT ? 1.0 : 1.0 / rho;
  const int step = a.step_override >= 0 ? a.step_override : *a.step_ptr;
  const bool src = on && cell == a.src_cell;
  const double* gp = src ? a.src_g + ln * B : nullptr;
  const double* stp = src ? a.src_tab + long(step) * (SPLIT ? B + 1 : B) : nullptr;
  if (src) {
#pragma unroll
    for (int x = 0; x < B; x++) P[x] += stp[0] * gp[x];  // level 0
  }
  double qh[4][B];
#pragma unroll
  for (int x = 0; x < B; x++) {
    qh[0][x] = P[x];
    qh[1][x] = U[x];
    qh[2][x] = Vv[x];
    qh[3][x] = Wv[x];
  }
  double an = 0.5 * a.dt;  // a_{i+1}
  constexpr int LEVELS = SPLIT ? B : B - 1;
#pragma unroll 1
  for (int i = 0; i < LEVELS; i++) {
    double* sb = sbuf[i & 1][cs];
    if (slot) {
#pragma unroll
      for (int x = 0; x < B; x++) {
        sb[ln * B + x] = P[x];
        sb[N + ln * B + x] = Vv[x];
        sb[2 * N + ln * B + x] = Wv[x];
      }
    }
    __syncthreads();
    double div[B], px[B], py[B], pz[B];
    line_derivatives<B>(Dm, invh, z, y, P, U, sb, div, px, py, pz);
    const double si = src ? stp[i + 1] : 0.0;  // s_{i+1} (generic row), s_i (split row)
#pragma unroll
    for (int x = 0; x < B; x++) {
      P[x] = lam * div[x] + (src ? si * gp[x] : 0.0);
      U[x] = irho * px[x];
      Vv[x] = irho * py[x];
      Wv[x] = irho * pz[x];
      qh[0][x] += an * P[x];
      qh[1][x] += an * U[x];
      qh[2][x] += an * Vv[x];
      qh[3][x] += an * Wv[x];
    }
    an *= a.dt / (SPLIT ? i + 2 : i + 3);
  }
  // A(qh): the volume update, and qh on the faces
  constexpr int LAST = LEVELS & 1;  // the buffer the last level did not use
  double* sb = sbuf[LAST][cs];
  if (slot) {
#pragma unroll
    for (int x = 0; x < B; x++) {
      sb[ln * B + x] = qh[0][x];
      sb[N + ln * B + x] = qh[2][x];
      sb[2 * N + ln * B + x] = qh[3][x];
    }
  }
  __syncthreads();  // also: every read of the other buffer (last level) is done
  double* su = sbuf[LAST ^ 1][cs];
  if (slot) {
#pragma unroll
    for (int x = 0; x < B; x++) su[ln * B + x] = qh[1][x];
  }
  if (UPDATE && on) {
    double div[B], px[B], py[B], pz[B];
    line_derivatives<B>(Dm, invh, z, y, qh[0], qh[1], sb, div, px, py, pz);
    const double s0 = src ? stp[0] : 0.0;
    double4* dst = reinterpret_cast<double4*>(a.u) + cell * N + ln * B;
#pragma unroll
    for (int x = 0; x < B; x++) {
      const double4 q = dst[x];
      dst[x] = make_double4(q.x + a.dt * (lam * div[x]) - (src ? s0 * gp[x] : 0.0), q.y + a.dt * (irho * px[x]),
                            q.z + a.dt * (irho * py[x]), q.w + a.dt * (irho * pz[x]));
    }
  }
  __syncthreads();  // u-component of qh in LDS
  if (on) {
    const double* F0 = a.tab->F0;
    const double* F1 = a.tab->F1;
    double4* qb = reinterpret_cast<double4*>(a.qb) + cell * 6 * F;
    // x-faces, node (z, y) = ln, from registers
    double l0[4] = {0, 0, 0, 0}, l1[4] = {0, 0, 0, 0};
#pragma unroll
    for (int x = 0; x < B; x++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        l0[v] += qh[v][x] * F0[x];
        l1[v] += qh[v][x] * F1[x];
      }
    qb[0 * F + ln] = make_double4(l0[0], l0[1], l0[2], l0[3]);
    qb[1 * F + ln] = make_double4(l1[0], l1[1], l1[2], l1[3]);
    // y-face node (z, x = y) and z-face node (y' = z, x' = y)
    double y0[4] = {0, 0, 0, 0}, y1[4] = {0, 0, 0, 0}, z0[4] = {0, 0, 0, 0}, z1[4] = {0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < B; l++) {
      const int ny = (z * B + l) * B + y;  // (z, l, x = y)
      const int nz = (l * B + z) * B + y;  // (l, y' = z, x' = y)
      const double f0 = F0[l], f1 = F1[l];
      const double vy[4] = {sb[ny], su[ny], sb[N + ny], sb[2 * N + ny]};
      const double vz[4] = {sb[nz], su[nz], sb[N + nz], sb[2 * N + nz]};
#pragma unroll
      for (int v = 0; v < 4; v++) {
        y0[v] += vy[v] * f0;
        y1[v] += vy[v] * f1;
        z0[v] += vz[v] * f0;
        z1[v] += vz[v] * f1;
      }
    }
    qb[2 * F + z * B + y] = make_double4(y0[0], y0[1], y0[2], y0[3]);
    qb[3 * F + z * B + y] = make_double4(y1[0], y1[1], y1[2], y1[3]);
    qb[4 * F + z * B + y] = make_double4(z0[0], z0[1], z0[2], z0[3]);
    qb[5 * F + z * B + y] = make_double4(z1[0], z1[1], z1[2], z1[3]);
  }
}

// ---------------------------------------------------------------------------
// forward corrector: Riemann fluctuations of the six faces + face update
// ---------------------------------------------------------------------------
template <int B>
__global__ void __launch_bounds__(kAcThreads) ac_corrector_kernel(AcousticArgs a) {
  using S = AcShape<B>;
  constexpr int N = S::N, F = S::F, CPB = S::CPB;
  __shared__ double sCf[2][B];  // dt/h * F_o[l] / w_l
  __shared__ double sFl[CPB][6][F][4];
  const int t = threadIdx.x;
  const int lc = t / N, nd = t - lc * N;
  const long cell = long(blockIdx.x) * CPB + lc;
  const bool on = lc < CPB && cell < a.ncells;
  if (t < B) {
    sCf[0][t] = a.dt / a.h * a.tab->F0[t] / a.tab->w[t];
    sCf[1][t] = a.dt / a.h * a.tab->F1[t] / a.tab->w[t];
  }
  if (on) {
    const double rho_o = a.mat[2 * cell], c_o = a.mat[2 * cell + 1];
    for (int task = nd; task < 6 * F; task += N) {
      const int f = task / F, fn = task - f * F, d = f >> 1, o = f & 1;
      const long nb = nb_cell(cell, a.n, d, o ? 1 : -1);
      const double rho_n = a.mat[2 * nb], c_n = a.mat[2 * nb + 1];
      const double4 qo = reinterpret_cast<const double4*>(a.qb)[(cell * 6 + f) * F + fn];
      const double4 qn = reinterpret_cast<const double4*>(a.qb)[(nb * 6 + (f ^ 1)) * F + fn];
      // own face o = 1: own cell is the left state; o = 0: own cell is the right state
      const double4 QL = o ? qo : qn, QR = o ? qn : qo;
      const double cL = o ? c_o : c_n, cR = o ? c_n : c_o;
      const double aL = fabs(cL), aR = fabs(cR);
      const double smax = aL > aR ? aL : aR;
      const double rhoa = 0.5 * ((o ? rho_n : rho_o) + (o ? rho_o : rho_n));
      const double ca = 0.5 * (cR + cL);
      const double lama = rhoa * ca * ca, irhoa = 1.0 / rhoa;
      double qd[4] = {0.5 * (QR.x - QL.x), 0.5 * (QR.y - QL.y), 0.5 * (QR.z - QL.z), 0.5 * (QR.w - QL.w)};
      double accv[4] = {0, 0, 0, 0};
      accv[1 + d] += irhoa * (-qd[0]);  // non-conservative product, times 1/rho
      accv[0] += lama * (-qd[1 + d]);   // flux, times lambda
      const double sgn = o ? -1.0 : 1.0;  // FL = acc - smax qd, FR = acc + smax qd
      for (int v = 0; v < 4; v++) sFl[lc][f][fn][v] = accv[v] + sgn * smax * qd[v];
    }
  }
  __syncthreads();
  if (on) {
    const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
    const int l3[3] = {x, y, z};
    const int fi[3] = {z * B + y, z * B + x, y * B + x};
    double du[4] = {0, 0, 0, 0};
    for (int d = 0; d < 3; d++)
      for (int o = 0; o < 2; o++) {
        const double cf = sCf[o][l3[d]];
        for (int v = 0; v < 4; v++) du[v] -= cf * sFl[lc][2 * d + o][fi[d]][v];
      }
    double4* up = reinterpret_cast<double4*>(a.u) + cell * N + nd;
    double4 r = *up;
    r.x += du[0];
    r.y += du[1];
    r.z += du[2];
    r.w += du[3];
    *up = r;
  }
}

// ---------------------------------------------------------------------------
// receivers: rec[step][r] = sum_node p * weight (one wave per receiver)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) ac_receivers_kernel(AcousticArgs a, int N) {
  const int r = blockIdx.x, lane = threadIdx.x;
  const long cell = a.rcell[r];
  double s = 0.0;
  for (int nd = lane; nd < N; nd += 64) s += a.u[(cell * N + nd) * 4] * a.rw[long(r) * N + nd];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  const int step = a.step_override >= 0 ? a.step_override : *a.step_ptr;
  if (lane == 0) a.rec[long(step) * a.nrec + r] = s;
}

// ubar[cell_r][node][0] += res[r] * weight (receivers in order; one block)
__global__ void __launch_bounds__(256) ac_receivers_adj_kernel(AcousticArgs a, int N, const double* res) {
  for (int r = 0; r < a.nrec; r++) {
    const long cell = a.rcell[r];
    const double rv = res[r];
    for (int nd = threadIdx.x; nd < N; nd += 256) a.ubar[(cell * N + nd) * 4] += rv * a.rw[long(r) * N + nd];
  }
}

__global__ void ac_advance_kernel(int* step) {
  if (threadIdx.x == 0) step[0] = step[0] + 1;
}

// ---------------------------------------------------------------------------
// adjoint 1: cotangent of the fluctuations, fbar[cell][f][fn][v] =
//   -dt/h * sum_l F_o[l]/w_l * ubar[cell][line node l][v]
// ---------------------------------------------------------------------------
template <int B>
__global__ void __launch_bounds__(kAcThreads) ac_fbar_kernel(AcousticArgs a) {
  using S = AcShape<B>;
  constexpr int N = S::N, F = S::F, CPB = S::CPB;
  __shared__ double sCf[2][B];
  __shared__ double sU[CPB][4][N];
  const int t = threadIdx.x;
  const int lc = t / N, nd = t - lc * N;
  const long cell = long(blockIdx.x) * CPB + lc;
  const bool on = lc < CPB && cell < a.ncells;
  if (t < B) {
    sCf[0][t] = a.dt / a.h * a.tab->F0[t] / a.tab->w[t];
    sCf[1][t] = a.dt / a.h * a.tab->F1[t] / a.tab->w[t];
  }
  if (on) {
    const double4 q = reinterpret_cast<const double4*>(a.ubar)[cell * N + nd];
    sU[lc][0][nd] = q.x;
    sU[lc][1][nd] = q.y;
    sU[lc][2][nd] = q.z;
    sU[lc][3][nd] = q.w;
  }
  __syncthreads();
  if (on) {
    for (int task = nd; task < 6 * F; task += N) {
      const int f = task / F, fn = task - f * F, d = f >> 1, o = f & 1;
      const int fa = fn / B, fb = fn - fa * B;
      double s[4] = {0, 0, 0, 0};
#pragma unroll
      for (int l = 0; l < B; l++) {
        const int node = face_line_node<B>(d, fa, fb, l);
        const double cf = sCf[o][l];
        for (int v = 0; v < 4; v++) s[v] -= cf * sU[lc][v][node];
      }
      reinterpret_cast<double4*>(a.fbar)[(cell * 6 + f) * F + fn] = make_double4(s[0], s[1], s[2], s[3]);
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint 2: Riemann transpose.  For every face of the cell: cotangent of the
// cell's extrapolated variables (qbbar) and the face's part of dJ/dc_cell.
// ---------------------------------------------------------------------------
template <int B>
__global__ void __launch_bounds__(kAcThreads) ac_riemann_adj_kernel(AcousticArgs a) {
  using S = AcShape<B>;
  constexpr int N = S::N, F = S::F, CPB = S::CPB;
  __shared__ double sRed[kAcThreads];
  const int t = threadIdx.x;
  const int lc = t / N, nd = t - lc * N;
  const long cell = long(blockIdx.x) * CPB + lc;
  const bool on = lc < CPB && cell < a.ncells;
  double gpart = 0.0;
  if (on) {
    const double rho_o = a.mat[2 * cell], c_o = a.mat[2 * cell + 1];
    for (int task = nd; task < 6 * F; task += N) {
      const int f = task / F, fn = task - f * F, d = f >> 1, o = f & 1;
      const long nb = nb_cell(cell, a.n, d, o ? 1 : -1);
      const double rho_n = a.mat[2 * nb], c_n = a.mat[2 * nb + 1];
      const long io = (cell * 6 + f) * F + fn, in = (nb * 6 + (f ^ 1)) * F + fn;
      const double4 qo = reinterpret_cast<const double4*>(a.qb)[io];
      const double4 qn = reinterpret_cast<const double4*>(a.qb)[in];
      const double4 bo = reinterpret_cast<const double4*>(a.fbar)[io];
      const double4 bn = reinterpret_cast<const double4*>(a.fbar)[in];
      const double4 QL = o ? qo : qn, QR = o ? qn : qo;
      const double4 FLb = o ? bo : bn, FRb = o ? bn : bo;
      const double cL = o ? c_o : c_n, cR = o ? c_n : c_o;
      const double aL = fabs(cL), aR = fabs(cR);
      const double smax = aL > aR ? aL : aR;
      const double rhoa = 0.5 * ((o ? rho_n : rho_o) + (o ? rho_o : rho_n));
      const double ca = 0.5 * (cR + cL);
      const double lama = rhoa * ca * ca, irhoa = 1.0 / rhoa;
      const double qd[4] = {0.5 * (QR.x - QL.x), 0.5 * (QR.y - QL.y), 0.5 * (QR.z - QL.z), 0.5 * (QR.w - QL.w)};
      const double sb[4] = {FLb.x + FRb.x, FLb.y + FRb.y, FLb.z + FRb.z, FLb.w + FRb.w};
      const double db[4] = {FRb.x - FLb.x, FRb.y - FLb.y, FRb.z - FLb.z, FRb.w - FLb.w};
      double qdb[4];
      double gs = 0.0;
      for (int v = 0; v < 4; v++) {
        qdb[v] = smax * db[v];
        gs += db[v] * qd[v];
      }
      qdb[1 + d] += -lama * sb[0];
      qdb[0] += -irhoa * sb[1 + d];
      const double gl = sb[0] * (-qd[1 + d]);
      // own share of d/dc: lambda_avg = rho_avg c_avg^2 with dc_avg/dc_own = 1/2;
      // smax = max(c_L, c_R) goes to the larger one (ties to the left cell)
      const bool own_max = o ? (aL >= aR) : (aR > aL);
      gpart += gl * rhoa * ca + (own_max ? gs : 0.0);
      const double sg = o ? -0.5 : 0.5;  // QL -= qdb/2, QR += qdb/2
      reinterpret_cast<double4*>(a.qbbar)[io] =
          make_double4(sg * qdb[0], sg * qdb[1], sg * qdb[2], sg * qdb[3]);
    }
  }
  // per-cell sum of gpart over its N threads
  sRed[t] = gpart;
  __syncthreads();
  if (on && nd == 0) {
    double s = 0.0;
    for (int k = 0; k < N; k++) s += sRed[lc * N + k];
    a.gc[cell] += s;
  }
}

// ---------------------------------------------------------------------------
// adjoint 3: predictor + volume update transpose (recomputes the forward
// levels of the cell).  ubar (cotangent of u^{k+1}) is replaced by the
// cotangent of u^k; gc[cell] += dJ/dc from the material matrix.
// ---------------------------------------------------------------------------
template <int B>
__global__ void __launch_bounds__(kAcThreads) ac_predict_adj_kernel(AcousticArgs a) {
  using S = AcShape<B>;
  constexpr int N = S::N, F = S::F, CPB = S::CPB;
  __shared__ double sD[B * B];
  __shared__ double sFc[2][B];
  __shared__ double sL[CPB][4][N];
  __shared__ double sQb[CPB][6][F][4];
  __shared__ double sRed[kAcThreads];
  const int t = threadIdx.x;
  const int lc = t / N, nd = t - lc * N;
  const long cell = long(blockIdx.x) * CPB + lc;
  const bool slot = lc < CPB;
  const bool on = slot && cell < a.ncells;
  if (t < B * B) sD[t] = a.tab->dudx[t] / a.h;
  if (t < B) {
    sFc[0][t] = a.tab->F0[t];
    sFc[1][t] = a.tab->F1[t];
  }
  const int x = nd % B, y = (nd / B) % B, z = nd / (B * B);
  double u0[4] = {0, 0, 0, 0}, ub[4] = {0, 0, 0, 0};
  double rho = 1.0, c = 1.0;
  if (on) {
    const double4 q = reinterpret_cast<const double4*>(a.u)[cell * N + nd];
    u0[0] = q.x;
    u0[1] = q.y;
    u0[2] = q.z;
    u0[3] = q.w;
    const double4 b = reinterpret_cast<const double4*>(a.ubar)[cell * N + nd];
    ub[0] = b.x;
    ub[1] = b.y;
    ub[2] = b.z;
    ub[3] = b.w;
    rho = a.mat[2 * cell];
    c = a.mat[2 * cell + 1];
    for (int task = nd; task < 6 * F; task += N) {
      const double4 qb = reinterpret_cast<const double4*>(a.qbbar)[cell * 6 * F + task];
      double* dst = &sQb[lc][0][0][0] + 4 * task;
      dst[0] = qb.x;
      dst[1] = qb.y;
      dst[2] = qb.z;
      dst[3] = qb.w;
    }
  }
  const double lam = rho * c * c, irho = 1.0 / rho;
  const int step = a.step_override >= 0 ? a.step_override : *a.step_ptr;
  const bool src = on && cell == a.src_cell;
  double g = 0.0;
  double st[B];
  for (int i = 0; i < B; i++) st[i] = 0.0;
  if (src) {
    g = a.src_g[nd];
    for (int i = 0; i < B; i++) st[i] = a.src_tab[long(step) * B + i];
  }
  double at[B];
  taylor<B>(a.dt, at);
  // forward levels: keep div(u) of every level
  double dv_[B];
  double q[4] = {u0[0] + st[0] * g, u0[1], u0[2], u0[3]};
#pragma unroll
  for (int i = 0; i < B; i++) {
    __syncthreads();
    if (slot)
      for (int v = 0; v < 4; v++) sL[lc][v][nd] = q[v];
    __syncthreads();
    double dp0 = 0, dp1 = 0, dp2 = 0, du = 0, dv = 0, dw = 0;
    if (on) {
#pragma unroll
      for (int m = 0; m < B; m++) {
        const double cx = sD[x * B + m], cy = sD[y * B + m], cz = sD[z * B + m];
        const int nx = (z * B + y) * B + m, ny = (z * B + m) * B + x, nz = (m * B + y) * B + x;
        dp0 += cx * sL[lc][0][nx];
        dp1 += cy * sL[lc][0][ny];
        dp2 += cz * sL[lc][0][nz];
        du += cx * sL[lc][1][nx];
        dv += cy * sL[lc][2][ny];
        dw += cz * sL[lc][3][nz];
      }
    }
    dv_[i] = du + dv + dw;
    q[0] = lam * dv_[i] + (i + 1 < B ? st[i + 1] * g : 0.0);
    q[1] = irho * dp0;
    q[2] = irho * dp1;
    q[3] = irho * dp2;
  }
  // cotangent of the time-averaged variables: E^T qbbar
  double qhb[4] = {0, 0, 0, 0};
  if (on) {
    const int l3[3] = {x, y, z};
    const int fi[3] = {z * B + y, z * B + x, y * B + x};
    for (int d = 0; d < 3; d++)
      for (int v = 0; v < 4; v++)
        qhb[v] += sFc[0][l3[d]] * sQb[lc][2 * d][fi[d]][v] + sFc[1][l3[d]] * sQb[lc][2 * d + 1][fi[d]][v];
  }
  // reverse sweep over the levels
  double lb[4] = {0, 0, 0, 0};  // cotangent of L_{i+1}
  double glam = 0.0;
#pragma unroll
  for (int i = B - 1; i >= 0; i--) {
    const double tv[4] = {-(lb[0] + a.dt * at[i] * ub[0]), -(lb[1] + a.dt * at[i] * ub[1]),
                          -(lb[2] + a.dt * at[i] * ub[2]), -(lb[3] + a.dt * at[i] * ub[3])};
    glam += tv[0] * (-dv_[i]);
    __syncthreads();
    if (slot) {
      sL[lc][0][nd] = -lam * tv[0];
      sL[lc][1][nd] = -irho * tv[1];
      sL[lc][2][nd] = -irho * tv[2];
      sL[lc][3][nd] = -irho * tv[3];
    }
    __syncthreads();
    double n0 = 0, n1 = 0, n2 = 0, n3 = 0;
    if (on) {
#pragma unroll
      for (int m = 0; m < B; m++) {
        const double cx = sD[m * B + x], cy = sD[m * B + y], cz = sD[m * B + z];
        const int nx = (z * B + y) * B + m, ny = (z * B + m) * B + x, nz = (m * B + y) * B + x;
        n1 += cx * sL[lc][0][nx];
        n2 += cy * sL[lc][0][ny];
        n3 += cz * sL[lc][0][nz];
        n0 += cx * sL[lc][1][nx] + cy * sL[lc][2][ny] + cz * sL[lc][3][nz];
      }
    }
    lb[0] = at[i] * qhb[0] + n0;
    lb[1] = at[i] * qhb[1] + n1;
    lb[2] = at[i] * qhb[2] + n2;
    lb[3] = at[i] * qhb[3] + n3;
  }
  if (on) {
    reinterpret_cast<double4*>(a.ubar)[cell * N + nd] =
        make_double4(ub[0] + lb[0], ub[1] + lb[1], ub[2] + lb[2], ub[3] + lb[3]);
  }
  sRed[t] = on ? glam : 0.0;
  __syncthreads();
  if (on && nd == 0) {
    double s = 0.0;
    for (int k = 0; k < N; k++) s += sRed[lc * N + k];
    a.gc[cell] += s * 2.0 * rho * c;
  }
}

template <int B>
int launch_all(int which, const AcousticArgs& a, hipStream_t s) {
  using S = AcShape<B>;
  const unsigned blocks = static_cast<unsigned>((a.ncells + S::CPB - 1) / S::CPB);
  const unsigned lblocks = static_cast<unsigned>((a.ncells + AcLineShape<B>::CPB - 1) / AcLineShape<B>::CPB);
  switch (which) {
    case kAcPredictUpdate:
      if (a.split_ck && a.pred_variant != 0) return -1;  // the split predictor has the x-line form only
      if (a.split_ck)
        ac_predict_lines_kernel<B, true, true><<<lblocks, kAcThreads, 0, s>>>(a);
      else if (a.pred_variant == 0)
        ac_predict_lines_kernel<B, true, false><<<lblocks, kAcThreads, 0, s>>>(a);
      else
        ac_predict_kernel<B, true><<<blocks, kAcThreads, 0, s>>>(a);
      break;
    case kAcPredictFaces:
      if (a.split_ck) return -1;  // the adjoint sweeps transpose the generic predictor only
      if (a.pred_variant == 0)
        ac_predict_lines_kernel<B, false, false><<<lblocks, kAcThreads, 0, s>>>(a);
      else
        ac_predict_kernel<B, false><<<blocks, kAcThreads, 0, s>>>(a);
      break;
    case kAcCorrector:
      ac_corrector_kernel<B><<<blocks, kAcThreads, 0, s>>>(a);
      break;
    case kAcFbar:
      ac_fbar_kernel<B><<<blocks, kAcThreads, 0, s>>>(a);
      break;
    case kAcRiemannAdj:
      ac_riemann_adj_kernel<B><<<blocks, kAcThreads, 0, s>>>(a);
      break;
    case kAcPredictAdj:
      ac_predict_adj_kernel<B><<<blocks, kAcThreads, 0, s>>>(a);
      break;
    case kAcReceivers:
      if (a.nrec > 0) ac_receivers_kernel<<<a.nrec, 64, 0, s>>>(a, S::N);
      break;
    default:
      return -1;
  }
  return static_cast<int>(hipGetLastError());
}

}  // namespace

bool acoustic_order_supported(int order) { return order >= 1 && order <= 5; }

int launch_acoustic(int which, int order, const AcousticArgs& a, hipStream_t s) {
  if (a.ncells <= 0 && which != kAcReceivers) return 0;
  switch (order) {
    case 1: return launch_all<2>(which, a, s);
    case 2: return launch_all<3>(which, a, s);
    case 3: return launch_all<4>(which, a, s);
    case 4: return launch_all<5>(which, a, s);
    case 5: return launch_all<6>(which, a, s);
    default: return -1;
  }
}

int launch_acoustic_receivers_adj(int order, const AcousticArgs& a, const double* res, hipStream_t s) {
  if (a.nrec <= 0) return 0;
  const int B = order + 1;
  ac_receivers_adj_kernel<<<1, 256, 0, s>>>(a, B * B * B, res);
  return static_cast<int>(hipGetLastError());
}

int launch_acoustic_advance(int* step, hipStream_t s) {
  ac_advance_kernel<<<1, 64, 0, s>>>(step);
  return static_cast<int>(hipGetLastError());
}

}  // namespace exa

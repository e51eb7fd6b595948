/*
 * Linear ADER-DG oracle -- TEST INFRASTRUCTURE ONLY, like aderdg_oracle.c:
 * used by tests/ and bench.py's cpu_baseline, never by the product library.
 *
 * Forward kernels restate the reference's generic linear kernels
 * (gc3 = ExaHyPE/kernels/aderdg/generic/c/3d; the 2D files in c/2d have the
 * same structure with one direction fewer), for dimension 2 or 3:
 *   orc_lin_stp        spaceTimePredictorLinear.cpph:30-647 (Cauchy-Kowalevski with
 *                      flux / NCP / material matrix switches, point sources,
 *                      Taylor time averages, extrapolation)
 *   orc_lin_volume     volumeIntegralLinear.cpph:21-70
 *   orc_lin_riemann    riemannSolverLinear.cpph:23-167
 *   orc_lin_face       faceIntegralLinear.cpph:8-101
 *   orc_lin_surface    surfaceIntegralLinear.cpph:26-119 (all faces of a cell)
 *   orc_lin_psi        deltaDistribution.cpph:37-146 (point-source moments, 3D)
 * PDE hooks:
 *   ORC_PDE_ACOUSTIC   ApplicationExamples/Linear/LinearWave3D/MyLinearWaveSolver.cpp
 *                      (variables p,u,v,w; parameters rho,c): eigenvalues :108-118,
 *                      flux :121-143, nonConservativeProduct :147-182,
 *                      multiplyMaterialParameterMatrix :188-209
 *   ORC_PDE_ELASTIC2D  the 2D elasticity test PDE of
 *                      ExaHyPE/exahype/tests/kernels/c/ElasticityKernelTest2D.cpp:32-205
 *                      (9 variables; parameters lambda, mu, rho) whose fixtures
 *                      (tests/kernels/testdata/elasticity_testdata.cpp) pin the
 *                      generic linear kernels.
 *
 * The acoustic mesh driver (periodic n^3 grid, fixed dt, Ricker point source,
 * receivers) and the discrete adjoint are this project's own: the reference
 * has no adjoint.  The adjoint is the exact transpose of the forward step
 * (linear in the state for fixed parameters) plus the derivative with respect
 * to each cell's wave speed c through the material matrix (lambda = rho c^2)
 * and the Rusanov speed max(c_L, c_R); dt is held fixed.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aderdg_oracle.h"

#define MAXLV 16
#define LIN_FLUX 1
#define LIN_NCP 2
#define LIN_MM 4

/* ---- PDE hooks -------------------------------------------------------------- */
static void lin_flux(int pde, const double* Q, double F[3][MAXLV]) {
  if (pde == ORC_PDE_ACOUSTIC) { /* MyLinearWaveSolver.cpp:121-143 */
    for (int d = 0; d < 3; d++) {
      F[d][0] = -Q[1 + d];
      F[d][1] = F[d][2] = F[d][3] = 0.0;
    }
  } else { /* ElasticityKernelTest2D.cpp:32-92 */
    double* f = F[0];
    double* g = F[1];
    for (int i = 0; i < 9; i++) f[i] = g[i] = 0.0;
    const double lam = Q[9], mu = Q[10], irho = 1.0 / Q[11];
    f[0] = -(lam + 2 * mu) * Q[6];
    f[1] = -lam * Q[6];
    f[2] = -lam * Q[6];
    f[3] = -mu * Q[7];
    f[5] = -mu * Q[8];
    f[6] = -irho * Q[0];
    f[7] = -irho * Q[3];
    f[8] = -irho * Q[5];
    g[0] = -lam * Q[7];
    g[1] = -(lam + 2 * mu) * Q[7];
    g[2] = -lam * Q[7];
    g[3] = -mu * Q[6];
    g[4] = -mu * Q[8];
    g[6] = -irho * Q[3];
    g[7] = -irho * Q[1];
    g[8] = -irho * Q[4];
  }
}

/* B(Q) gradQ per direction */
static void lin_ncp(int pde, const double* Q, const double G[3][MAXLV], double R[3][MAXLV]) {
  if (pde == ORC_PDE_ACOUSTIC) { /* MyLinearWaveSolver.cpp:147-182 */
    for (int d = 0; d < 3; d++) {
      R[d][0] = R[d][1] = R[d][2] = R[d][3] = 0.0;
      R[d][1 + d] = -G[d][0];
    }
  } else { /* ElasticityKernelTest2D.cpp:131-181 */
    const double lam = Q[9], mu = Q[10], irho = 1.0 / Q[11];
    const double* gx = G[0];
    const double* gy = G[1];
    double* bx = R[0];
    double* by = R[1];
    for (int i = 0; i < 9; i++) bx[i] = by[i] = 0.0;
    bx[0] = -(lam + 2 * mu) * gx[6];
    bx[1] = -lam * gx[6];
    bx[2] = -lam * gx[6];
    bx[3] = -mu * gx[7];
    bx[5] = -mu * gx[8];
    bx[6] = -irho * gx[0];
    bx[7] = -irho * gx[3];
    bx[8] = -irho * gx[5];
    by[0] = -lam * gy[7];
    by[1] = -(lam + 2 * mu) * gy[7];
    by[2] = -lam * gy[7];
    by[3] = -mu * gy[6];
    by[4] = -mu * gy[8];
    by[6] = -irho * gy[3];
    by[7] = -irho * gy[1];
    by[8] = -irho * gy[4];
  }
}

/* rhs <- M(Q) rhs (MyLinearWaveSolver.cpp:188-209; the elasticity test's is empty) */
static void lin_mm(int pde, const double* Q, double R[3][MAXLV]) {
  if (pde != ORC_PDE_ACOUSTIC) return;
  const double cp = Q[5], rho = Q[4];
  const double lam = rho * cp * cp;
  for (int d = 0; d < 3; d++) {
    R[d][0] = lam * R[d][0];
    R[d][1] = 1.0 / rho * R[d][1];
    R[d][2] = 1.0 / rho * R[d][2];
    R[d][3] = 1.0 / rho * R[d][3];
  }
}

/* max |eigenvalue| (acoustic: (c,-c,0,0); elasticity: +-cp, +-cs, 0) */
static double lin_smax(int pde, const double* Q) {
  if (pde == ORC_PDE_ACOUSTIC) return fabs(Q[5]);
  return sqrt((Q[9] + 2 * Q[10]) / Q[11]);
}

/* ---- index helpers (nodes x fastest) ---------------------------------------- */
static int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}
static int coord(int n, int d, int B) { return (n / ipow(B, d)) % B; }
static int with_coord(int n, int d, int m, int B) { return n + (m - coord(n, d, B)) * ipow(B, d); }
/* index on the face normal to d of volume node n (remaining coordinates, lower faster) */
static int face_index(int n, int d, int B, int D) {
  int f = 0, s = 1;
  for (int e = 0; e < D; e++)
    if (e != d) {
      f += coord(n, e, B) * s;
      s *= B;
    }
  return f;
}

/* Lagrange basis phi_k(s) on the nodes x[0..B) */
static double lagr(const double* x, int B, int k, double s) {
  double r = 1.0;
  for (int m = 0; m < B; m++)
    if (m != k) r *= (s - x[m]) / (x[k] - x[m]);
  return r;
}

/* Ricker wavelet with peak frequency f0, delayed by t0 */
double orc_ricker(double t, double f0, double t0) {
  const double pi = 3.14159265358979323846;
  const double a = pi * pi * f0 * f0 * (t - t0) * (t - t0);
  return (1.0 - 2.0 * a) * exp(-a);
}

/* deltaDistribution (deltaDistribution.cpph:74-146) for one pressure source
 * of strength amp*ricker(t; f0, t0) at xs in the cube cell [centre -/+ h/2]^3:
 * PSi[n][B^3][V] for n = 0 (time t) and n = 1..B (t + dt*node[n-1]). */
void orc_lin_psi(const orc_solver* s, const double* centre, double h, double t, double dt, const double* xs,
                 double amp, double f0, double t0, double* PSi) {
  const int B = s->b.order + 1, B3 = B * B * B, V = s->nvar;
  const double* w = s->b.w;
  const double* x = s->b.x;
  memset(PSi, 0, sizeof(double) * (B + 1) * B3 * V);
  double xr[3];
  for (int d = 0; d < 3; d++) xr[d] = (xs[d] - centre[d]) / h + 0.5;
  for (int n = 0; n < B + 1; n++) {
    const double tn = (n == 0) ? t : dt * x[n - 1] + t;
    const double f = amp * orc_ricker(tn, f0, t0);
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double basis = lagr(x, B, k, xr[0]) * lagr(x, B, j, xr[1]) * lagr(x, B, i, xr[2]);
          const double winv = 1.0 / (w[k] * h * w[j] * h * w[i] * h);
          PSi[((size_t)(n * B3 + (i * B + j) * B + k)) * V + 0] += winv * basis * f;
        }
  }
}

/* Point-source moments of the split predictor (fusedSPTVI_linear_split_ck_cpp.template:112-155):
 * PSi[t][B^3][V] at the Gauss time nodes t + dt node[t], t = 0..B-1, with the
 * spatial factor of orc_lin_psi. */
void orc_lin_psi_split(const orc_solver* s, const double* centre, double h, double t, double dt, const double* xs,
                       double amp, double f0, double t0, double* PSi) {
  const int B = s->b.order + 1, B3 = B * B * B, V = s->nvar;
  double* full = (double*)malloc(sizeof(double) * (B + 1) * B3 * V);
  /* slots 1..B of the generic moments are the Gauss time nodes */
  orc_lin_psi(s, centre, h, t, dt, xs, amp, f0, t0, full);
  memcpy(PSi, full + (size_t)B3 * V, sizeof(double) * B * B3 * V);
  free(full);
}

/* Taylor averaging factors a_0 = 1, a_i = dt^i/(i+1)! (spaceTimePredictorLinear.cpph:409-421) */
static void taylor_factors(int B, double dt, double* a) {
  a[0] = 1.0;
  double f = 0.5 * dt;
  for (int i = 1; i < B; i++) {
    a[i] = f;
    f *= dt / (i + 2);
  }
}

/* spaceTimePredictorLinear for one cell (dimension s->dim).
 *   luh[B^D][nData], dx[D], PSi (3D only, may be NULL)
 *   out lQhi[B^D][nData], lFhi[(D+1)][B^D][V] (slot D, the algebraic source, stays 0),
 *       lQbnd[2D][B^(D-1)][nData], lFbnd[2D][B^(D-1)][V]; any output may be NULL.
 *   levels (optional) receives lQi[(B+1)][B^D][nData]. */
void orc_lin_stp(const orc_solver* s, int flags, const double* luh, const double* dx, double dt, const double* PSi,
                 double* lQhi_out, double* lFhi_out, double* lQbnd, double* lFbnd, double* levels) {
  const int D = s->dim, B = s->b.order + 1, V = s->nvar, ND = s->nvar + s->npar;
  const int BD = ipow(B, D), BF = ipow(B, D - 1);
  const double* w = s->b.w;
  const double* dudx = s->b.dudx;
  const double* F0 = s->b.F0;
  const double* F1 = s->b.F1;
  const double invDt = 1.0 / dt;
  double* lQi = (double*)calloc((size_t)(B + 1) * BD * ND, sizeof(double));
  double* lFi = (double*)calloc((size_t)B * BD * D * V, sizeof(double));
  double* lQhi = (double*)calloc((size_t)BD * ND, sizeof(double));
  double* lFhi = (double*)calloc((size_t)(D + 1) * BD * V, sizeof(double));
  double* Fn = (double*)malloc(sizeof(double) * BD * 3 * MAXLV);
#define QI(i, n, v) lQi[((size_t)(i) * BD + (n)) * ND + (v)]
#define FI(i, n, d, v) lFi[(((size_t)(i) * BD + (n)) * D + (d)) * V + (v)]
  memcpy(lQi, luh, sizeof(double) * BD * ND);
  if (PSi && D == 3) {
    /* :111-182 */
    for (int n = 0; n < BD; n++)
      for (int m = 0; m < V; m++) {
        double sum = 0.0;
        for (int t = 1; t < B + 1; t++) sum = sum + w[t - 1] * PSi[((size_t)t * BD + n) * V + m];
        QI(0, n, m) += dt * sum;
      }
    double* PSd = (double*)malloc(sizeof(double) * B * BD * V);
    double* tmp = (double*)malloc(sizeof(double) * B * BD * V);
    memcpy(PSd, PSi + (size_t)BD * V, sizeof(double) * B * BD * V);
    for (int n = 0; n < BD; n++)
      for (int m = 0; m < V; m++) QI(1, n, m) += PSi[(size_t)n * V + m];
    for (int i = 2; i < B; i++) {
      memset(tmp, 0, sizeof(double) * B * BD * V);
      for (int t = 0; t < B; t++)
        for (int n = 0; n < BD; n++)
          for (int m = 0; m < V; m++)
            for (int q = 0; q < B; q++)
              tmp[((size_t)t * BD + n) * V + m] += PSd[((size_t)q * BD + n) * V + m] * dudx[t * B + q] * invDt;
      for (int n = 0; n < BD; n++)
        for (int m = 0; m < V; m++)
          for (int t = 0; t < B; t++) QI(i, n, m) += tmp[((size_t)t * BD + n) * V + m] * F0[t];
      memcpy(PSd, tmp, sizeof(double) * B * BD * V);
    }
    free(PSd);
    free(tmp);
  }
  /* Cauchy-Kowalevski levels (:185-392) */
  for (int i = 0; i < B; i++) {
    if (flags & LIN_FLUX) /* fluxes at every node of level i (:187-199) */
      for (int n = 0; n < BD; n++) {
        double F[3][MAXLV];
        lin_flux(s->pde, &QI(i, n, 0), F);
        for (int d = 0; d < D; d++)
          for (int v = 0; v < V; v++) Fn[((size_t)n * 3 + d) * MAXLV + v] = F[d][v];
      }
    for (int n = 0; n < BD; n++) {
      double F[3][MAXLV], G[3][MAXLV], R[3][MAXLV];
      memset(F, 0, sizeof(F));
      memset(G, 0, sizeof(G));
      for (int d = 0; d < D; d++) {
        const int c = coord(n, d, B);
        for (int m = 0; m < B; m++) {
          const int nn = with_coord(n, d, m, B);
          const double dd = dudx[c * B + m] / dx[d];
          for (int v = 0; v < V; v++) {
            if (flags & LIN_FLUX) F[d][v] += Fn[((size_t)nn * 3 + d) * MAXLV + v] * dd; /* :201-262 */
            G[d][v] += QI(i, nn, v) * dd;                                             /* :264-304 */
          }
        }
      }
      if (flags & LIN_NCP) { /* :306-329 */
        lin_ncp(s->pde, &QI(i, n, 0), (const double(*)[MAXLV])G, R);
        for (int d = 0; d < D; d++)
          for (int v = 0; v < V; v++) F[d][v] += R[d][v];
      }
      if (flags & LIN_MM) lin_mm(s->pde, &QI(i, n, 0), F); /* :332-344 */
      for (int d = 0; d < D; d++)
        for (int v = 0; v < V; v++) FI(i, n, d, v) = F[d][v];
    }
    for (int n = 0; n < BD; n++) { /* :361-392 */
      for (int d = 0; d < D; d++)
        for (int v = 0; v < V; v++) QI(i + 1, n, v) -= FI(i, n, d, v);
      for (int v = V; v < ND; v++) QI(i + 1, n, v) = QI(i, n, v);
    }
  }
  /* time averages (:395-480) */
  double a[MAXLV];
  taylor_factors(B, dt, a);
  memcpy(lQhi, lQi, sizeof(double) * BD * ND);
  for (int i = 1; i < B; i++)
    for (int n = 0; n < BD; n++)
      for (int v = 0; v < V; v++) lQhi[n * ND + v] += a[i] * QI(i, n, v);
  for (int i = 0; i < B; i++)
    for (int d = 0; d < D; d++)
      for (int n = 0; n < BD; n++)
        for (int v = 0; v < V; v++) lFhi[((size_t)d * BD + n) * V + v] += a[i] * FI(i, n, d, v);
  if (PSi && D == 3) { /* :482-503 */
    for (int d = 0; d < D; d++)
      for (int n = 0; n < BD; n++)
        for (int v = 0; v < V; v++) {
          double sum = 0.0;
          for (int t = 1; t < B + 1; t++) sum += w[t - 1] * PSi[((size_t)t * BD + n) * V + v];
          lFhi[((size_t)d * BD + n) * V + v] += sum / D;
        }
  }
  /* extrapolation (:507-640) */
  if (lQbnd) memset(lQbnd, 0, sizeof(double) * 2 * D * BF * ND);
  if (lFbnd) memset(lFbnd, 0, sizeof(double) * 2 * D * BF * V);
  for (int n = 0; n < BD; n++)
    for (int d = 0; d < D; d++) {
      const int l = coord(n, d, B), fn = face_index(n, d, B, D);
      if (lQbnd)
        for (int k = 0; k < ND; k++) {
          lQbnd[((size_t)(2 * d) * BF + fn) * ND + k] += lQhi[n * ND + k] * F0[l];
          lQbnd[((size_t)(2 * d + 1) * BF + fn) * ND + k] += lQhi[n * ND + k] * F1[l];
        }
      if (lFbnd)
        for (int k = 0; k < V; k++) {
          lFbnd[((size_t)(2 * d) * BF + fn) * V + k] += lFhi[((size_t)d * BD + n) * V + k] * F0[l];
          lFbnd[((size_t)(2 * d + 1) * BF + fn) * V + k] += lFhi[((size_t)d * BD + n) * V + k] * F1[l];
        }
    }
  if (lQhi_out) memcpy(lQhi_out, lQhi, sizeof(double) * BD * ND);
  if (lFhi_out) memcpy(lFhi_out, lFhi, sizeof(double) * (D + 1) * BD * V);
  if (levels) memcpy(levels, lQi, sizeof(double) * (B + 1) * BD * ND);
  free(lQi);
  free(lFi);
  free(lQhi);
  free(lFhi);
  free(Fn);
#undef QI
#undef FI
}

/* Split Cauchy-Kowalevski predictor of the optimised linear kernels
 * (KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_linear_split_ck_cpp.template,
 * "tpl" below) with its volume integral, dimension 2 or 3, the flux / NCP /
 * material-matrix switches of orc_lin_stp.  Only the current time-derivative
 * level is kept:
 *   L_0 = luh + dt sum_t w_t PSi_t                                    tpl:179-199
 *   L_{k+1} = sum_d [ MM(-D_d F_d(L_k)) - ncp_d(L_k, D_d L_k) ]
 *             + sum_t derPS_k[t] PSi_t,   derPS_0 = FLCoeff,
 *             derPS_{k+1}[j] = sum_i derPS_k[i] dudx[i][j]             tpl:201-384
 *   lQhi = L_0 + sum_{k=0}^{B-1} dt^(k+1)/(k+2)! L_{k+1}               tpl:386-395
 * (B derivative levels where the generic kernel sums B - 1; the material
 * matrix multiplies the flux part only, the NCP enters unscaled; the source
 * derivatives are those on the unit time interval).  Then lQbnd = lQhi and
 * the parameters extrapolated (tpl:402-470) and per direction d
 *   lFhi_d = sum_t w_t/D PSi_t + MM(D_d F_d(lQhi)) + ncp_d(lQhi, D_d lQhi),
 * lFbnd from lFhi_d, lduh -= W lFhi_d (tpl:474-640).
 * D_d f(node) = sum_m (1/dx_d) dudx[c_d][m] f(line node m).
 *   luh[B^D][nData]; PSi [B][B^D][V] or NULL; out lduh[B^D][V],
 *   lQhi_out[B^D][nData] (may be NULL), lQbnd[2D][B^(D-1)][nData], lFbnd[2D][B^(D-1)][V]. */
void orc_lin_stp_split(const orc_solver* s, int flags, const double* luh, const double* dx, double dt,
                       const double* PSi, double* lduh, double* lQhi_out, double* lQbnd, double* lFbnd) {
  const int D = s->dim, B = s->b.order + 1, V = s->nvar, ND = s->nvar + s->npar;
  const int BD = ipow(B, D), BF = ipow(B, D - 1);
  const double* w = s->b.w;
  const double* dudx = s->b.dudx;
  const double* F0 = s->b.F0;
  const double* F1 = s->b.F1;
  double* Q = (double*)malloc(sizeof(double) * BD * ND);  /* lQi with the parameters */
  double* Qh = (double*)malloc(sizeof(double) * BD * ND); /* lQhi */
  double* Qn = (double*)malloc(sizeof(double) * BD * V);  /* lQi_next */
  double* Fd = (double*)malloc(sizeof(double) * BD * V);  /* flux of one direction */
  double* Fh = (double*)malloc(sizeof(double) * BD * V);  /* lFhi of one direction */
  double* PS0 = (double*)calloc((size_t)BD * V, sizeof(double));
  double derPS[ORC_MAXB], buf[ORC_MAXB];
  memcpy(Q, luh, sizeof(double) * BD * ND);
  if (PSi) /* tpl:181-188 */
    for (int t = 0; t < B; t++)
      for (int n = 0; n < BD; n++)
        for (int v = 0; v < V; v++) Q[n * ND + v] += dt * w[t] * PSi[((size_t)t * BD + n) * V + v];
  memcpy(Qh, Q, sizeof(double) * BD * ND);
  for (int j = 0; j < B; j++) derPS[j] = F0[j];
  double dtav = 0.5 * dt;
  for (int t = 0; t < B; t++) {
    memset(Qn, 0, sizeof(double) * BD * V);
    for (int d = 0; d < D; d++) {
      const double invDx = 1.0 / dx[d];
      if (flags & LIN_FLUX) { /* tpl:204-223 (x), 237-257 (y), 271-289 (z) */
        for (int n = 0; n < BD; n++) {
          double F[3][MAXLV];
          lin_flux(s->pde, &Q[n * ND], F);
          for (int v = 0; v < V; v++) Fd[n * V + v] = F[d][v];
        }
        for (int n = 0; n < BD; n++) {
          const int c = coord(n, d, B);
          double R[3][MAXLV];
          memset(R, 0, sizeof(R));
          for (int m = 0; m < B; m++) {
            const double cn = -(invDx * dudx[c * B + m]); /* negativeDudxT_by_dx */
            const int nn = with_coord(n, d, m, B);
            for (int v = 0; v < V; v++) R[d][v] += Fd[nn * V + v] * cn;
          }
          if (flags & LIN_MM) lin_mm(s->pde, &Q[n * ND], R);
          for (int v = 0; v < V; v++) Qn[n * V + v] += R[d][v];
        }
      }
      if (flags & LIN_NCP) /* tpl:225-235, 259-269, 291-301 */
        for (int n = 0; n < BD; n++) {
          const int c = coord(n, d, B);
          double G[3][MAXLV], R[3][MAXLV];
          memset(G, 0, sizeof(G));
          for (int m = 0; m < B; m++) {
            const double cp = invDx * dudx[c * B + m]; /* dudxT_by_dx */
            const int nn = with_coord(n, d, m, B);
            for (int v = 0; v < V; v++) G[d][v] += Q[nn * ND + v] * cp;
          }
          lin_ncp(s->pde, &Q[n * ND], (const double(*)[MAXLV])G, R);
          for (int v = 0; v < V; v++) Qn[n * V + v] -= R[d][v];
        }
    }
    if (PSi) { /* tpl:313-331 */
      for (int k = 0; k < B; k++)
        for (int n = 0; n < BD; n++)
          for (int v = 0; v < V; v++) Qn[n * V + v] += derPS[k] * PSi[((size_t)k * BD + n) * V + v];
      for (int j = 0; j < B; j++) buf[j] = 0.0;
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++) buf[j] += derPS[i] * dudx[i * B + j];
      for (int j = 0; j < B; j++) derPS[j] = buf[j];
    }
    for (int n = 0; n < BD; n++) /* tpl:333-342 */
      for (int v = 0; v < V; v++) {
        Qh[n * ND + v] += dtav * Qn[n * V + v];
        Q[n * ND + v] = Qn[n * V + v];
      }
    dtav *= dt / (t + 2);
  }
  /* extrapolation of lQhi with the parameters (tpl:350-470) */
  memset(lQbnd, 0, sizeof(double) * 2 * D * BF * ND);
  for (int n = 0; n < BD; n++)
    for (int d = 0; d < D; d++) {
      const int l = coord(n, d, B), fn = face_index(n, d, B, D);
      for (int k = 0; k < ND; k++) {
        lQbnd[((size_t)(2 * d) * BF + fn) * ND + k] += Qh[n * ND + k] * F0[l];
        lQbnd[((size_t)(2 * d + 1) * BF + fn) * ND + k] += Qh[n * ND + k] * F1[l];
      }
    }
  if (PSi) /* tpl:477-494: time integral of the moments, shared by the directions */
    for (int t = 0; t < B; t++)
      for (int k = 0; k < BD * V; k++) PS0[k] += w[t] / D * PSi[(size_t)t * BD * V + k];
  memset(lduh, 0, sizeof(double) * BD * V);
  memset(lFbnd, 0, sizeof(double) * 2 * D * BF * V);
  for (int d = 0; d < D; d++) {
    const double invDx = 1.0 / dx[d];
    memcpy(Fh, PS0, sizeof(double) * BD * V);
    if (flags & LIN_FLUX) {
      for (int n = 0; n < BD; n++) {
        double F[3][MAXLV];
        lin_flux(s->pde, &Qh[n * ND], F);
        for (int v = 0; v < V; v++) Fd[n * V + v] = F[d][v];
      }
      for (int n = 0; n < BD; n++) {
        const int c = coord(n, d, B);
        double R[3][MAXLV];
        memset(R, 0, sizeof(R));
        for (int m = 0; m < B; m++) {
          const double cp = invDx * dudx[c * B + m];
          const int nn = with_coord(n, d, m, B);
          for (int v = 0; v < V; v++) R[d][v] += Fd[nn * V + v] * cp;
        }
        if (flags & LIN_MM) lin_mm(s->pde, &Qh[n * ND], R);
        for (int v = 0; v < V; v++) Fh[n * V + v] += R[d][v];
      }
    }
    if (flags & LIN_NCP)
      for (int n = 0; n < BD; n++) {
        const int c = coord(n, d, B);
        double G[3][MAXLV], R[3][MAXLV];
        memset(G, 0, sizeof(G));
        for (int m = 0; m < B; m++) {
          const double cp = invDx * dudx[c * B + m];
          const int nn = with_coord(n, d, m, B);
          for (int v = 0; v < V; v++) G[d][v] += Qh[nn * ND + v] * cp;
        }
        lin_ncp(s->pde, &Qh[n * ND], (const double(*)[MAXLV])G, R);
        for (int v = 0; v < V; v++) Fh[n * V + v] += R[d][v];
      }
    for (int n = 0; n < BD; n++) {
      const int x = coord(n, 0, B), y = coord(n, 1, B), l = coord(n, d, B), fn = face_index(n, d, B, D);
      const double W = (D == 3) ? w[coord(n, 2, B)] * w[y] * w[x] : w[y] * w[x];
      for (int v = 0; v < V; v++) {
        lFbnd[((size_t)(2 * d) * BF + fn) * V + v] += Fh[n * V + v] * F0[l];
        lFbnd[((size_t)(2 * d + 1) * BF + fn) * V + v] += Fh[n * V + v] * F1[l];
        lduh[n * V + v] -= Fh[n * V + v] * W;
      }
    }
  }
  if (lQhi_out) memcpy(lQhi_out, Qh, sizeof(double) * BD * ND);
  free(Q);
  free(Qh);
  free(Qn);
  free(Fd);
  free(Fh);
  free(PS0);
}

/* volumeIntegralLinear (volumeIntegralLinear.cpph:21-70): lduh = -W sum_d lFhi[d] */
void orc_lin_volume(const orc_solver* s, double* lduh, const double* lFhi) {
  const int D = s->dim, B = s->b.order + 1, V = s->nvar, BD = ipow(B, D);
  const double* w = s->b.w;
  for (int n = 0; n < BD; n++) {
    double W = 1.0;
    for (int d = 0; d < D; d++) W *= w[coord(n, d, B)];
    for (int v = 0; v < V; v++) {
      lduh[n * V + v] = 0.0;
      for (int d = 0; d < D; d++) lduh[n * V + v] -= W * lFhi[((size_t)d * BD + n) * V + v];
    }
  }
}

/* faceIntegralLinear (faceIntegralLinear.cpph:8-101):
 * lduh -= (prod of the face weights / dx_d) FCoeff[orientation][k] lFhbnd. */
void orc_lin_face(const orc_solver* s, double* lduh, const double* lFhbnd, int direction, int orientation,
                  const double* dx) {
  const int D = s->dim, B = s->b.order + 1, V = s->nvar, BD = ipow(B, D);
  const double* w = s->b.w;
  const double* FC = orientation ? s->b.F1 : s->b.F0;
  for (int n = 0; n < BD; n++) {
    double sw = 1.0 / dx[direction];
    for (int e = 0; e < D; e++)
      if (e != direction) sw *= w[coord(n, e, B)];
    const int fn = face_index(n, direction, B, D);
    const double c = sw * FC[coord(n, direction, B)];
    for (int v = 0; v < V; v++) lduh[n * V + v] -= c * lFhbnd[fn * V + v];
  }
}

/* surfaceIntegralLinear: all 2D faces of a cell, lFhbnd[2D][B^(D-1)][V] */
void orc_lin_surface(const orc_solver* s, double* lduh, const double* lFhbnd, const double* dx) {
  const int D = s->dim, B = s->b.order + 1, V = s->nvar, BF = ipow(B, D - 1);
  for (int f = 0; f < 2 * D; f++) orc_lin_face(s, lduh, lFhbnd + (size_t)f * BF * V, f / 2, f % 2, dx);
}

/* riemannSolverLinear (riemannSolverLinear.cpph:23-167), fluctuation form.
 * QL, QR [B^(D-1)][nData]; FL, FR [B^(D-1)][V] out. */
void orc_lin_riemann(const orc_solver* s, int flags, double* FL, double* FR, const double* QL, const double* QR,
                     int direction) {
  const int D = s->dim, B = s->b.order + 1, V = s->nvar, ND = s->nvar + s->npar, BF = ipow(B, D - 1);
  const double* w = s->b.w;
  double QavL[MAXLV] = {0}, QavR[MAXLV] = {0};
  for (int fn = 0; fn < BF; fn++) { /* averageRiemannInputs */
    double W = 1.0;
    for (int e = 0; e < D - 1; e++) W *= w[coord(fn, e, B)];
    for (int k = 0; k < ND; k++) {
      QavL[k] += W * QL[fn * ND + k];
      QavR[k] += W * QR[fn * ND + k];
    }
  }
  const double sL = lin_smax(s->pde, QavL), sR = lin_smax(s->pde, QavR);
  const double smax = sL > sR ? sL : sR;
  double Qavg[MAXLV], Qdiff[MAXLV];
  for (int l = 0; l < ND; l++) Qavg[l] = 0.5 * (QavR[l] + QavL[l]);
  for (int l = V; l < ND; l++) Qdiff[l] = Qavg[l];
  for (int fn = 0; fn < BF; fn++) {
    for (int l = 0; l < V; l++) Qdiff[l] = 0.5 * (QR[fn * ND + l] - QL[fn * ND + l]);
    double acc[MAXLV] = {0};
    if (flags & LIN_NCP) {
      double G[3][MAXLV], R[3][MAXLV];
      memset(G, 0, sizeof(G));
      for (int d = 0; d < D; d++)
        for (int l = 0; l < V; l++) G[d][l] = Qdiff[l];
      lin_ncp(s->pde, Qavg, (const double(*)[MAXLV])G, R);
      if (flags & LIN_MM) lin_mm(s->pde, Qavg, R);
      for (int l = 0; l < V; l++) acc[l] += R[direction][l];
    }
    if (flags & LIN_FLUX) {
      double F[3][MAXLV];
      lin_flux(s->pde, Qdiff, F);
      if (flags & LIN_MM) lin_mm(s->pde, Qavg, F);
      for (int l = 0; l < V; l++) acc[l] += F[direction][l];
    }
    for (int l = 0; l < V; l++) {
      FL[fn * V + l] = acc[l] - smax * Qdiff[l];
      FR[fn * V + l] = acc[l] + smax * Qdiff[l];
    }
  }
}

/* ---------------------------------------------------------------------------- */
/* Acoustic mesh: periodic n^3 cells of edge h, fixed dt, one point source,     */
/* receivers recording p after every step.                                      */
/* ---------------------------------------------------------------------------- */
#define AV 4 /* p, u, v, w */
#define AD 6 /* + rho, c */
#define ACF (LIN_FLUX | LIN_NCP | LIN_MM)

typedef struct {
  int cell;
  double phi[3][MAXLV]; /* phi_k(xi_d) at the receiver's reference coordinates */
} orc_receiver;

static void setup_receivers(const orc_solver* s, int n, double h, int nrec, const double* xr, orc_receiver* R) {
  const int B = s->b.order + 1;
  for (int r = 0; r < nrec; r++) {
    int ic[3];
    for (int d = 0; d < 3; d++) {
      ic[d] = (int)floor(xr[3 * r + d] / h);
      if (ic[d] < 0) ic[d] = 0;
      if (ic[d] >= n) ic[d] = n - 1;
      const double xi = (xr[3 * r + d] - ic[d] * h) / h;
      for (int k = 0; k < B; k++) R[r].phi[d][k] = lagr(s->b.x, B, k, xi);
    }
    R[r].cell = (ic[2] * n + ic[1]) * n + ic[0];
  }
}

static double receiver_p(const orc_solver* s, const double* luh, const orc_receiver* R) {
  const int B = s->b.order + 1, B3 = B * B * B;
  const double* u = luh + (size_t)R->cell * B3 * AD;
  double p = 0.0;
  for (int z = 0; z < B; z++)
    for (int y = 0; y < B; y++)
      for (int x = 0; x < B; x++) p += u[((z * B + y) * B + x) * AD] * R->phi[0][x] * R->phi[1][y] * R->phi[2][z];
  return p;
}

static long source_cell(int n, double h, const double* src_x) {
  if (!src_x) return -1;
  int ic[3];
  for (int d = 0; d < 3; d++) {
    ic[d] = (int)floor(src_x[d] / h);
    if (ic[d] < 0) ic[d] = 0;
    if (ic[d] >= n) ic[d] = n - 1;
  }
  return ((long)ic[2] * n + ic[1]) * n + ic[0];
}

static long neighbour(long c, int n, int d) {
  long ix = c % n, iy = (c / n) % n, iz = c / ((long)n * n);
  if (d == 0) ix = (ix + 1) % n;
  if (d == 1) iy = (iy + 1) % n;
  if (d == 2) iz = (iz + 1) % n;
  return (iz * n + iy) * n + ix;
}

/* predictor + volume integral of every cell (lFb, levels optional) */
static void lin_predict_all(const orc_solver* s, const double* luh, int n, double h, double t, double dt, long src_cell,
                            const double* src_x, double amp, double f0, double t0, double* lduh, double* lQb,
                            double* lFb, double* levels) {
  const int B = s->b.order + 1, B2 = B * B, B3 = B2 * B;
  const long nc = (long)n * n * n;
  const double dxv[3] = {h, h, h};
#pragma omp parallel for schedule(dynamic, 8)
  for (long c = 0; c < nc; c++) {
    double* PSi = NULL;
    if (c == src_cell) {
      const long ix = c % n, iy = (c / n) % n, iz = c / ((long)n * n);
      const double centre[3] = {(ix + 0.5) * h, (iy + 0.5) * h, (iz + 0.5) * h};
      PSi = (double*)malloc(sizeof(double) * (B + 1) * B3 * AV);
      if (s->split_ck)
        orc_lin_psi_split(s, centre, h, t, dt, src_x, amp, f0, t0, PSi);
      else
        orc_lin_psi(s, centre, h, t, dt, src_x, amp, f0, t0, PSi);
    }
    if (s->split_ck) { /* lFb is overwritten by the Riemann fluctuations: scratch here */
      double* lFs = (double*)malloc(sizeof(double) * 6 * B2 * AV);
      orc_lin_stp_split(s, ACF, luh + c * B3 * AD, dxv, dt, PSi, lduh + c * B3 * AV, NULL, lQb + c * 6 * B2 * AD,
                        lFb ? lFb + c * 6 * B2 * AV : lFs);
      free(lFs);
      free(PSi);
      continue;
    }
    double* lFhi = (double*)malloc(sizeof(double) * 4 * B3 * AV);
    orc_lin_stp(s, ACF, luh + c * B3 * AD, dxv, dt, PSi, NULL, lFhi, lQb + c * 6 * B2 * AD,
                lFb ? lFb + c * 6 * B2 * AV : NULL, levels ? levels + (size_t)c * (B + 1) * B3 * AD : NULL);
    orc_lin_volume(s, lduh + c * B3 * AV, lFhi);
    free(lFhi);
    free(PSi);
  }
}

/* One linear step on the periodic mesh: predictor+volume, Riemann per face,
 * face integrals and update (the order of ADERDGSolver.cpp's fused step). */
static void lin_step(const orc_solver* s, double* luh, int n, double h, double t, double dt, long src_cell,
                     const double* src_x, double amp, double f0, double t0, double* lduh, double* lQb, double* lFb) {
  const int B = s->b.order + 1, B2 = B * B, B3 = B2 * B;
  const long nc = (long)n * n * n;
  const double dxv[3] = {h, h, h};
  lin_predict_all(s, luh, n, h, t, dt, src_cell, src_x, amp, f0, t0, lduh, lQb, lFb, NULL);
  for (int d = 0; d < 3; d++) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; c++) {
      const long r = neighbour(c, n, d);
      orc_lin_riemann(s, ACF, lFb + (c * 6 + 2 * d + 1) * B2 * AV, lFb + (r * 6 + 2 * d) * B2 * AV,
                      lQb + (c * 6 + 2 * d + 1) * B2 * AD, lQb + (r * 6 + 2 * d) * B2 * AD, d);
    }
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; c++) {
    orc_lin_surface(s, lduh + c * B3 * AV, lFb + c * 6 * B2 * AV, dxv);
    const double* w = s->b.w;
    for (int nd = 0; nd < B3; nd++) {
      const double upd = dt / (w[nd % B] * w[(nd / B) % B] * w[nd / B2]);
      for (int v = 0; v < AV; v++) luh[(c * B3 + nd) * AD + v] += lduh[(c * B3 + nd) * AV + v] * upd;
    }
  }
}

/* Forward run of nsteps fixed-dt steps from time 0 (s->nvar must be 4 and
 * s->npar 2); luh advanced in place; rec[nsteps][nrec] receives p at the
 * receivers after every step; states (optional) receives the state before
 * every step. */
int orc_lin_run(const orc_solver* s, double* luh, int n, double h, double dt, int nsteps, const double* src_x,
                double amp, double f0, double t0, int nrec, const double* xr, double* rec, double* states, int threads) {
  const int B = s->b.order + 1, B2 = B * B, B3 = B2 * B;
  const long nc = (long)n * n * n;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  double* lduh = (double*)malloc(sizeof(double) * nc * B3 * AV);
  double* lQb = (double*)malloc(sizeof(double) * nc * 6 * B2 * AD);
  double* lFb = (double*)malloc(sizeof(double) * nc * 6 * B2 * AV);
  orc_receiver* R = (orc_receiver*)malloc(sizeof(orc_receiver) * (nrec > 0 ? nrec : 1));
  if (!lduh || !lQb || !lFb || !R) return -1;
  setup_receivers(s, n, h, nrec, xr, R);
  const long src_cell = source_cell(n, h, src_x);
  for (int k = 0; k < nsteps; k++) {
    if (states) memcpy(states + (size_t)k * nc * B3 * AD, luh, sizeof(double) * nc * B3 * AD);
    lin_step(s, luh, n, h, k * dt, dt, src_cell, src_x, amp, f0, t0, lduh, lQb, lFb);
    for (int r = 0; r < nrec; r++) rec[(size_t)k * nrec + r] = receiver_p(s, luh, &R[r]);
  }
  free(lduh);
  free(lQb);
  free(lFb);
  free(R);
  return 0;
}

/* ---------------------------------------------------------------------------- */
/* Discrete adjoint (acoustic, 3D)                                               */
/* ---------------------------------------------------------------------------- */

/* Transpose of predictor + volume integral for one cell.  L: forward levels
 * lQi[(B+1)][B^3][6]; lduh_bar[B^3][4]: cotangent of the volume integral;
 * Qb_bar[6][B^2][4]: cotangent of the extrapolated variables.  Accumulates
 * u_bar[B^3][4]; returns dJ/dlambda for the cell's lambda = rho c^2.
 * Forward per level i and direction d (acoustic hooks):
 *   lFi[i][d][0] = lambda * (-D_d u_d),  lFi[i][d][1+d] = (1/rho) * (-D_d p)
 *   L[i+1] = -sum_d lFi[i][d];  lQhi = sum_i a_i L[i];  lFhi[d] = sum_i a_i lFi[i][d]
 *   lduh = -W sum_d lFhi[d];    lQbnd = E lQhi. */
static double lin_stp_adjoint(const orc_solver* s, const double* L, double h, double dt, const double* lduh_bar,
                              const double* Qb_bar, double* u_bar, double rho, double lam) {
  const int B = s->b.order + 1, B2 = B * B, B3 = B2 * B;
  const double* w = s->b.w;
  const double* dudx = s->b.dudx;
  const double* F0 = s->b.F0;
  const double* F1 = s->b.F1;
  const double invDx = 1.0 / h, irho = 1.0 / rho;
  double* Fhb = (double*)calloc((size_t)3 * B3 * AV, sizeof(double)); /* lFhi_bar[d][n][v] */
  double* Qhb = (double*)calloc((size_t)B3 * AV, sizeof(double));     /* lQhi_bar[n][v] */
  double* Lb = (double*)calloc((size_t)(B + 1) * B3 * AV, sizeof(double));
  double* T = (double*)malloc(sizeof(double) * 3 * B3 * AV);
  for (int nd = 0; nd < B3; nd++) {
    const double W = w[nd % B] * w[(nd / B) % B] * w[nd / B2];
    for (int d = 0; d < 3; d++)
      for (int v = 0; v < AV; v++) Fhb[((size_t)d * B3 + nd) * AV + v] = -W * lduh_bar[nd * AV + v];
    for (int d = 0; d < 3; d++) {
      const int l = coord(nd, d, B), fn = face_index(nd, d, B, 3);
      for (int v = 0; v < AV; v++)
        Qhb[nd * AV + v] +=
            F0[l] * Qb_bar[((2 * d) * B2 + fn) * AV + v] + F1[l] * Qb_bar[((2 * d + 1) * B2 + fn) * AV + v];
    }
  }
  double a[MAXLV];
  taylor_factors(B, dt, a);
  for (int i = 0; i < B; i++)
    for (int k = 0; k < B3 * AV; k++) Lb[(size_t)i * B3 * AV + k] += a[i] * Qhb[k];
  double glam = 0.0;
  for (int i = B - 1; i >= 0; i--) {
    /* cotangent of lFi[i][d]: -Lb[i+1] + a_i Fhb[d] */
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < B3 * AV; k++)
        T[(size_t)d * B3 * AV + k] = -Lb[(size_t)(i + 1) * B3 * AV + k] + a[i] * Fhb[(size_t)d * B3 * AV + k];
    for (int nd = 0; nd < B3; nd++)
      for (int d = 0; d < 3; d++) {
        const int c = coord(nd, d, B);
        const double tp = T[((size_t)d * B3 + nd) * AV + 0];
        const double tu = T[((size_t)d * B3 + nd) * AV + 1 + d];
        double Du = 0.0;
        for (int m = 0; m < B; m++) {
          const int nn = with_coord(nd, d, m, B);
          const double dd = dudx[c * B + m] * invDx;
          Du += dd * L[((size_t)i * B3 + nn) * AD + 1 + d];
          Lb[((size_t)i * B3 + nn) * AV + 1 + d] += dd * (-lam * tp);
          Lb[((size_t)i * B3 + nn) * AV + 0] += dd * (-irho * tu);
        }
        glam += tp * (-Du);
      }
  }
  for (int k = 0; k < B3 * AV; k++) u_bar[k] += Lb[k];
  free(T);
  free(Fhb);
  free(Qhb);
  free(Lb);
  return glam;
}

/* Transpose of one step u^{k+1} = Step(u^k; c).  ubar_next: cotangent of
 * u^{k+1} (variables, [nc][B^3][4]); writes the cotangent of u^k into ubar
 * and accumulates dJ/dc per cell into gc.  u: the forward state u^k. */
static void lin_step_adjoint(const orc_solver* s, const double* u, int n, double h, double t, double dt, long src_cell,
                             const double* src_x, double amp, double f0, double t0, const double* ubar_next,
                             double* ubar, double* gc) {
  const int B = s->b.order + 1, B2 = B * B, B3 = B2 * B;
  const long nc = (long)n * n * n;
  const double* w = s->b.w;
  double* lduh_bar = (double*)malloc(sizeof(double) * nc * B3 * AV);
  double* Fb_bar = (double*)calloc((size_t)nc * 6 * B2 * AV, sizeof(double));
  double* Qb_bar = (double*)calloc((size_t)nc * 6 * B2 * AV, sizeof(double));
  double* lQb = (double*)malloc(sizeof(double) * nc * 6 * B2 * AD);
  double* lduh = (double*)malloc(sizeof(double) * nc * B3 * AV);
  double* levels = (double*)malloc(sizeof(double) * nc * (B + 1) * B3 * AD);
  lin_predict_all(s, u, n, h, t, dt, src_cell, src_x, amp, f0, t0, lduh, lQb, NULL, levels);
  /* update u' = u + dt/W lduh */
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; c++)
    for (int nd = 0; nd < B3; nd++) {
      const double upd = dt / (w[nd % B] * w[(nd / B) % B] * w[nd / B2]);
      for (int v = 0; v < AV; v++) {
        const size_t k = ((size_t)c * B3 + nd) * AV + v;
        ubar[k] = ubar_next[k];
        lduh_bar[k] = upd * ubar_next[k];
      }
    }
  /* face integrals: transpose of lduh -= sw FC[k] F  ->  F_bar -= sw FC[k] lduh_bar */
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; c++)
    for (int nd = 0; nd < B3; nd++)
      for (int d = 0; d < 3; d++) {
        const int l = coord(nd, d, B), fn = face_index(nd, d, B, 3);
        double sw = 1.0 / h;
        for (int e = 0; e < 3; e++)
          if (e != d) sw *= w[coord(nd, e, B)];
        for (int o = 0; o < 2; o++) {
          const double cf = sw * (o ? s->b.F1[l] : s->b.F0[l]);
          for (int v = 0; v < AV; v++)
            Fb_bar[(((size_t)c * 6 + 2 * d + o) * B2 + fn) * AV + v] -= cf * lduh_bar[((size_t)c * B3 + nd) * AV + v];
        }
      }
  /* Riemann transpose (serial: both sides accumulate) */
  for (int d = 0; d < 3; d++)
    for (long c = 0; c < nc; c++) {
      const long r = neighbour(c, n, d);
      const double* QL = lQb + (c * 6 + 2 * d + 1) * B2 * AD;
      const double* QR = lQb + (r * 6 + 2 * d) * B2 * AD;
      const double* FLb = Fb_bar + (c * 6 + 2 * d + 1) * B2 * AV;
      const double* FRb = Fb_bar + (r * 6 + 2 * d) * B2 * AV;
      double* QLb = Qb_bar + (c * 6 + 2 * d + 1) * B2 * AV;
      double* QRb = Qb_bar + (r * 6 + 2 * d) * B2 * AV;
      double QavL[AD] = {0}, QavR[AD] = {0};
      for (int fn = 0; fn < B2; fn++)
        for (int k = 0; k < AD; k++) {
          QavL[k] += w[fn % B] * w[fn / B] * QL[fn * AD + k];
          QavR[k] += w[fn % B] * w[fn / B] * QR[fn * AD + k];
        }
      const double cL = fabs(QavL[5]), cR = fabs(QavR[5]);
      const double smax = cL > cR ? cL : cR;
      const double rhoa = 0.5 * (QavL[4] + QavR[4]), ca = 0.5 * (QavL[5] + QavR[5]);
      const double lama = rhoa * ca * ca, irhoa = 1.0 / rhoa;
      double gsmax = 0.0, glama = 0.0;
      for (int fn = 0; fn < B2; fn++) {
        double Qd[AV], sb[AV], db[AV], Qdb[AV];
        for (int l = 0; l < AV; l++) {
          Qd[l] = 0.5 * (QR[fn * AD + l] - QL[fn * AD + l]);
          sb[l] = FLb[fn * AV + l] + FRb[fn * AV + l];
          db[l] = FRb[fn * AV + l] - FLb[fn * AV + l];
          Qdb[l] = smax * db[l];
          gsmax += db[l] * Qd[l];
        }
        /* acc[0] = -lam Qd[1+d], acc[1+d] = -(1/rho) Qd[0] (flux and NCP parts) */
        Qdb[1 + d] += -lama * sb[0];
        Qdb[0] += -irhoa * sb[1 + d];
        glama += sb[0] * (-Qd[1 + d]);
        for (int l = 0; l < AV; l++) {
          QRb[fn * AV + l] += 0.5 * Qdb[l];
          QLb[fn * AV + l] -= 0.5 * Qdb[l];
        }
      }
      /* lam_avg = rho_avg c_avg^2 with c_avg = (c_L + c_R)/2; smax = max(c_L, c_R) */
      gc[c] += glama * rhoa * ca;
      gc[r] += glama * rhoa * ca;
      if (cL >= cR)
        gc[c] += gsmax;
      else
        gc[r] += gsmax;
    }
#pragma omp parallel for schedule(dynamic, 8)
  for (long c = 0; c < nc; c++) {
    const double rho = u[c * B3 * AD + 4], cc = u[c * B3 * AD + 5];
    const double glam = lin_stp_adjoint(s, levels + (size_t)c * (B + 1) * B3 * AD, h, dt, lduh_bar + c * B3 * AV,
                                        Qb_bar + c * 6 * B2 * AV, ubar + c * B3 * AV, rho, rho * cc * cc);
    gc[c] += glam * 2.0 * rho * cc;
  }
  free(lduh_bar);
  free(Fb_bar);
  free(Qb_bar);
  free(lQb);
  free(lduh);
  free(levels);
}

/* Misfit J = 1/2 sum_k sum_r (p_r(t_{k+1}) - obs[k][r])^2 of a forward run
 * from u0 and (if gc) its gradient with respect to each cell's wave speed
 * (parameter slot 5, equal at all nodes of a cell).  Stores every state
 * (tiny meshes only). */
double orc_lin_misfit_grad(const orc_solver* s, const double* u0, int n, double h, double dt, int nsteps,
                           const double* src_x, double amp, double f0, double t0, int nrec, const double* xr,
                           const double* obs, double* gc, int threads) {
  const int B = s->b.order + 1, B3 = B * B * B;
  const long nc = (long)n * n * n;
  const size_t ns = (size_t)nc * B3 * AD;
  if (s->split_ck) return NAN; /* the adjoint transposes the generic predictor only */
  double* states = (double*)malloc(sizeof(double) * ns * nsteps);
  double* u = (double*)malloc(sizeof(double) * ns);
  double* rec = (double*)malloc(sizeof(double) * nsteps * (nrec > 0 ? nrec : 1));
  memcpy(u, u0, sizeof(double) * ns);
  orc_lin_run(s, u, n, h, dt, nsteps, src_x, amp, f0, t0, nrec, xr, rec, states, threads);
  double J = 0.0;
  for (int k = 0; k < nsteps * nrec; k++) J += 0.5 * (rec[k] - obs[k]) * (rec[k] - obs[k]);
  if (gc) {
    orc_receiver* R = (orc_receiver*)malloc(sizeof(orc_receiver) * (nrec > 0 ? nrec : 1));
    setup_receivers(s, n, h, nrec, xr, R);
    const long src_cell = source_cell(n, h, src_x);
    const size_t nv = (size_t)nc * B3 * AV;
    double* ub = (double*)calloc(nv, sizeof(double));
    double* ub2 = (double*)calloc(nv, sizeof(double));
    memset(gc, 0, sizeof(double) * nc);
    for (int k = nsteps - 1; k >= 0; k--) {
      for (int r = 0; r < nrec; r++) {
        const double res = rec[(size_t)k * nrec + r] - obs[(size_t)k * nrec + r];
        for (int nd = 0; nd < B3; nd++)
          ub[((size_t)R[r].cell * B3 + nd) * AV + 0] +=
              res * R[r].phi[0][nd % B] * R[r].phi[1][(nd / B) % B] * R[r].phi[2][nd / (B * B)];
      }
      lin_step_adjoint(s, states + (size_t)k * ns, n, h, k * dt, dt, src_cell, src_x, amp, f0, t0, ub, ub2, gc);
      double* tmp = ub;
      ub = ub2;
      ub2 = tmp;
    }
    free(ub);
    free(ub2);
    free(R);
  }
  free(states);
  free(u);
  free(rec);
  return J;
}

/*
 * ADER-DG oracle -- plain-C FP64 restatement of the reference's generic
 * nonlinear ADER-DG kernels.  TEST INFRASTRUCTURE ONLY (see aderdg_oracle.h):
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are the
 * only users; the product library never touches it.
 *
 * The loop nests and the order of the floating-point operations follow the
 * cited reference loops so that results agree to rounding level.  Paths are
 * relative to the ExaHyPE-Engine root; "gc3" = ExaHyPE/kernels/aderdg/generic/c/3d,
 * "gc2" = .../generic/c/2d.
 */
#include "aderdg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GAMMA 1.4

/* ------------------------------------------------------------------------ */
/* PDE hooks                                                                  */
/* ------------------------------------------------------------------------ */

/* Euler flux: ApplicationExamples/Euler/Euler_ADERDG/EulerSolver_ADERDG.cpp:61-115.
 * In 2D the reference keeps 5 variables and drops Q[3] from |j|^2 (:81-85).
 * Linear advection (ORC_PDE_ADVECTION) is a test PDE: F_d = a_d Q. */
void orc_flux(const orc_solver* s, const double* Q, double* F) {
  const int V = s->nvar;
  if (s->pde == ORC_PDE_ADVECTION || s->pde == ORC_PDE_ADV_NCP_SRC) {
    for (int d = 0; d < s->dim; d++)
      for (int v = 0; v < V; v++) F[d * V + v] = s->adv[d] * Q[v];
    return;
  }
  const double gamma = GAMMA;
  const double irho = 1. / Q[0];
  const double j2 = (s->dim == 3) ? Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]
                                  : Q[1] * Q[1] + Q[2] * Q[2];
  const double p = (gamma - 1) * (Q[4] - 0.5 * irho * j2);
  double* f = F;
  double* g = F + V;
  f[0] = Q[1];
  f[1] = irho * Q[1] * Q[1] + p;
  f[2] = irho * Q[2] * Q[1];
  f[3] = irho * Q[3] * Q[1];
  f[4] = irho * (Q[4] + p) * Q[1];
  g[0] = Q[2];
  g[1] = irho * Q[1] * Q[2];
  g[2] = irho * Q[2] * Q[2] + p;
  g[3] = irho * Q[3] * Q[2];
  g[4] = irho * (Q[4] + p) * Q[2];
  if (s->dim == 3) {
    double* h = F + 2 * V;
    h[0] = Q[3];
    h[1] = irho * Q[1] * Q[3];
    h[2] = irho * Q[2] * Q[3];
    h[3] = irho * Q[3] * Q[3] + p;
    h[4] = irho * (Q[4] + p) * Q[3];
  }
}

/* Euler eigenvalues: EulerSolver_ADERDG.cpp:117-147. */
void orc_eigenvalues(const orc_solver* s, const double* Q, int d, double* lambda) {
  const int V = s->nvar;
  if (s->pde == ORC_PDE_ADVECTION) {
    for (int v = 0; v < V; v++) lambda[v] = s->adv[d];
    return;
  }
  if (s->pde == ORC_PDE_ADV_NCP_SRC) {  /* eigenvalues of a_d + b_d (flux and NCP together) */
    for (int v = 0; v < V; v++) lambda[v] = s->adv[d] + s->ncp_b[d];
    return;
  }
  const double gamma = GAMMA;
  const double irho = 1. / Q[0];
  const double j2 = (s->dim == 3) ? Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]
                                  : Q[1] * Q[1] + Q[2] * Q[2];
  const double p = (gamma - 1) * (Q[4] - 0.5 * irho * j2);
  const double u_n = Q[d + 1] * irho;
  const double c = sqrt(gamma * p * irho);
  for (int v = 0; v < 5; v++) lambda[v] = u_n;
  lambda[0] -= c;
  lambda[4] += c;
}

/* Test PDE ORC_PDE_ADV_NCP_SRC: BgradQ_v = sum_d b_d dQ_v/dx_d, S_v = -kappa Q_v. */
void orc_ncp(const orc_solver* s, const double* Q, const double* gradQ, double* BgradQ) {
  const int V = s->nvar;
  (void)Q;
  for (int v = 0; v < V; v++) {
    double a = 0.0;
    for (int d = 0; d < s->dim; d++) a += s->ncp_b[d] * gradQ[d * V + v];
    BgradQ[v] = a;
  }
}

void orc_source(const orc_solver* s, const double* Q, double* S) {
  for (int v = 0; v < s->nvar; v++) S[v] = -s->kappa * Q[v];
}

/* Source and NCP terms of the Picard loop at one space-time point
 * (gc3/spaceTimePredictorNonlinear.cpph:323-367, gc2 :211-243): S =
 * algebraicSource(Q) - nonConservativeProduct(Q, gradQ). */
static int has_ncs(const orc_solver* s) { return s->pde == ORC_PDE_ADV_NCP_SRC; }

static void ncs_term(const orc_solver* s, const double* Q, const double* gradQ, double* S) {
  double ncp[64];
  orc_source(s, Q, S);
  orc_ncp(s, Q, gradQ, ncp);
  for (int v = 0; v < s->nvar; v++) S[v] -= ncp[v];
}

/* ------------------------------------------------------------------------ */
/* CERK initial guess of the optimised kernels                                 */
/* (KernelGenerator/kernelgenerator/templates/aderdg/                          */
/*  fusedSPTVI_nonlinear_cpp.template:121-173 and                              */
/*  templates/subtemplates/RK_loop.template:12-40):                            */
/*   K1 = sum_d dudx_d F_d(luh) / dx_d   (flux derivative in x, then y, z)     */
/*   lwh = luh - dt K1 (parameters copied),  K2 = the same of lwh,             */
/*   lQi(t) = luh - dt xi_t K1 - 1/2 dt xi_t^2 (K2 - K1)  (parameters copied). */
/* The template stores dudx_T scaled by 1/dx and accumulates the three         */
/* directions into K with beta = 1; lQi is written in this oracle's generic    */
/* layout [space][t][nData].                                                   */
/* ------------------------------------------------------------------------ */
static void flux_divergence(const orc_solver* s, const double* u, int ldu, const double* invDx,
                            double* K /* [space][V] */) {
  const int B = s->b.order + 1, V = s->nvar, D = s->dim;
  const int NS = (D == 3) ? B * B * B : B * B;
  double* F = (double*)malloc(sizeof(double) * NS * D * V);
  for (int i = 0; i < NS; i++) orc_flux(s, u + i * ldu, F + i * D * V);
  for (int i = 0; i < NS * V; i++) K[i] = 0.0;
  for (int d = 0; d < D; d++) {
    const int stride = (d == 0) ? 1 : (d == 1 ? B : B * B);
    for (int i = 0; i < NS; i++) {
      const int c = (i / stride) % B;       /* coordinate of node i along d */
      const int base = i - c * stride;      /* node of the line at coordinate 0 */
      for (int v = 0; v < V; v++) {
        double a = 0.0;
        for (int n = 0; n < B; n++)
          a += F[(base + n * stride) * D * V + d * V + v] * (invDx[d] * s->b.dudx[c * B + n]);
        K[i * V + v] += a;
      }
    }
  }
  free(F);
}

static void cerk_guess(const orc_solver* s, const double* luh, const double* invDx, double dt,
                       double* lQi) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar, D = s->dim;
  const int NS = (D == 3) ? B * B * B : B * B;
  double* K1 = (double*)malloc(sizeof(double) * NS * V);
  double* K2 = (double*)malloc(sizeof(double) * NS * V);
  double* lwh = (double*)malloc(sizeof(double) * NS * ND);
  flux_divergence(s, luh, ND, invDx, K1);
  for (int i = 0; i < NS; i++) {
    for (int v = 0; v < V; v++) lwh[i * ND + v] = luh[i * ND + v] - dt * K1[i * V + v];
    for (int v = V; v < ND; v++) lwh[i * ND + v] = luh[i * ND + v];
  }
  flux_divergence(s, lwh, ND, invDx, K2);
  const double* xi = s->b.x;
  for (int i = 0; i < NS; i++)
    for (int t = 0; t < B; t++) {
      double* q = lQi + (i * B + t) * ND;
      for (int v = 0; v < V; v++)
        q[v] = luh[i * ND + v] - (dt * xi[t] * K1[i * V + v]) -
               (0.5 * dt * xi[t] * xi[t] * (K2[i * V + v] - K1[i * V + v]));
      for (int v = V; v < ND; v++) q[v] = luh[i * ND + v];
    }
  free(K1);
  free(K2);
  free(lwh);
}

/* MaxIterations of the Picard loop: generic (gc3 :85-86) or optimised rule. */
static int picard_max_iterations(const orc_solver* s) {
  const int nDof = s->b.order + 1;
  if (s->opt_iterations) return s->cerk_guess ? (nDof - 3 <= 1 ? 1 : nDof - 3) : 2 * nDof + 1;
  return s->use_max_picard ? s->max_picard : 2 * (s->b.order + 1);
}

/* ------------------------------------------------------------------------ */
/* Space-time predictor, 3D: gc3/spaceTimePredictorNonlinear.cpph             */
/* ------------------------------------------------------------------------ */

static int picard3(const orc_solver* s, const double* luh, const double* invDx, double dt,
                   double* lQi, double* rhs, double* lFi) {
  const int B = s->b.order + 1, V = s->nvar, P = s->npar, ND = V + P, D = 3;
  const double* w = s->b.w;
  const double* Kxi = s->b.Kxi;
  const double* iK1 = s->b.iK1;
  const double* F0 = s->b.F0;
#define LUH(z, y, x, v) luh[(((z) * B + (y)) * B + (x)) * ND + (v)]
#define LQI(z, y, x, t, v) lQi[((((z) * B + (y)) * B + (x)) * B + (t)) * ND + (v)]
#define LFI(t, z, y, x, d, v) lFi[(((((t) * B + (z)) * B + (y)) * B + (x)) * (D + 1) + (d)) * V + (v)]
#define RHS(t, z, y, x, v) rhs[((((t) * B + (z)) * B + (y)) * B + (x)) * V + (v)]
  /* 1. trivial initial guess (:71-82), or the optimised kernels' CERK guess */
  if (s->cerk_guess) {
    cerk_guess(s, luh, invDx, dt, lQi);
  } else {
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++)
            for (int m = 0; m < ND; m++) LQI(i, j, k, l, m) = LUH(i, j, k, m);
  }
  /* 2. Picard iterations (:84-432) */
  const int maxIter = picard_max_iterations(s);
  int iter = 0;
  double Fnode[3 * 64];
  for (; iter < maxIter; iter++) {
    for (int i = 0; i < B; i++) {
      /* fluxes (:196-227) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++) {
            orc_flux(s, &LQI(j, k, l, i, 0), Fnode);
            for (int d = 0; d < D; d++)
              for (int m = 0; m < V; m++) LFI(i, j, k, l, d, m) = Fnode[d * V + m];
            for (int m = 0; m < V; m++) LFI(i, j, k, l, 3, m) = 0.0;
          }
      /* rhs0 (:231-246) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++) {
            const double weight = w[j] * w[k] * w[l];
            for (int m = 0; m < V; m++) RHS(i, j, k, l, m) = weight * F0[i] * LUH(j, k, l, m);
          }
      /* x direction (:248-271) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = w[i] * w[j] * w[k];
          const double updateSize = weight * dt * invDx[0];
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++)
              for (int n = 0; n < B; n++)
                RHS(i, j, k, l, m) -= updateSize * LFI(i, j, k, n, 0, m) * Kxi[n * B + l];
        }
      /* y direction (:273-296) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = w[i] * w[j] * w[k];
          const double updateSize = weight * dt * invDx[1];
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++)
              for (int n = 0; n < B; n++)
                RHS(i, j, l, k, m) -= updateSize * LFI(i, j, n, k, 1, m) * Kxi[n * B + l];
        }
      /* z direction (:298-321) */
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = w[i] * w[j] * w[k];
          const double updateSize = weight * dt * invDx[2];
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++)
              for (int n = 0; n < B; n++)
                RHS(i, l, j, k, m) -= updateSize * LFI(i, n, j, k, 2, m) * Kxi[n * B + l];
        }
      if (has_ncs(s)) {
        /* gradQCur of time node i (:113-176): invDx_d lQi dudx along x, y, z;
         * then S = source - ncp into slot 3 of lFi and w^4 dt S into rhs (:323-367) */
        double gq[3 * 64], S[64];
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++)
            for (int l = 0; l < B; l++) {
              for (int m = 0; m < V; m++) {
                double gx = 0.0, gy = 0.0, gz = 0.0;
                for (int n = 0; n < B; n++) {
                  gx += 1.0 * invDx[0] * LQI(j, k, n, i, m) * s->b.dudx[l * B + n];
                  gy += 1.0 * invDx[1] * LQI(j, n, l, i, m) * s->b.dudx[k * B + n];
                  gz += 1.0 * invDx[2] * LQI(n, k, l, i, m) * s->b.dudx[j * B + n];
                }
                gq[0 * V + m] = gx;
                gq[1 * V + m] = gy;
                gq[2 * V + m] = gz;
              }
              ncs_term(s, &LQI(j, k, l, i, 0), gq, S);
              const double weight = w[i] * w[j] * w[k] * w[l];
              const double updateSize = weight * dt;
              for (int m = 0; m < V; m++) {
                LFI(i, j, k, l, 3, m) = S[m];
                RHS(i, j, k, l, m) += updateSize * S[m];
              }
            }
      }
    }
    /* 3. multiply with iK1 (:370-405) */
    double sq_res = 0.0;
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = w[i] * w[j] * w[k];
          const double iweight = 1.0 / weight;
          for (int l = 0; l < B; l++)
            for (int m = 0; m < V; m++) {
              double lQi_new = 0;
              for (int n = 0; n < B; n++) lQi_new += iweight * RHS(n, i, j, k, m) * iK1[l * B + n];
              const double dq = lQi_new - LQI(i, j, k, l, m);
              sq_res += dq * dq;
              LQI(i, j, k, l, m) = lQi_new;
            }
        }
    /* 4. exit condition (:407-423) */
    if (!s->use_max_picard || s->opt_iterations) {
      const double tol = 1e-7;
      if (sq_res < tol * tol) break;
    }
  }
  if (s->opt_iterations) return iter + 1 < maxIter ? iter + 1 : maxIter;  /* template :607 */
  return iter;
#undef LUH
#undef LQI
#undef LFI
#undef RHS
}

/* aderPredictorNonlinear (:441-515) and aderExtrapolatorNonlinear (:517-687). */
static void predictor_extrapolator3(const orc_solver* s, const double* lQi, const double* lFi,
                                    double* lQhi, double* lFhi, double* lQhbnd, double* lFhbnd) {
  const int B = s->b.order + 1, V = s->nvar, P = s->npar, ND = V + P, D = 3;
  const int B2 = B * B, B3 = B2 * B;
  const double* w = s->b.w;
  const double* F0 = s->b.F0;
  const double* F1 = s->b.F1;
  double* lFhi_x = lFhi;
  double* lFhi_y = lFhi + 1 * B3 * V;
  double* lFhi_z = lFhi + 2 * B3 * V;
  double* lShi = lFhi + 3 * B3 * V;
#define LQI(z, y, x, t, v) lQi[((((z) * B + (y)) * B + (x)) * B + (t)) * ND + (v)]
#define LFI(t, z, y, x, d, v) lFi[(((((t) * B + (z)) * B + (y)) * B + (x)) * (D + 1) + (d)) * V + (v)]
#define IQH(a, b, c, v) ((((a) * B + (b)) * B + (c)) * ND + (v))
#define IFH(a, b, c, v) ((((a) * B + (b)) * B + (c)) * V + (v))
#define IQB(f, a, b, v) ((((f) * B + (a)) * B + (b)) * ND + (v))
#define IFB(f, a, b, v) ((((f) * B + (a)) * B + (b)) * V + (v))
  memset(lQhi, 0, sizeof(double) * B3 * ND);
  memset(lFhi, 0, sizeof(double) * 4 * B3 * V);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        for (int l = 0; l < V; l++)
          for (int m = 0; m < B; m++) {
            lFhi_x[IFH(i, j, k, l)] += LFI(m, i, j, k, 0, l) * w[m];
            lFhi_y[IFH(i, k, j, l)] += LFI(m, i, j, k, 1, l) * w[m];
            lFhi_z[IFH(j, k, i, l)] += LFI(m, i, j, k, 2, l) * w[m];
          }
        for (int l = 0; l < ND; l++)
          for (int m = 0; m < B; m++) lQhi[IQH(i, j, k, l)] += LQI(i, j, k, m, l) * w[m];
        if (has_ncs(s))  /* lShi (:494-499) */
          for (int l = 0; l < V; l++)
            for (int m = 0; m < B; m++) lShi[IFH(i, j, k, l)] += LFI(m, i, j, k, 3, l) * w[m];
      }
  memset(lQhbnd, 0, sizeof(double) * 2 * D * B2 * ND);
  memset(lFhbnd, 0, sizeof(double) * 2 * D * B2 * V);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      for (int k = 0; k < V; k++)
        for (int l = 0; l < B; l++) {
          lFhbnd[IFB(0, i, j, k)] += lFhi_x[IFH(i, j, l, k)] * F0[l];
          lFhbnd[IFB(1, i, j, k)] += lFhi_x[IFH(i, j, l, k)] * F1[l];
        }
      for (int k = 0; k < ND; k++)
        for (int l = 0; l < B; l++) {
          lQhbnd[IQB(0, i, j, k)] += lQhi[IQH(i, j, l, k)] * F0[l];
          lQhbnd[IQB(1, i, j, k)] += lQhi[IQH(i, j, l, k)] * F1[l];
        }
    }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      for (int k = 0; k < V; k++)
        for (int l = 0; l < B; l++) {
          lFhbnd[IFB(2, i, j, k)] += lFhi_y[IFH(i, j, l, k)] * F0[l];
          lFhbnd[IFB(3, i, j, k)] += lFhi_y[IFH(i, j, l, k)] * F1[l];
        }
      for (int k = 0; k < ND; k++)
        for (int l = 0; l < B; l++) {
          lQhbnd[IQB(2, i, j, k)] += lQhi[IQH(i, l, j, k)] * F0[l];
          lQhbnd[IQB(3, i, j, k)] += lQhi[IQH(i, l, j, k)] * F1[l];
        }
    }
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++) {
      for (int k = 0; k < V; k++)
        for (int l = 0; l < B; l++) {
          lFhbnd[IFB(4, i, j, k)] += lFhi_z[IFH(i, j, l, k)] * F0[l];
          lFhbnd[IFB(5, i, j, k)] += lFhi_z[IFH(i, j, l, k)] * F1[l];
        }
      for (int k = 0; k < ND; k++)
        for (int l = 0; l < B; l++) {
          lQhbnd[IQB(4, i, j, k)] += lQhi[IQH(l, i, j, k)] * F0[l];
          lQhbnd[IQB(5, i, j, k)] += lQhi[IQH(l, i, j, k)] * F1[l];
        }
    }
#undef LQI
#undef LFI
#undef IQH
#undef IFH
#undef IQB
#undef IFB
}

/* ------------------------------------------------------------------------ */
/* Space-time predictor, 2D: gc2/spaceTimePredictorNonlinear.cpph             */
/* ------------------------------------------------------------------------ */

static int picard2(const orc_solver* s, const double* luh, const double* invDx, double dt,
                   double* lQi, double* rhs, double* lFi) {
  const int B = s->b.order + 1, V = s->nvar, P = s->npar, ND = V + P, D = 2;
  const double* w = s->b.w;
  const double* Kxi = s->b.Kxi;
  const double* iK1 = s->b.iK1;
  const double* F0 = s->b.F0;
#define LUH(y, x, v) luh[((y) * B + (x)) * ND + (v)]
#define LQI(y, x, t, v) lQi[(((y) * B + (x)) * B + (t)) * ND + (v)]
#define LFI(t, y, x, d, v) lFi[((((t) * B + (y)) * B + (x)) * (D + 1) + (d)) * V + (v)]
#define RHS(t, y, x, v) rhs[((((t) * B + (y)) * B + (x))) * V + (v)]
  /* initial guess (:61-69), or the optimised kernels' CERK guess */
  if (s->cerk_guess) {
    cerk_guess(s, luh, invDx, dt, lQi);
  } else {
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++)
          for (int m = 0; m < ND; m++) LQI(j, k, l, m) = LUH(j, k, m);
  }
  const int maxIter = picard_max_iterations(s);
  int iter = 0;
  double Fnode[2 * 64];
  for (; iter < maxIter; iter++) {
    for (int i = 0; i < B; i++) {
      /* fluxes (:134-150) */
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++) {
          orc_flux(s, &LQI(k, l, i, 0), Fnode);
          for (int d = 0; d < D; d++)
            for (int m = 0; m < V; m++) LFI(i, k, l, d, m) = Fnode[d * V + m];
          for (int m = 0; m < V; m++) LFI(i, k, l, 2, m) = 0.0;
        }
      /* rhs0 (:152-161) */
      for (int k = 0; k < B; k++)
        for (int l = 0; l < B; l++) {
          const double weight = w[k] * w[l];
          for (int m = 0; m < V; m++) RHS(i, k, l, m) = weight * F0[i] * LUH(k, l, m);
        }
      /* x direction (:163-178) */
      for (int k = 0; k < B; k++) {
        const double weight = w[i] * w[k];
        const double updateSize = weight * dt * invDx[0];
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++)
            for (int n = 0; n < B; n++)
              RHS(i, k, l, m) -= updateSize * LFI(i, k, n, 0, m) * Kxi[n * B + l];
      }
      /* y direction (:180-196) */
      for (int k = 0; k < B; k++) {
        const double weight = w[i] * w[k];
        const double updateSize = weight * dt * invDx[1];
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++)
            for (int n = 0; n < B; n++)
              RHS(i, l, k, m) -= updateSize * LFI(i, n, k, 1, m) * Kxi[n * B + l];
      }
      if (has_ncs(s)) {
        /* gradQCur of time node i (:99-135), then S = source - ncp (:211-243) */
        double gq[2 * 64], S[64];
        for (int k = 0; k < B; k++)
          for (int l = 0; l < B; l++) {
            for (int m = 0; m < V; m++) {
              double gx = 0.0, gy = 0.0;
              for (int n = 0; n < B; n++) {
                gx += 1.0 * invDx[0] * LQI(k, n, i, m) * s->b.dudx[l * B + n];
                gy += 1.0 * invDx[1] * LQI(n, l, i, m) * s->b.dudx[k * B + n];
              }
              gq[0 * V + m] = gx;
              gq[1 * V + m] = gy;
            }
            ncs_term(s, &LQI(k, l, i, 0), gq, S);
            const double weight = w[i] * w[k] * w[l];
            const double updateSize = weight * dt;
            for (int m = 0; m < V; m++) {
              LFI(i, k, l, 2, m) = S[m];
              RHS(i, k, l, m) += updateSize * S[m];
            }
          }
      }
    }
    /* iK1 solve and residual (:230-248) */
    double sq_res = 0.0;
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double weight = w[j] * w[k];
        const double iweight = 1.0 / weight;
        for (int l = 0; l < B; l++)
          for (int m = 0; m < V; m++) {
            double lQi_new = 0;
            for (int n = 0; n < B; n++) lQi_new += iweight * RHS(n, j, k, m) * iK1[l * B + n];
            const double dq = lQi_new - LQI(j, k, l, m);
            sq_res += dq * dq;
            LQI(j, k, l, m) = lQi_new;
          }
      }
    /* exit condition: the 2D kernel forces iter = MaxIterations + 1 (:296-306);
     * the optimised kernel breaks (template :307-309) */
    if (s->opt_iterations) {
      if (sq_res < 1e-7 * 1e-7) break;
    } else if (!s->use_max_picard) {
      const double tol = 1e-7;
      if (sq_res < tol * tol) iter = maxIter + 1;
    }
  }
  if (s->opt_iterations) return iter + 1 < maxIter ? iter + 1 : maxIter;  /* template :607 */
  return iter;
#undef LUH
#undef LQI
#undef LFI
#undef RHS
}

/* gc2 aderPredictorNonlinear (:332-380) and aderExtrapolatorNonlinear (:382-470).
 * Note: the reference extrapolates only the first NumberOfVariables entries
 * of lQhbnd on the x faces in 2D (parameters stay zero there), but all
 * NumberOfVariables+NumberOfParameters on the y faces.  Kept as is. */
static void predictor_extrapolator2(const orc_solver* s, const double* lQi, const double* lFi,
                                    double* lQhi, double* lFhi, double* lQhbnd, double* lFhbnd) {
  const int B = s->b.order + 1, V = s->nvar, P = s->npar, ND = V + P, D = 2;
  const int B2 = B * B;
  const double* w = s->b.w;
  const double* F0 = s->b.F0;
  const double* F1 = s->b.F1;
  double* lFhi_x = lFhi;
  double* lFhi_y = lFhi + 1 * B2 * V;
#define LQI(y, x, t, v) lQi[(((y) * B + (x)) * B + (t)) * ND + (v)]
#define LFI(t, y, x, d, v) lFi[((((t) * B + (y)) * B + (x)) * (D + 1) + (d)) * V + (v)]
#define IQH(a, b, v) (((a) * B + (b)) * ND + (v))
#define IFH(a, b, v) (((a) * B + (b)) * V + (v))
#define IQB(f, a, v) (((f) * B + (a)) * ND + (v))
#define IFB(f, a, v) (((f) * B + (a)) * V + (v))
  memset(lQhi, 0, sizeof(double) * B2 * ND);
  memset(lFhi, 0, sizeof(double) * 3 * B2 * V);
  for (int j = 0; j < B; j++)
    for (int k = 0; k < B; k++) {
      for (int l = 0; l < V; l++)
        for (int m = 0; m < B; m++) {
          lFhi_x[IFH(j, k, l)] += LFI(m, j, k, 0, l) * w[m];
          lFhi_y[IFH(k, j, l)] += LFI(m, j, k, 1, l) * w[m];
        }
      for (int l = 0; l < ND; l++)
        for (int m = 0; m < B; m++) lQhi[IQH(j, k, l)] += LQI(j, k, m, l) * w[m];
      if (has_ncs(s))  /* lShi (gc2 :370-374) */
        for (int l = 0; l < V; l++)
          for (int m = 0; m < B; m++) lFhi[2 * B2 * V + IFH(j, k, l)] += LFI(m, j, k, 2, l) * w[m];
    }
  memset(lQhbnd, 0, sizeof(double) * 2 * D * B * ND);
  memset(lFhbnd, 0, sizeof(double) * 2 * D * B * V);
  for (int j = 0; j < B; j++) {
    for (int k = 0; k < V; k++)
      for (int l = 0; l < B; l++) {
        lFhbnd[IFB(0, j, k)] += lFhi_x[IFH(j, l, k)] * F0[l];
        lFhbnd[IFB(1, j, k)] += lFhi_x[IFH(j, l, k)] * F1[l];
      }
    for (int k = 0; k < V; k++)
      for (int l = 0; l < B; l++) {
        lQhbnd[IQB(0, j, k)] += lQhi[IQH(j, l, k)] * F0[l];
        lQhbnd[IQB(1, j, k)] += lQhi[IQH(j, l, k)] * F1[l];
      }
  }
  for (int j = 0; j < B; j++) {
    for (int k = 0; k < V; k++)
      for (int l = 0; l < B; l++) {
        lFhbnd[IFB(2, j, k)] += lFhi_y[IFH(j, l, k)] * F0[l];
        lFhbnd[IFB(3, j, k)] += lFhi_y[IFH(j, l, k)] * F1[l];
      }
    for (int k = 0; k < ND; k++)
      for (int l = 0; l < B; l++) {
        lQhbnd[IQB(2, j, k)] += lQhi[IQH(l, j, k)] * F0[l];
        lQhbnd[IQB(3, j, k)] += lQhi[IQH(l, j, k)] * F1[l];
      }
  }
#undef LQI
#undef LFI
#undef IQH
#undef IFH
#undef IQB
#undef IFB
}

/* spaceTimePredictorNonlinear entry (gc3 :816-864, gc2 :598-645). */
int orc_stp_nonlinear(const orc_solver* s, const double* luh, const double* invDx, double dt,
                      double* lQi, double* rhs, double* lFi, double* lQhi, double* lFhi,
                      double* lQhbnd, double* lFhbnd) {
  if (s->dim == 3) {
    const int it = picard3(s, luh, invDx, dt, lQi, rhs, lFi);
    predictor_extrapolator3(s, lQi, lFi, lQhi, lFhi, lQhbnd, lFhbnd);
    return it;
  }
  const int it = picard2(s, luh, invDx, dt, lQi, rhs, lFi);
  predictor_extrapolator2(s, lQi, lFi, lQhi, lFhi, lQhbnd, lFhbnd);
  return it;
}

/* ------------------------------------------------------------------------ */
/* Volume integral (time-averaged branch): gc3/volumeIntegralNonlinear.cpph    */
/* :133-223, gc2/volumeIntegralNonlinear.cpph (same structure).                */
/* ------------------------------------------------------------------------ */
void orc_volume_integral_nonlinear(const orc_solver* s, double* lduh, const double* lFhi,
                                   const double* dx) {
  const int B = s->b.order + 1, V = s->nvar;
  const double* w = s->b.w;
  const double* Kxi = s->b.Kxi;
  if (s->dim == 3) {
    const int B3 = B * B * B;
    const double invDx[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
#define I4(a, b, c, v) ((((a) * B + (b)) * B + (c)) * V + (v))
    memset(lduh, 0, sizeof(double) * B3 * V);
    const double* fx = lFhi;
    const double* fy = lFhi + B3 * V;
    const double* fz = lFhi + 2 * B3 * V;
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double updateSize = w[i] * w[j] * invDx[0];
        for (int k = 0; k < B; k++)
          for (int l = 0; l < V; l++)
            for (int m = 0; m < B; m++)
              lduh[I4(i, j, k, l)] += Kxi[k * B + m] * fx[I4(i, j, m, l)] * updateSize;
      }
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double updateSize = w[i] * w[j] * invDx[1];
        for (int k = 0; k < B; k++)
          for (int l = 0; l < V; l++)
            for (int m = 0; m < B; m++)
              lduh[I4(i, k, j, l)] += Kxi[k * B + m] * fy[I4(i, j, m, l)] * updateSize;
      }
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double updateSize = w[i] * w[j] * invDx[2];
        for (int k = 0; k < B; k++)
          for (int l = 0; l < V; l++)
            for (int m = 0; m < B; m++)
              lduh[I4(k, i, j, l)] += Kxi[k * B + m] * fz[I4(i, j, m, l)] * updateSize;
      }
    if (has_ncs(s)) {  /* lduh += w lShi (:204-222) */
      const double* sh = lFhi + 3 * B3 * V;
      for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++)
          for (int k = 0; k < B; k++) {
            const double weight = w[i] * w[j] * w[k];
            for (int l = 0; l < V; l++) lduh[I4(i, j, k, l)] += weight * sh[I4(i, j, k, l)];
          }
    }
#undef I4
  } else {
    const int B2 = B * B;
    const double invDx[2] = {1.0 / dx[0], 1.0 / dx[1]};
#define I3(a, b, v) (((a) * B + (b)) * V + (v))
    memset(lduh, 0, sizeof(double) * B2 * V);
    const double* fx = lFhi;
    const double* fy = lFhi + B2 * V;
    for (int j = 0; j < B; j++) {
      const double updateSize = w[j] * invDx[0];
      for (int k = 0; k < B; k++)
        for (int l = 0; l < V; l++)
          for (int m = 0; m < B; m++) lduh[I3(j, k, l)] += Kxi[k * B + m] * fx[I3(j, m, l)] * updateSize;
    }
    for (int j = 0; j < B; j++) {
      const double updateSize = w[j] * invDx[1];
      for (int k = 0; k < B; k++)
        for (int l = 0; l < V; l++)
          for (int m = 0; m < B; m++) lduh[I3(k, j, l)] += Kxi[k * B + m] * fy[I3(j, m, l)] * updateSize;
    }
    if (has_ncs(s)) {  /* lduh += w lShi (gc2 volumeIntegralNonlinear) */
      const double* sh = lFhi + 2 * B2 * V;
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = w[j] * w[k];
          for (int l = 0; l < V; l++) lduh[I3(j, k, l)] += weight * sh[I3(j, k, l)];
        }
    }
#undef I3
  }
}

/* ------------------------------------------------------------------------ */
/* Rusanov: gc3/riemannSolverNonlinear.cpph:37-128 + averageRiemannInputs      */
/* (ExaHyPE/kernels/RiemannSolverUtils.h:152-177); 2D twin in gc2.             */
/* ------------------------------------------------------------------------ */
void orc_riemann_nonlinear(const orc_solver* s, double* FL, double* FR, const double* QL,
                           const double* QR, int direction) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar;
  const double* w = s->b.w;
  double QavL[64] = {0}, QavR[64] = {0}, LL[64], LR[64];
  const int nFace = (s->dim == 3) ? B * B : B;
  if (s->dim == 3) {
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double wij = w[i] * w[j];
        for (int k = 0; k < ND; k++) {
          QavL[k] += wij * QL[(i * B + j) * ND + k];
          QavR[k] += wij * QR[(i * B + j) * ND + k];
        }
      }
  } else {
    for (int j = 0; j < B; j++) {
      const double wj = w[j];
      for (int k = 0; k < ND; k++) {
        QavL[k] += wj * QL[j * ND + k];
        QavR[k] += wj * QR[j * ND + k];
      }
    }
  }
  orc_eigenvalues(s, QavL, direction, LL);
  orc_eigenvalues(s, QavR, direction, LR);
  double mL = fabs(LL[0]), mR = fabs(LR[0]);
  for (int k = 1; k < V; k++) {
    if (fabs(LL[k]) > mL) mL = fabs(LL[k]);
    if (fabs(LR[k]) > mR) mR = fabs(LR[k]);
  }
  const double smax = mL > mR ? mL : mR;
  for (int n = 0; n < nFace; n++) {
    double ncp[64] = {0};
    if (has_ncs(s)) {
      /* fluctuation (riemannSolverNonlinear.cpph:95-120): gradQ[direction] = QR - QL,
       * the NCP of the averaged state; FR = F - ncp/2, FL = F + ncp/2 */
      double gq[3 * 64] = {0}, Qavg[64];
      for (int l = 0; l < V; l++) gq[direction * V + l] = QR[n * ND + l] - QL[n * ND + l];
      for (int l = 0; l < ND; l++) Qavg[l] = 0.5 * (QR[n * ND + l] + QL[n * ND + l]);
      orc_ncp(s, Qavg, gq, ncp);
    }
    for (int k = 0; k < V; k++) {
      const double f = 0.5 * (FR[n * V + k] + FL[n * V + k]) - 0.5 * smax * (QR[n * ND + k] - QL[n * ND + k]);
      if (has_ncs(s)) {
        FR[n * V + k] = f - 0.5 * ncp[k];
        FL[n * V + k] = f + 0.5 * ncp[k];
      } else {
        FL[n * V + k] = f;
        FR[n * V + k] = f;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* faceIntegralNonlinear: gc3/faceIntegralNonlinear.cpph:21-103, gc2 twin.     */
/* ------------------------------------------------------------------------ */
void orc_face_integral_nonlinear(const orc_solver* s, double* lduh, const double* lFhbnd,
                                 int direction, int orientation, const double* dx) {
  const int B = s->b.order + 1, V = s->nvar;
  const double* w = s->b.w;
  const double* FC = orientation ? s->b.F1 : s->b.F0;
  const double scaling = 1.0 / dx[direction] * (2.0 * orientation - 1.0);
  if (s->dim == 3) {
#define I4(a, b, c, v) ((((a) * B + (b)) * B + (c)) * V + (v))
#define I3(a, b, v) (((a) * B + (b)) * V + (v))
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          if (direction == 0) {
            const double sw = w[i] * w[j] * scaling;
            for (int l = 0; l < V; l++) lduh[I4(i, j, k, l)] -= sw * FC[k] * lFhbnd[I3(i, j, l)];
          } else if (direction == 1) {
            const double sw = w[i] * w[k] * scaling;
            for (int l = 0; l < V; l++) lduh[I4(i, j, k, l)] -= sw * FC[j] * lFhbnd[I3(i, k, l)];
          } else {
            const double sw = w[j] * w[k] * scaling;
            for (int l = 0; l < V; l++) lduh[I4(i, j, k, l)] -= sw * FC[i] * lFhbnd[I3(j, k, l)];
          }
        }
#undef I4
#undef I3
  } else {
#define I3(a, b, v) (((a) * B + (b)) * V + (v))
    if (direction == 0) {
      for (int j = 0; j < B; j++) {
        const double sw = w[j] * scaling;
        for (int k = 0; k < B; k++)
          for (int l = 0; l < V; l++) lduh[I3(j, k, l)] -= sw * FC[k] * lFhbnd[j * V + l];
      }
    } else {
      for (int k = 0; k < B; k++) {
        const double sw = w[k] * scaling;
        for (int j = 0; j < B; j++)
          for (int l = 0; l < V; l++) lduh[I3(j, k, l)] -= sw * FC[j] * lFhbnd[k * V + l];
      }
    }
#undef I3
  }
}

/* surfaceIntegralNonlinear: gc3/surfaceIntegralNonlinear.cpph:21-90, gc2 twin. */
void orc_surface_integral_nonlinear(const orc_solver* s, double* lduh, const double* lFhbnd,
                                    const double* dx) {
  const int B = s->b.order + 1, V = s->nvar;
  const double* w = s->b.w;
  const double* F0 = s->b.F0;
  const double* F1 = s->b.F1;
  if (s->dim == 3) {
    const double invDx[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
#define I4(a, b, c, v) ((((a) * B + (b)) * B + (c)) * V + (v))
#define IFB(f, a, b, v) ((((f) * B + (a)) * B + (b)) * V + (v))
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double updateSize = w[i] * w[j] * invDx[0];
        for (int k = 0; k < B; k++)
          for (int l = 0; l < V; l++)
            lduh[I4(i, j, k, l)] -= (lFhbnd[IFB(1, i, j, l)] * F1[k] - lFhbnd[IFB(0, i, j, l)] * F0[k]) * updateSize;
      }
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double updateSize = w[i] * w[k] * invDx[1];
          for (int l = 0; l < V; l++)
            lduh[I4(i, j, k, l)] -= (lFhbnd[IFB(3, i, k, l)] * F1[j] - lFhbnd[IFB(2, i, k, l)] * F0[j]) * updateSize;
        }
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double updateSize = w[j] * w[k] * invDx[2];
          for (int l = 0; l < V; l++)
            lduh[I4(i, j, k, l)] -= (lFhbnd[IFB(5, j, k, l)] * F1[i] - lFhbnd[IFB(4, j, k, l)] * F0[i]) * updateSize;
        }
#undef I4
#undef IFB
  } else {
    const double invDx[2] = {1.0 / dx[0], 1.0 / dx[1]};
#define I3(a, b, v) (((a) * B + (b)) * V + (v))
#define IFB(f, a, v) (((f) * B + (a)) * V + (v))
    for (int j = 0; j < B; j++) {
      const double updateSize = w[j] * invDx[0];
      for (int k = 0; k < B; k++)
        for (int l = 0; l < V; l++)
          lduh[I3(j, k, l)] -= (lFhbnd[IFB(1, j, l)] * F1[k] - lFhbnd[IFB(0, j, l)] * F0[k]) * updateSize;
    }
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++) {
        const double updateSize = w[k] * invDx[1];
        for (int l = 0; l < V; l++)
          lduh[I3(j, k, l)] -= (lFhbnd[IFB(3, k, l)] * F1[j] - lFhbnd[IFB(2, k, l)] * F0[j]) * updateSize;
      }
#undef I3
#undef IFB
  }
}

/* solutionUpdate: gc3/solutionUpdate.cpph:28-67, gc2 twin. */
void orc_solution_update(const orc_solver* s, double* luh, const double* luhOld,
                         const double* lduh, double dt) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar;
  const double* w = s->b.w;
  if (s->dim == 3) {
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++)
        for (int k = 0; k < B; k++) {
          const double weight = w[i] * w[j] * w[k];
          const double updateSize = dt / weight;
          const int n = (i * B + j) * B + k;
          for (int l = 0; l < V; l++) luh[n * ND + l] = luhOld[n * ND + l] + lduh[n * V + l] * updateSize;
        }
  } else {
    for (int i = 0; i < B; i++)
      for (int j = 0; j < B; j++) {
        const double weight = w[i] * w[j];
        const double updateSize = dt / weight;
        const int n = i * B + j;
        for (int l = 0; l < V; l++) luh[n * ND + l] = luhOld[n * ND + l] + lduh[n * V + l] * updateSize;
      }
  }
}

/* stableTimeStepSize: gc3/stableTimeStepSize.cpph:14-62, gc2 twin. */
double orc_stable_dt(const orc_solver* s, const double* luh, const double* dx) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar, D = s->dim;
  const int nNodes = (D == 3) ? B * B * B : B * B;
  double invDx[3];
  for (int d = 0; d < D; d++) invDx[d] = 1.0 / dx[d];
  double dt = DBL_MAX;
  double lambda[64];
  for (int n = 0; n < nNodes; n++) {
    double denominator = 0.0;
    for (int d = 0; d < D; d++) {
      orc_eigenvalues(s, luh + n * ND, d, lambda);
      double maxEigenvalue = 0.0;
      for (int v = 0; v < V; v++) maxEigenvalue = fmax(fabs(lambda[v]), maxEigenvalue);
      denominator += maxEigenvalue * invDx[d];
    }
    const double c = s->cfl * s->pnpm / denominator;
    if (c < dt) dt = c;
  }
  return dt;
}

/* Glue fusedSpaceTimePredictorVolumeIntegral, non-linear branch with
 * addVolumeIntegralResultToUpdate == true:
 * Toolkit/exahype/toolkit/templates/solvers/AbstractGenericADERDGSolverImplementation.template:155-210. */
int orc_fused_stp_vi(const orc_solver* s, double* lduh, double* lQhbnd, double* lFhbnd,
                     const double* luh, const double* dx, double dt) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar, D = s->dim;
  const int sp = (D == 3) ? B * B * B : B * B;
  const int st = sp * B;
  double* lQi = (double*)malloc(sizeof(double) * st * ND);
  double* rhs = (double*)malloc(sizeof(double) * st * ND);
  double* lFi = (double*)malloc(sizeof(double) * st * (D + 1) * V);
  double* lQhi = (double*)malloc(sizeof(double) * sp * ND);
  double* lFhi = (double*)malloc(sizeof(double) * sp * (D + 1) * V);
  double invDx[3];
  for (int d = 0; d < D; d++) invDx[d] = 1.0 / dx[d];
  const int it = orc_stp_nonlinear(s, luh, invDx, dt, lQi, rhs, lFi, lQhi, lFhi, lQhbnd, lFhbnd);
  orc_volume_integral_nonlinear(s, lduh, lFhi, dx);
  free(lQi);
  free(rhs);
  free(lFi);
  free(lQhi);
  free(lFhi);
  return it;
}

/* ------------------------------------------------------------------------ */
/* Uniform periodic mesh, one global time step (order of ADERDGSolver.cpp:     */
/* prediction (:887-947), Riemann per face (:1764-1823), surface integral in   */
/* face order 0..2D-1 into the update (:1190-1239), update (:1241-1256), and   */
/* the new dt as the min of stableTimeStepSize (:1013-1049, :458-462)).        */
/* Periodic wrap is this driver's own neighbour rule (the reference has no     */
/* periodic boundary in the ADER-DG path).                                    */
/* ------------------------------------------------------------------------ */
int orc_run_uniform(const orc_solver* s, double* luh, int n, double h, int steps, double* dt_io,
                    int* picard_out, int threads) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar, D = s->dim;
  const int B1 = (D == 3) ? B * B : B;      /* nodes per face */
  const int sp = B1 * B;                    /* nodes per cell */
  const long ncell = (D == 3) ? (long)n * n * n : (long)n * n;
  const double dx[3] = {h, h, h};
  double* lduh = (double*)malloc(sizeof(double) * ncell * sp * V);
  double* lQhbnd = (double*)malloc(sizeof(double) * ncell * 2 * D * B1 * ND);
  double* lFhbnd = (double*)malloc(sizeof(double) * ncell * 2 * D * B1 * V);
  if (!lduh || !lQhbnd || !lFhbnd) return -1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  double dt = *dt_io;
  if (dt <= 0) {
    dt = DBL_MAX;
    for (long c = 0; c < ncell; c++) {
      const double d = orc_stable_dt(s, luh + c * sp * ND, dx);
      if (d < dt) dt = d;
    }
  }
  for (int step = 0; step < steps; step++) {
    /* 1. predictor + volume integral */
#pragma omp parallel for schedule(dynamic, 16)
    for (long c = 0; c < ncell; c++) {
      const int it = orc_fused_stp_vi(s, lduh + c * sp * V, lQhbnd + c * 2 * D * B1 * ND,
                                      lFhbnd + c * 2 * D * B1 * V, luh + c * sp * ND, dx, dt);
      if (picard_out) picard_out[c] = it;
    }
    /* 2. Riemann on every face: cell c is the left cell, its +d neighbour the right */
    for (int d = 0; d < D; d++) {
#pragma omp parallel for schedule(static)
      for (long c = 0; c < ncell; c++) {
        long ix = c % n, iy = (c / n) % n, iz = (D == 3) ? c / ((long)n * n) : 0;
        if (d == 0) ix = (ix + 1) % n;
        if (d == 1) iy = (iy + 1) % n;
        if (d == 2) iz = (iz + 1) % n;
        const long r = (D == 3) ? (iz * n + iy) * n + ix : iy * n + ix;
        double* FL = lFhbnd + (c * 2 * D + 2 * d + 1) * B1 * V;
        double* FR = lFhbnd + (r * 2 * D + 2 * d) * B1 * V;
        const double* QL = lQhbnd + (c * 2 * D + 2 * d + 1) * B1 * ND;
        const double* QR = lQhbnd + (r * 2 * D + 2 * d) * B1 * ND;
        orc_riemann_nonlinear(s, FL, FR, QL, QR, d);
      }
    }
    /* 3. face integrals into the update, then the solution update */
    double dt_new = DBL_MAX;
#pragma omp parallel for schedule(static) reduction(min : dt_new)
    for (long c = 0; c < ncell; c++) {
      double* du = lduh + c * sp * V;
      for (int d = 0; d < D; d++)
        for (int o = 0; o < 2; o++)
          orc_face_integral_nonlinear(s, du, lFhbnd + (c * 2 * D + 2 * d + o) * B1 * V, d, o, dx);
      orc_solution_update(s, luh + c * sp * ND, luh + c * sp * ND, du, dt);
      const double dc = orc_stable_dt(s, luh + c * sp * ND, dx);
      if (dc < dt_new) dt_new = dc;
    }
    dt = dt_new;
  }
  *dt_io = dt;
  free(lduh);
  free(lQhbnd);
  free(lFhbnd);
  return 0;
}

/*
 * ADER-DG oracle, boundary conditions -- plain-C FP64 restatement of the
 * reference's boundary path.  TEST INFRASTRUCTURE ONLY (see aderdg_oracle.h).
 *
 *   kernel   ExaHyPE/kernels/aderdg/generic/c/3d/boundaryConditions.cpph:29-122
 *            (2D twin generic/c/2d/boundaryConditions.cpph:28-100): the user's
 *            boundaryValues hook at every face node, node coordinates
 *            x[normal] = centre + (orientation - 1/2) dx, x[d2] from the slow
 *            face index, x[d1] from the fast one.
 *   glue     Toolkit/exahype/toolkit/templates/solvers/
 *            AbstractGenericADERDGSolverImplementation.template:307-350: the
 *            outside state/flux go to the left of the Riemann problem on a
 *            left face (orientation 0) and to the right on a right face; the
 *            Riemann result overwrites the cell's own flux.
 *   caller   ExaHyPE/exahype/solvers/ADERDGSolver.cpp:1836-1900
 *            (applyBoundaryConditions: t and dt of the cell's predictor).
 *   hooks    ApplicationExamples/Euler/Euler/EulerSolver_ADERDG.cpp:525-554
 *            (boundaryValues: reflecting wall; Dirichlet data = Gauss time
 *            average of the entropy wave) and :266-286 (entropyWave).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aderdg_oracle.h"

#define GAMMA 1.4

/* entropyWave (EulerSolver_ADERDG.cpp:266-286): density bump advected with
 * v0 = (1/2, 0, 0) at constant pressure 1; rho = 1/2 + 3/10 exp(-|x - x0 -
 * v0 t| / width^D), width = 3/10, x0 = (1/2, 1/2, 1/2). */
void orc_entropy_wave(const orc_solver* s, const double* x, double t, double* Q) {
  const int D = s->dim;
  const double width = 0.3, v0x = 0.5;
  double r2 = 0.0;
  for (int d = 0; d < D; d++) {
    const double c = x[d] - 0.5 - (d == 0 ? v0x * t : 0.0);
    r2 += c * c;
  }
  const double distance = sqrt(r2);
  Q[0] = 0.5 + 0.3 * exp(-distance / pow(width, D));
  Q[1] = Q[0] * v0x;
  Q[2] = 0.0;
  Q[3] = 0.0;
  Q[4] = 1.0 / (GAMMA - 1) + 0.5 * Q[0] * (v0x * v0x);
}

/* boundaryValues of the Euler example (EulerSolver_ADERDG.cpp:525-554) at one
 * face node.  Parameters (NumberOfParameters > 0) are not set by the hook; the
 * outside state keeps the inside ones. */
void orc_boundary_values(const orc_solver* s, int kind, const double* x, double t, double dt, int faceIndex,
                         int direction, const double* FIn, const double* QIn, double* FOut, double* QOut) {
  const int V = s->nvar, ND = V + s->npar;
  const int B = s->b.order + 1;
  double F[3 * 64];
  (void)faceIndex;
  (void)FIn;
  for (int k = V; k < ND; k++) QOut[k] = QIn[k];
  if (kind == ORC_BC_WALL) { /* :529-536 */
    for (int v = 0; v < V; v++) QOut[v] = QIn[v];
    QOut[1 + direction] = -QOut[1 + direction];
    orc_flux(s, QOut, F);
    for (int v = 0; v < V; v++) FOut[v] = F[direction * V + v];
  } else { /* ORC_BC_DIRICHLET_ENTROPY_WAVE, :537-554 */
    double Q[64];
    for (int v = 0; v < V; v++) QOut[v] = FOut[v] = 0.0;
    for (int i = 0; i < B; i++) {
      const double ti = t + dt * s->b.x[i];
      orc_entropy_wave(s, x, ti, Q);
      for (int k = V; k < ND; k++) Q[k] = QIn[k];
      orc_flux(s, Q, F);
      for (int v = 0; v < V; v++) {
        QOut[v] += Q[v] * s->b.w[i];
        FOut[v] += F[direction * V + v] * s->b.w[i];
      }
    }
  }
}

/* boundaryConditions kernel + glue for one boundary face of one cell:
 * FIn [nodes][V] (in: the cell's extrapolated flux; out: the Riemann flux),
 * QIn [nodes][ND]; nodes = B^(D-1), face-node order of the extrapolated arrays. */
void orc_boundary_conditions(const orc_solver* s, int kind, double* FIn, const double* QIn, const double* centre,
                             const double* dx, double t, double dt, int direction, int orientation) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar, D = s->dim;
  const int NF = (D == 3) ? B * B : B;
  const int faceIndex = 2 * direction + orientation;
  double* QOut = (double*)malloc(sizeof(double) * NF * ND);
  double* FOut = (double*)malloc(sizeof(double) * NF * V);
  /* free directions: d1 fast, d2 slow (boundaryConditions.cpph:49-70) */
  int d1 = 0, d2 = 0;
  if (D == 3) {
    d1 = (direction == 0) ? 1 : 0;
    d2 = (direction == 2) ? 1 : 2;
  } else {
    d1 = (direction == 0) ? 1 : 0;
  }
  double x[3];
  x[direction] = centre[direction] + (-0.5 + orientation) * dx[direction];
  if (D == 3) {
    for (int i = 0; i < B; i++) {
      x[d2] = centre[d2] + dx[d2] * (s->b.x[i] - 0.5);
      for (int j = 0; j < B; j++) {
        x[d1] = centre[d1] + dx[d1] * (s->b.x[j] - 0.5);
        const int n = i * B + j;
        orc_boundary_values(s, kind, x, t, dt, faceIndex, direction, FIn + n * V, QIn + n * ND, FOut + n * V,
                            QOut + n * ND);
      }
    }
  } else {
    for (int i = 0; i < B; i++) {
      x[d1] = centre[d1] + dx[d1] * (s->b.x[i] - 0.5);
      orc_boundary_values(s, kind, x, t, dt, faceIndex, direction, FIn + i * V, QIn + i * ND, FOut + i * V,
                          QOut + i * ND);
    }
  }
  if (orientation == 0) { /* outside on the left */
    orc_riemann_nonlinear(s, FOut, FIn, QOut, QIn, direction);
  } else {
    orc_riemann_nonlinear(s, FIn, FOut, QIn, QOut, direction);
  }
  free(QOut);
  free(FOut);
}

/* Uniform box of nx x ny (x nz) cells of size h with a boundary kind per
 * domain face (bc[2d+o], ORC_BC_*): one global step per iteration in the
 * order of orc_run_uniform, with the boundary faces of the non-periodic
 * directions solved by orc_boundary_conditions at the step's (t, dt).  A
 * direction is periodic iff both of its faces are ORC_BC_PERIODIC.  Cells x
 * fastest. */
int orc_run_box(const orc_solver* s, double* luh, const int* ncell_d, double h, int steps, double* dt_io,
                double* t_io, const int* bc, int threads) {
  const int B = s->b.order + 1, V = s->nvar, ND = V + s->npar, D = s->dim;
  const int B1 = (D == 3) ? B * B : B;
  const int sp = B1 * B;
  const long nx = ncell_d[0], ny = ncell_d[1], nz = (D == 3) ? ncell_d[2] : 1;
  const long nd[3] = {nx, ny, nz};
  const long ncell = nx * ny * nz;
  const double dx[3] = {h, h, h};
  for (int d = 0; d < D; d++)
    if ((bc[2 * d] == ORC_BC_PERIODIC) != (bc[2 * d + 1] == ORC_BC_PERIODIC)) return -2;
  double* lduh = (double*)malloc(sizeof(double) * ncell * sp * V);
  double* lQhbnd = (double*)malloc(sizeof(double) * ncell * 2 * D * B1 * ND);
  double* lFhbnd = (double*)malloc(sizeof(double) * ncell * 2 * D * B1 * V);
  if (!lduh || !lQhbnd || !lFhbnd) return -1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  double dt = *dt_io, t = *t_io;
  if (dt <= 0) {
    dt = DBL_MAX;
    for (long c = 0; c < ncell; c++) {
      const double d = orc_stable_dt(s, luh + c * sp * ND, dx);
      if (d < dt) dt = d;
    }
  }
  for (int step = 0; step < steps; step++) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long c = 0; c < ncell; c++)
      orc_fused_stp_vi(s, lduh + c * sp * V, lQhbnd + c * 2 * D * B1 * ND, lFhbnd + c * 2 * D * B1 * V,
                       luh + c * sp * ND, dx, dt);
    for (int d = 0; d < D; d++) {
      const int periodic = bc[2 * d] == ORC_BC_PERIODIC;
      const long n = nd[d];
#pragma omp parallel for schedule(static)
      for (long c = 0; c < ncell; c++) {
        long ix = c % nx, iy = (c / nx) % ny, iz = (D == 3) ? c / (nx * ny) : 0;
        const long id = (d == 0) ? ix : (d == 1 ? iy : iz);
        double centre[3] = {(ix + 0.5) * h, (iy + 0.5) * h, (iz + 0.5) * h};
        if (!periodic && id == 0) /* the cell's left face is a domain face */
          orc_boundary_conditions(s, bc[2 * d], lFhbnd + (c * 2 * D + 2 * d) * B1 * V,
                                  lQhbnd + (c * 2 * D + 2 * d) * B1 * ND, centre, dx, t, dt, d, 0);
        if (!periodic && id == n - 1) {
          orc_boundary_conditions(s, bc[2 * d + 1], lFhbnd + (c * 2 * D + 2 * d + 1) * B1 * V,
                                  lQhbnd + (c * 2 * D + 2 * d + 1) * B1 * ND, centre, dx, t, dt, d, 1);
          continue;
        }
        if (d == 0) ix = (ix + 1) % nx;
        if (d == 1) iy = (iy + 1) % ny;
        if (d == 2) iz = (iz + 1) % nz;
        const long r = (D == 3) ? (iz * ny + iy) * nx + ix : iy * nx + ix;
        orc_riemann_nonlinear(s, lFhbnd + (c * 2 * D + 2 * d + 1) * B1 * V, lFhbnd + (r * 2 * D + 2 * d) * B1 * V,
                              lQhbnd + (c * 2 * D + 2 * d + 1) * B1 * ND, lQhbnd + (r * 2 * D + 2 * d) * B1 * ND, d);
      }
    }
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
    t += dt;
    dt = dt_new;
  }
  *dt_io = dt;
  *t_io = t;
  free(lduh);
  free(lQhbnd);
  free(lFhbnd);
  return 0;
}

/* The n^D cube of orc_run_box. */
int orc_run_bc(const orc_solver* s, double* luh, int n, double h, int steps, double* dt_io, double* t_io,
               const int* bc, int threads) {
  const int nd[3] = {n, n, n};
  return orc_run_box(s, luh, nd, h, steps, dt_io, t_io, bc, threads);
}

/*
 * AMR oracle -- TEST INFRASTRUCTURE ONLY (see aderdg_oracle.h).
 *
 * Restates the reference's generic AMR projections
 * (ExaHyPE/kernels/aderdg/generic/c/{2d,3d}/amrRoutines.cpph) for any level
 * difference:
 *   orc_face_prolongation     faceUnknownsProlongation      3d :23-143   (overwrites the fine face)
 *   orc_face_restriction      faceUnknownsRestriction (Q,F) 3d :248-314  (accumulates into the coarse face)
 *   orc_face_flux_restriction faceUnknownsRestriction (F)   3d :178-246  (overwrites; used by faceIntegral
 *                                                                        with levelDelta > 0)
 *   orc_volume_prolongation   volumeUnknownsProlongation    3d :337-419  (overwrites)
 *   orc_volume_restriction    volumeUnknownsRestriction     3d :483-554  (accumulates)
 * Single-level operators: prolongation F[m] = sum_n C[n] prod_d fgp[s_d][n_d][m_d];
 * restriction C[m] += sum_n prod_d (w[n_d] fgp[s_d][m_d][n_d] / w[m_d]) F[n] / 3^(D-1)
 * (faces) or / 3^D (volumes).
 *
 * And a static two-level 3D mesh step (orc_amr_run) with the reference's
 * coarse-fine coupling under global time stepping: the coarse leaf's face
 * data are prolongated to virtual fine cells (ADERDGSolver.cpp:1340-1399,
 * prolongateFaceDataToVirtualCell), the Riemann problem is solved on the fine
 * subfaces, and the virtual cells' fluctuations are restricted back and
 * face-integrated with the coarse cell size (restrictToTopMostParent,
 * ADERDGSolver.cpp:1472-1563, and the faceIntegral glue with levelDelta > 0,
 * AbstractGenericADERDGSolverImplementation.template:247-285).  The coarse
 * leaf's own face integral on a face adjacent to a refined cell is skipped
 * (surfaceIntegral, ADERDGSolver.cpp:1190-1239: faces whose neighbour is an
 * ancestor).  Periodic unit cube; leaves ordered by coarse cell (x fastest),
 * a refined coarse cell contributing its 27 children (x fastest).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "aderdg_oracle.h"

static int ipow3(int e) {
  int r = 1;
  while (e-- > 0) r *= 3;
  return r;
}

/* ---- single-level operators on K = D-1 (faces) or D (volumes) axes -------- */
static void single_prolong(const orc_solver* s, int K, int nv, double* fine, const double* coarse, const int* sub) {
  const int B = s->b.order + 1;
  const double* fgp = s->b.fgp; /* fgp[(sub*B + n)*B + m] */
  int Bk = 1;
  for (int k = 0; k < K; k++) Bk *= B;
  for (int m = 0; m < Bk; m++)
    for (int v = 0; v < nv; v++) {
      double acc = 0.0;
      for (int n = 0; n < Bk; n++) {
        double f = 1.0;
        int mm = m, nn = n;
        for (int k = 0; k < K; k++) {
          f *= fgp[(sub[k] * B + nn % B) * B + mm % B];
          mm /= B;
          nn /= B;
        }
        acc += coarse[n * nv + v] * f;
      }
      fine[m * nv + v] += acc;
    }
}

static void single_restrict(const orc_solver* s, int K, int nv, double* coarse, const double* fine, const int* sub) {
  const int B = s->b.order + 1;
  const double* fgp = s->b.fgp;
  const double* w = s->b.w;
  int Bk = 1;
  for (int k = 0; k < K; k++) Bk *= B;
  const double scale = (double)ipow3(K);
  for (int m = 0; m < Bk; m++)
    for (int v = 0; v < nv; v++) {
      double acc = 0.0;
      for (int n = 0; n < Bk; n++) {
        double f = 1.0;
        int mm = m, nn = n;
        for (int k = 0; k < K; k++) {
          f *= w[nn % B] * fgp[(sub[k] * B + mm % B) * B + nn % B] / w[mm % B];
          mm /= B;
          nn /= B;
        }
        acc += f * fine[n * nv + v];
      }
      coarse[m * nv + v] += acc / scale;
    }
}

/* prolongation over levelDelta levels: digits of the subindex from the most
 * significant (3^(delta-1)) down (amrRoutines.cpph:69-90) */
static void prolong_levels(const orc_solver* s, int K, int nv, double* fine, const double* coarse, int delta,
                           const int* subidx) {
  const int B = s->b.order + 1;
  int Bk = 1;
  for (int k = 0; k < K; k++) Bk *= B;
  double* a = (double*)malloc(sizeof(double) * Bk * nv);
  double* b = (double*)malloc(sizeof(double) * Bk * nv);
  memcpy(a, coarse, sizeof(double) * Bk * nv);
  int prev[3] = {subidx[0], K > 1 ? subidx[1] : 0, K > 2 ? subidx[2] : 0};
  for (int l = 1; l <= delta; l++) {
    const int sig = ipow3(delta - l);
    int cur[3], si[3];
    for (int k = 0; k < K; k++) {
      cur[k] = prev[k] % sig;
      si[k] = (prev[k] - cur[k]) / sig;
    }
    memset(b, 0, sizeof(double) * Bk * nv);
    single_prolong(s, K, nv, b, a, si);
    for (int k = 0; k < K; k++) prev[k] = cur[k];
    double* t = a;
    a = b;
    b = t;
  }
  memcpy(fine, a, sizeof(double) * Bk * nv);
  free(a);
  free(b);
}

/* restriction over levelDelta levels: digits from the least significant up
 * (amrRoutines.cpph:196-224); result returned in out (overwritten) */
static void restrict_levels(const orc_solver* s, int K, int nv, double* out, const double* fine, int delta,
                            const int* subidx) {
  const int B = s->b.order + 1;
  int Bk = 1;
  for (int k = 0; k < K; k++) Bk *= B;
  double* a = (double*)malloc(sizeof(double) * Bk * nv);
  double* b = (double*)malloc(sizeof(double) * Bk * nv);
  memcpy(a, fine, sizeof(double) * Bk * nv);
  int cur[3] = {subidx[0], K > 1 ? subidx[1] : 0, K > 2 ? subidx[2] : 0};
  for (int l = 1; l <= delta; l++) {
    int si[3];
    for (int k = 0; k < K; k++) {
      si[k] = cur[k] % 3;
      cur[k] = (cur[k] - si[k]) / 3;
    }
    memset(b, 0, sizeof(double) * Bk * nv);
    single_restrict(s, K, nv, b, a, si);
    double* t = a;
    a = b;
    b = t;
  }
  memcpy(out, a, sizeof(double) * Bk * nv);
  free(a);
  free(b);
}

static int nbf(const orc_solver* s) { /* nodes per face */
  const int B = s->b.order + 1;
  return s->dim == 3 ? B * B : B;
}
static int nbv(const orc_solver* s) { /* nodes per cell */
  const int B = s->b.order + 1;
  return s->dim == 3 ? B * B * B : B * B;
}

void orc_face_prolongation(const orc_solver* s, double* lQf, double* lFf, const double* lQc, const double* lFc,
                           int level_coarse, int level_fine, const int* subface) {
  const int ND = s->nvar + s->npar, K = s->dim - 1, delta = level_fine - level_coarse;
  prolong_levels(s, K, ND, lQf, lQc, delta, subface);
  prolong_levels(s, K, s->nvar, lFf, lFc, delta, subface);
}

void orc_face_restriction(const orc_solver* s, double* lQc, double* lFc, const double* lQf, const double* lFf,
                          int level_coarse, int level_fine, const int* subface) {
  const int ND = s->nvar + s->npar, K = s->dim - 1, delta = level_fine - level_coarse, NF = nbf(s);
  double* tq = (double*)malloc(sizeof(double) * NF * ND);
  double* tf = (double*)malloc(sizeof(double) * NF * s->nvar);
  restrict_levels(s, K, ND, tq, lQf, delta, subface);
  restrict_levels(s, K, s->nvar, tf, lFf, delta, subface);
  for (int i = 0; i < NF * ND; i++) lQc[i] += tq[i];
  for (int i = 0; i < NF * s->nvar; i++) lFc[i] += tf[i];
  free(tq);
  free(tf);
}

void orc_face_flux_restriction(const orc_solver* s, double* lFc, const double* lFf, const int* subface,
                               int level_delta) {
  restrict_levels(s, s->dim - 1, s->nvar, lFc, lFf, level_delta, subface);
}

void orc_volume_prolongation(const orc_solver* s, double* luhF, const double* luhC, int level_coarse,
                             int level_fine, const int* subcell) {
  prolong_levels(s, s->dim, s->nvar + s->npar, luhF, luhC, level_fine - level_coarse, subcell);
}

void orc_volume_restriction(const orc_solver* s, double* luhC, const double* luhF, int level_coarse,
                            int level_fine, const int* subcell) {
  const int ND = s->nvar + s->npar, NS = nbv(s);
  double* t = (double*)malloc(sizeof(double) * NS * ND);
  restrict_levels(s, s->dim, ND, t, luhF, level_fine - level_coarse, subcell);
  for (int i = 0; i < NS * ND; i++) luhC[i] += t[i];
  free(t);
}

/* ---------------------------------------------------------------------------- */
/* static two-level 3D mesh                                                     */
/* ---------------------------------------------------------------------------- */
typedef struct {
  int level;   /* 0 coarse, 1 fine */
  int c[3];    /* integer coordinates on its level */
} amr_leaf;

typedef struct {
  int n;
  const unsigned char* refined;
  int nleaves;
  amr_leaf* leaves;
  int* coarse_leaf; /* [n^3] leaf index of an unrefined coarse cell, -1 if refined */
  int* fine_leaf;   /* [(3n)^3] leaf index of a fine cell, -1 if its coarse cell is unrefined */
} amr_mesh;

static int amr_build(amr_mesh* m, int n, const unsigned char* refined) {
  m->n = n;
  m->refined = refined;
  const long nc = (long)n * n * n, N3 = 3L * n, nf = N3 * N3 * N3;
  m->coarse_leaf = (int*)malloc(sizeof(int) * nc);
  m->fine_leaf = (int*)malloc(sizeof(int) * nf);
  for (long i = 0; i < nf; i++) m->fine_leaf[i] = -1;
  int cnt = 0;
  for (long c = 0; c < nc; c++) cnt += refined[c] ? 27 : 1;
  m->nleaves = cnt;
  m->leaves = (amr_leaf*)malloc(sizeof(amr_leaf) * cnt);
  int li = 0;
  for (long c = 0; c < nc; c++) {
    const int ix = (int)(c % n), iy = (int)((c / n) % n), iz = (int)(c / ((long)n * n));
    if (!refined[c]) {
      m->coarse_leaf[c] = li;
      m->leaves[li].level = 0;
      m->leaves[li].c[0] = ix;
      m->leaves[li].c[1] = iy;
      m->leaves[li].c[2] = iz;
      li++;
    } else {
      m->coarse_leaf[c] = -1;
      for (int sz = 0; sz < 3; sz++)
        for (int sy = 0; sy < 3; sy++)
          for (int sx = 0; sx < 3; sx++) {
            const int X = 3 * ix + sx, Y = 3 * iy + sy, Z = 3 * iz + sz;
            m->fine_leaf[((long)Z * N3 + Y) * N3 + X] = li;
            m->leaves[li].level = 1;
            m->leaves[li].c[0] = X;
            m->leaves[li].c[1] = Y;
            m->leaves[li].c[2] = Z;
            li++;
          }
    }
  }
  return 0;
}

static void amr_free(amr_mesh* m) {
  free(m->coarse_leaf);
  free(m->fine_leaf);
  free(m->leaves);
}

int orc_amr_nleaves(int n, const unsigned char* refined) {
  long cnt = 0;
  for (long c = 0; c < (long)n * n * n; c++) cnt += refined[c] ? 27 : 1;
  return (int)cnt;
}

/* leaf levels and coordinates (for initial conditions in tests) */
void orc_amr_leaves(int n, const unsigned char* refined, int* level, int* coords) {
  amr_mesh m;
  amr_build(&m, n, refined);
  for (int i = 0; i < m.nleaves; i++) {
    level[i] = m.leaves[i].level;
    for (int d = 0; d < 3; d++) coords[3 * i + d] = m.leaves[i].c[d];
  }
  amr_free(&m);
}

/* One global step; dt_io in: dt (<= 0: from the state), out: dt of the next step */
int orc_amr_run(const orc_solver* s, int n, const unsigned char* refined, double* luh, int steps, double* dt_io,
                int threads) {
  if (s->dim != 3) return -2;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  amr_mesh m;
  amr_build(&m, n, refined);
  const int B = s->b.order + 1, NS = B * B * B, NF = B * B, ND = s->nvar + s->npar, V = s->nvar;
  const int L = m.nleaves;
  const double hc = 1.0 / n, hf = hc / 3.0;
  const long N3 = 3L * n;
  double* lduh = (double*)malloc(sizeof(double) * L * NS * V);
  double* lQb = (double*)malloc(sizeof(double) * L * 6 * NF * ND);
  double* lFb = (double*)malloc(sizeof(double) * L * 6 * NF * V);
  double* vQ = (double*)malloc(sizeof(double) * NF * ND);
  double* vF = (double*)malloc(sizeof(double) * NF * V);
  double* rF = (double*)malloc(sizeof(double) * NF * V);
  double dt = *dt_io;
  if (!(dt > 0)) {
    dt = DBL_MAX;
    for (int i = 0; i < L; i++) {
      const double hh = m.leaves[i].level ? hf : hc;
      const double dxv[3] = {hh, hh, hh};
      const double d = orc_stable_dt(s, luh + (size_t)i * NS * ND, dxv);
      if (d < dt) dt = d;
    }
  }
  for (int step = 0; step < steps; step++) {
    /* 1. predictor + volume integral of every leaf */
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < L; i++) {
      const double hh = m.leaves[i].level ? hf : hc;
      const double dxv[3] = {hh, hh, hh};
      orc_fused_stp_vi(s, lduh + (size_t)i * NS * V, lQb + (size_t)i * 6 * NF * ND, lFb + (size_t)i * 6 * NF * V,
                       luh + (size_t)i * NS * ND, dxv, dt);
    }
    /* 2. Riemann: same-level faces (left leaf's + side) and coarse-fine subfaces */
    for (int i = 0; i < L; i++) {
      const amr_leaf* a = &m.leaves[i];
      for (int d = 0; d < 3; d++) {
        int c[3] = {a->c[0], a->c[1], a->c[2]};
        if (a->level == 0) {
          c[d] = (c[d] + 1) % n;
          const long cc = ((long)c[2] * n + c[1]) * n + c[0];
          const int j = m.coarse_leaf[cc];
          if (j >= 0)
            orc_riemann_nonlinear(s, lFb + ((size_t)i * 6 + 2 * d + 1) * NF * V, lFb + ((size_t)j * 6 + 2 * d) * NF * V,
                                  lQb + ((size_t)i * 6 + 2 * d + 1) * NF * ND, lQb + ((size_t)j * 6 + 2 * d) * NF * ND, d);
        } else {
          c[d] = (int)((c[d] + 1) % N3);
          const int j = m.fine_leaf[((long)c[2] * N3 + c[1]) * N3 + c[0]];
          if (j >= 0)
            orc_riemann_nonlinear(s, lFb + ((size_t)i * 6 + 2 * d + 1) * NF * V, lFb + ((size_t)j * 6 + 2 * d) * NF * V,
                                  lQb + ((size_t)i * 6 + 2 * d + 1) * NF * ND, lQb + ((size_t)j * 6 + 2 * d) * NF * ND, d);
        }
      }
    }
    /* coarse leaves next to refined cells: virtual subfaces */
    for (int i = 0; i < L; i++) {
      const amr_leaf* a = &m.leaves[i];
      if (a->level != 0) continue;
      double* lduhC = lduh + (size_t)i * NS * V;
      for (int f = 0; f < 6; f++) {
        const int d = f / 2, o = f % 2;
        int c[3] = {a->c[0], a->c[1], a->c[2]};
        c[d] = (c[d] + (o ? 1 : n - 1)) % n;
        const long cc = ((long)c[2] * n + c[1]) * n + c[0];
        if (!refined[cc]) continue;
        const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2; /* transverse axes, lower first */
        for (int s1 = 0; s1 < 3; s1++)
          for (int s0 = 0; s0 < 3; s0++) {
            const int sub[2] = {s0, s1};
            orc_face_prolongation(s, vQ, vF, lQb + ((size_t)i * 6 + f) * NF * ND, lFb + ((size_t)i * 6 + f) * NF * V, 0,
                                  1, sub);
            /* the fine leaf across the subface */
            int X[3];
            X[d] = o ? 3 * c[d] : 3 * c[d] + 2;
            X[t0] = 3 * a->c[t0] + s0;
            X[t1] = 3 * a->c[t1] + s1;
            const int j = m.fine_leaf[((long)X[2] * N3 + X[1]) * N3 + X[0]];
            const int fo = 2 * d + (1 - o); /* the fine leaf's face */
            if (o)
              orc_riemann_nonlinear(s, vF, lFb + ((size_t)j * 6 + fo) * NF * V, vQ, lQb + ((size_t)j * 6 + fo) * NF * ND, d);
            else
              orc_riemann_nonlinear(s, lFb + ((size_t)j * 6 + fo) * NF * V, vF, lQb + ((size_t)j * 6 + fo) * NF * ND, vQ, d);
            /* restrictToTopMostParent: restricted flux, face integral with the coarse size */
            orc_face_flux_restriction(s, rF, vF, sub, 1);
            const double dxc[3] = {hc, hc, hc};
            orc_face_integral_nonlinear(s, lduhC, rF, d, o, dxc);
          }
      }
    }
    /* 3. surface integrals (skipping coarse faces next to refined cells) and update */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L; i++) {
      const amr_leaf* a = &m.leaves[i];
      const double hh = a->level ? hf : hc;
      const double dxv[3] = {hh, hh, hh};
      for (int f = 0; f < 6; f++) {
        const int d = f / 2, o = f % 2;
        if (a->level == 0) {
          int c[3] = {a->c[0], a->c[1], a->c[2]};
          c[d] = (c[d] + (o ? 1 : n - 1)) % n;
          if (refined[((long)c[2] * n + c[1]) * n + c[0]]) continue;
        }
        orc_face_integral_nonlinear(s, lduh + (size_t)i * NS * V, lFb + ((size_t)i * 6 + f) * NF * V, d, o, dxv);
      }
      orc_solution_update(s, luh + (size_t)i * NS * ND, luh + (size_t)i * NS * ND, lduh + (size_t)i * NS * V, dt);
    }
    /* 4. next dt */
    double dtn = DBL_MAX;
    for (int i = 0; i < L; i++) {
      const double hh = m.leaves[i].level ? hf : hc;
      const double dxv[3] = {hh, hh, hh};
      const double dd = orc_stable_dt(s, luh + (size_t)i * NS * ND, dxv);
      if (dd < dtn) dtn = dd;
    }
    dt = dtn;
  }
  *dt_io = dt;
  free(lduh);
  free(lQb);
  free(lFb);
  free(vQ);
  free(vF);
  free(rF);
  amr_free(&m);
  return 0;
}

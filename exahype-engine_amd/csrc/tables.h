// Operator tables of the nodal ADER-DG scheme on Gauss-Legendre points,
// derived on the host from closed forms in long double and rounded once to
// double.  Index conventions are those of the reference tables
// (ExaHyPE/kernels/GaussLegendreBasis.h; generator Miscellaneous/aderdg/aderdg.py:14-227):
//   Kxi[k][l]   = w_l * phi_k'(xi_l)                 (stiffness, aderdg.py:53-106)
//   dudx[i][j]  = phi_j'(xi_i)                       (= MM^-1 Kxi^T, aderdg.py:177-196)
//   iK1         = (phi(1) phi(1)^T - Kxi)^-1         (aderdg.py:134-160)
//   F0[k]       = phi_k(0),  F1[k] = phi_k(1)        (FCoeff, aderdg.py:169-172)
//   fgp[j][m][n]= phi_m((j + xi_n) / 3)              (fineGridProjector, aderdg.py:197-227)
//   egp[m][n]   = phi_m(n / N)                       (equidistantGridProjector)
#pragma once

namespace exa {

constexpr int kMaxBasis = 16;

struct Tables {
  int order = 0;
  int B = 1;
  double w[kMaxBasis];
  double x[kMaxBasis];
  double Kxi[kMaxBasis * kMaxBasis];
  double iK1[kMaxBasis * kMaxBasis];
  double dudx[kMaxBasis * kMaxBasis];
  double F0[kMaxBasis];
  double F1[kMaxBasis];
  double fgp[3 * kMaxBasis * kMaxBasis];
  double egp[kMaxBasis * kMaxBasis];
};

// Returns false for unsupported orders (order < 0 or order + 1 > kMaxBasis).
bool build_tables(int order, Tables* out);

// PNPM factor of the reference spec defaults for a given order
// (Toolkit/exahype/toolkit/solverController.py:17-23).
double pnpm_for_order(int order);

}  // namespace exa

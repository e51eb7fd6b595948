/*
 * Device PDE hooks of the nonlinear ADER-DG path: the compile-time plug-in for
 * a user PDE (pde id EXA_PDE_USER in exahype_amd.h).
 *
 * Reference interface replaced: the user hooks of a generated ADER-DG solver,
 *   flux, eigenvalues, nonConservativeProduct, algebraicSource
 *   (Toolkit/exahype/toolkit/templates/solvers/AbstractGenericADERDGSolverHeader.template:220,311-313,325;
 *    called from ExaHyPE/kernels/aderdg/generic/c/3d/spaceTimePredictorNonlinear.cpph:196-227,
 *    :323-367 and riemannSolverNonlinear.cpph:61-120).
 * Device code cannot call host virtual functions, so the hooks are static
 * device functions of one struct, ExaUserPde, compiled into the library:
 * build the library with EXA_USER_PDE_HEADER="path/to/my_pde.h" (make
 * USER_PDE=path/to/my_pde.h) where that header defines struct ExaUserPde with
 * the members below; without it the test PDE defined here is compiled.
 *
 * Contract (N = kNVar conserved variables, 5 when the struct does not
 * declare kNVar; no material parameters; R is double in the FP64 form):
 *   static constexpr int kNVar;     // optional: the number of variables N
 *                                   // (compiled for orders 3 and 4; 5 variables
 *                                   // for every compiled order)
 *   static constexpr bool kNcp;     // nonConservativeProduct is used
 *   static constexpr bool kSource;  // algebraicSource is used
 *   template <int D, typename R> static __device__ void
 *       flux(const double* prm, const R* Q, R (&F)[D][N]);
 *   template <int D> static __device__ double
 *       max_abs_eigenvalue(const double* prm, const double* Q, int d);
 *       // max_v |lambda_v(Q, d)| (the reference's eigenvalues followed by the
 *       // max of absolute values, riemannSolverNonlinear.cpph:61-70,
 *       // stableTimeStepSize.cpph:14-62)
 *   template <int D, typename R> static __device__ void
 *       ncp(const double* prm, const R* Q, const R (&gradQ)[D][N], R* BgradQ);
 *   template <typename R> static __device__ void
 *       source(const double* prm, const R* Q, R* S);
 *       // algebraicSource without x and t (the device mesh keeps no cell
 *       // centres on the predictor path)
 * prm: the 16 user parameters set with exa_ctx_set_pde_params.
 *
 * This header declares no extern "C" entry points.
 */
#ifndef EXAHYPE_AMD_PDE_H
#define EXAHYPE_AMD_PDE_H

#ifdef EXA_USER_PDE_HEADER
#include EXA_USER_PDE_HEADER
#else

/* Test PDE: dQ/dt + div(a Q) + b . grad Q = -kappa Q for each variable, with
 * prm = {a_x, a_y, a_z, b_x, b_y, b_z, kappa}.  Exact periodic solution
 * Q0(x - (a + b) t) exp(-kappa t). */
struct ExaUserPde {
  static constexpr bool kNcp = true;
  static constexpr bool kSource = true;

  template <int D, typename R>
  static __device__ __forceinline__ void flux(const double* prm, const R* Q, R (&F)[D][5]) {
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < 5; v++) F[d][v] = static_cast<R>(prm[d]) * Q[v];
  }

  template <int D>
  static __device__ __forceinline__ double max_abs_eigenvalue(const double* prm, const double* Q, int d) {
    (void)Q;
    return fabs(prm[d] + prm[3 + d]);
  }

  template <int D, typename R>
  static __device__ __forceinline__ void ncp(const double* prm, const R* Q, const R (&gradQ)[D][5], R* BgradQ) {
    (void)Q;
#pragma unroll
    for (int v = 0; v < 5; v++) {
      R a = R(0);
#pragma unroll
      for (int d = 0; d < D; d++) a += static_cast<R>(prm[3 + d]) * gradQ[d][v];
      BgradQ[v] = a;
    }
  }

  template <typename R>
  static __device__ __forceinline__ void source(const double* prm, const R* Q, R* S) {
#pragma unroll
    for (int v = 0; v < 5; v++) S[v] = -static_cast<R>(prm[6]) * Q[v];
  }
};

#endif /* EXA_USER_PDE_HEADER */
#endif /* EXAHYPE_AMD_PDE_H */

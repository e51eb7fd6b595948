// Device PDE functors for the nonlinear ADER-DG path.
//
// Euler (gamma = 1.4), ApplicationExamples/Euler/Euler_ADERDG/EulerSolver_ADERDG.cpp:
//   flux        :61-115  (2D keeps 5 variables, |j|^2 drops Q[3] in 2D, :81-85)
//   eigenvalues :117-147 (u_n -+ c, u_n x3)
// Linear advection (test PDE): F_d = a_d Q, eigenvalues a_d.
// User PDE (kPdeUser): the static hooks of ExaUserPde (include/exahype_amd_pde.h,
// or the header named by EXA_USER_PDE_HEADER) with the 16 user parameters.
//
// The expression order follows the reference hooks so that FP64 results
// agree to rounding level.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/exahype_amd_pde.h"

namespace exa {

enum PdeId { kPdeEuler = 0, kPdeAdvection = 2, kPdeUser = 3 };

constexpr int kUserParams = 16;

// Variables of the user PDE: ExaUserPde::kNVar where the header declares it, else 5.
template <class T, class = void>
struct UserNVar {
  static constexpr int value = 5;
};
template <class T>
struct UserNVar<T, std::void_t<decltype(T::kNVar)>> {
  static constexpr int value = T::kNVar;
};
constexpr int kUserNVar = UserNVar<ExaUserPde>::value;

struct PdeParams {
  int id;
  double a[3];            // advection velocity
  double u[kUserParams];  // user PDE parameters
};

// The user PDE adds source and / or non-conservative terms to the predictor
// (and, with kNcp, fluctuations to the Riemann solve).
__host__ __device__ inline bool pde_has_ncs(const PdeParams& P) {
  return P.id == kPdeUser && (ExaUserPde::kNcp || ExaUserPde::kSource);
}
__host__ __device__ inline bool pde_has_ncp(const PdeParams& P) { return P.id == kPdeUser && ExaUserPde::kNcp; }

// V variables: Euler is compiled for V = 5, the user PDE for V = kUserNVar,
// linear advection for any V (the host accepts no other combination).
template <int D, typename R, int V>
__device__ __forceinline__ void pde_flux(const PdeParams& P, const R* __restrict__ Q, R (&F)[D][V]) {
  if constexpr (V == kUserNVar) {
    if (P.id == kPdeUser) {
      ExaUserPde::flux<D, R>(P.u, Q, F);
      return;
    }
  }
  if (P.id == kPdeAdvection || V != 5) {
#pragma unroll
    for (int d = 0; d < D; d++)
#pragma unroll
      for (int v = 0; v < V; v++) F[d][v] = static_cast<R>(P.a[d]) * Q[v];
    return;
  }
  if constexpr (V == 5) {
    constexpr R gamma = static_cast<R>(1.4);
    R irho;
    if constexpr (sizeof(R) == 8) {
      // 1/rho: hardware reciprocal and two Newton steps (error <= 1 ulp; the
      // IEEE division's scaling and fix-up steps cost twice the instructions)
      irho = __builtin_amdgcn_rcp(Q[0]);
      R e = __builtin_fma(-Q[0], irho, R(1));
      irho = __builtin_fma(irho, e, irho);
      e = __builtin_fma(-Q[0], irho, R(1));
      irho = __builtin_fma(irho, e, irho);
    } else {
      irho = R(1) / Q[0];
    }
    const R j2 = (D == 3) ? Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3] : Q[1] * Q[1] + Q[2] * Q[2];
    const R p = (gamma - R(1)) * (Q[4] - R(0.5) * irho * j2);
    F[0][0] = Q[1];
    F[0][1] = irho * Q[1] * Q[1] + p;
    F[0][2] = irho * Q[2] * Q[1];
    F[0][3] = irho * Q[3] * Q[1];
    F[0][4] = irho * (Q[4] + p) * Q[1];
    F[1][0] = Q[2];
    F[1][1] = irho * Q[1] * Q[2];
    F[1][2] = irho * Q[2] * Q[2] + p;
    F[1][3] = irho * Q[3] * Q[2];
    F[1][4] = irho * (Q[4] + p) * Q[2];
    if constexpr (D == 3) {
      F[2][0] = Q[3];
      F[2][1] = irho * Q[1] * Q[3];
      F[2][2] = irho * Q[2] * Q[3];
      F[2][3] = irho * Q[3] * Q[3] + p;
      F[2][4] = irho * (Q[4] + p) * Q[3];
    }
  }
}

// max_v |lambda_v(Q, d)| with the reference's eigenvalue expressions (Q: V values).
template <int D, int V>
__device__ __forceinline__ double pde_max_abs_eig(const PdeParams& P, const double* __restrict__ Q, int d) {
  if constexpr (V == kUserNVar) {
    if (P.id == kPdeUser) return ExaUserPde::max_abs_eigenvalue<D>(P.u, Q, d);
  }
  if constexpr (V != 5) {
    return fabs(P.a[d]);
  } else {
    if (P.id == kPdeAdvection) return fabs(P.a[d]);
    constexpr double gamma = 1.4;
    const double irho = 1. / Q[0];
    const double j2 = (D == 3) ? Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3] : Q[1] * Q[1] + Q[2] * Q[2];
    const double p = (gamma - 1) * (Q[4] - 0.5 * irho * j2);
    const double u_n = Q[d + 1] * irho;
    const double c = sqrt(gamma * p * irho);
    const double l0 = fabs(u_n - c), l1 = fabs(u_n), l4 = fabs(u_n + c);
    return fmax(fmax(l0, l1), l4);
  }
}

// Source minus non-conservative product at one point (spaceTimePredictorNonlinear.cpph:341-362):
// S = algebraicSource(Q) - nonConservativeProduct(Q, gradQ).
template <int D, typename R, int V>
__device__ __forceinline__ void pde_ncs(const PdeParams& P, const R* Q, const R (&gradQ)[D][V], R* S) {
  static_assert(V == kUserNVar, "source / NCP terms: the user PDE's variables");
  if constexpr (ExaUserPde::kSource) {
    ExaUserPde::source<R>(P.u, Q, S);
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) S[v] = R(0);
  }
  if constexpr (ExaUserPde::kNcp) {
    R ncp[V];
    ExaUserPde::ncp<D, R>(P.u, Q, gradQ, ncp);
#pragma unroll
    for (int v = 0; v < V; v++) S[v] -= ncp[v];
  }
}

// Exact entropy wave of the Euler example (ApplicationExamples/Euler/Euler/
// EulerSolver_ADERDG.cpp:266-286): a density bump advected with v0 = (1/2, 0, 0)
// at pressure 1, rho = 1/2 + 3/10 exp(-|x - x0 - v0 t| / 0.3^D), x0 = (1/2, 1/2, 1/2).
template <int D>
__device__ __forceinline__ void euler_entropy_wave(const double* x, double t, double* Q) {
  const double width = 0.3, v0x = 0.5;
  double r2 = 0.0;
#pragma unroll
  for (int d = 0; d < D; d++) {
    const double c = x[d] - 0.5 - (d == 0 ? v0x * t : 0.0);
    r2 += c * c;
  }
  const double distance = sqrt(r2);
  Q[0] = 0.5 + 0.3 * exp(-distance / pow(width, D));
  Q[1] = Q[0] * v0x;
  Q[2] = 0.0;
  Q[3] = 0.0;
  Q[4] = 1.0 / (1.4 - 1) + 0.5 * Q[0] * (v0x * v0x);
}

}  // namespace exa

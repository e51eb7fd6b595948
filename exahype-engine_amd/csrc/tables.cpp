// Closed-form derivation of the Gauss-Legendre ADER-DG operator tables.
// Everything is computed in long double and rounded to double at the end.
#include "tables.h"

#include <cmath>

namespace exa {
namespace {

using ld = long double;

// Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_n.
void gauss_legendre(int n, ld* x, ld* w) {
  const ld pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; i++) {
    ld z = std::cos(pi * (i + 0.75L) / (n + 0.5L));
    ld dp = 0;
    for (int it = 0; it < 100; it++) {
      ld p0 = 1, p1 = z;
      for (int k = 2; k <= n; k++) {
        const ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) {
        p1 = z;
        p0 = 1;
      }
      // P_n(z) = p1, P_{n-1}(z) = p0
      dp = n * (z * p1 - p0) / (z * z - 1);
      const ld dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-30L) break;
    }
    // recompute derivative at the converged root
    ld p0 = 1, p1 = z;
    for (int k = 2; k <= n; k++) {
      const ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) p0 = 1;
    dp = n * (z * p1 - p0) / (z * z - 1);
    x[i] = z;  // descending in z
    w[i] = 2 / ((1 - z * z) * dp * dp);
  }
}

ld lagrange(const ld* xi, int B, int k, ld s) {
  ld r = 1;
  for (int m = 0; m < B; m++)
    if (m != k) r *= (s - xi[m]) / (xi[k] - xi[m]);
  return r;
}

ld lagrange_deriv(const ld* xi, int B, int k, ld s) {
  ld tot = 0;
  for (int j = 0; j < B; j++) {
    if (j == k) continue;
    ld r = 1 / (xi[k] - xi[j]);
    for (int m = 0; m < B; m++)
      if (m != k && m != j) r *= (s - xi[m]) / (xi[k] - xi[m]);
    tot += r;
  }
  return tot;
}

// In-place Gauss-Jordan inversion with partial pivoting.
bool invert(ld* a, int n) {
  ld inv[kMaxBasis * kMaxBasis];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) inv[i * n + j] = (i == j) ? 1 : 0;
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0) return false;
    if (p != c)
      for (int j = 0; j < n; j++) {
        std::swap(a[p * n + j], a[c * n + j]);
        std::swap(inv[p * n + j], inv[c * n + j]);
      }
    const ld d = a[c * n + c];
    for (int j = 0; j < n; j++) {
      a[c * n + j] /= d;
      inv[c * n + j] /= d;
    }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const ld f = a[r * n + c];
      if (f == 0) continue;
      for (int j = 0; j < n; j++) {
        a[r * n + j] -= f * a[c * n + j];
        inv[r * n + j] -= f * inv[c * n + j];
      }
    }
  }
  for (int i = 0; i < n * n; i++) a[i] = inv[i];
  return true;
}

}  // namespace

bool build_tables(int order, Tables* t) {
  if (order < 0 || order + 1 > kMaxBasis) return false;
  const int B = order + 1;
  t->order = order;
  t->B = B;
  ld z[kMaxBasis], wz[kMaxBasis], xi[kMaxBasis], w[kMaxBasis];
  if (B == 1) {
    xi[0] = 0.5L;
    w[0] = 1;
  } else {
    gauss_legendre(B, z, wz);
    // map to [0,1] in ascending order
    for (int i = 0; i < B; i++) {
      xi[i] = (1 - z[i]) / 2;  // z descending -> xi ascending
      w[i] = wz[i] / 2;
    }
  }
  ld Kxi[kMaxBasis * kMaxBasis], K1[kMaxBasis * kMaxBasis];
  ld F0[kMaxBasis], F1[kMaxBasis];
  for (int k = 0; k < B; k++) {
    F0[k] = (B == 1) ? 1 : lagrange(xi, B, k, 0);
    F1[k] = (B == 1) ? 1 : lagrange(xi, B, k, 1);
  }
  for (int k = 0; k < B; k++)
    for (int l = 0; l < B; l++) {
      const ld d = (B == 1) ? 0 : lagrange_deriv(xi, B, k, xi[l]);
      Kxi[k * B + l] = w[l] * d;
      t->dudx[l * B + k] = static_cast<double>(d);  // dudx[i][j] = phi_j'(xi_i)
    }
  for (int k = 0; k < B; k++)
    for (int l = 0; l < B; l++) K1[k * B + l] = F1[k] * F1[l] - Kxi[k * B + l];
  if (!invert(K1, B)) return false;
  for (int i = 0; i < B; i++) {
    t->w[i] = static_cast<double>(w[i]);
    t->x[i] = static_cast<double>(xi[i]);
    t->F0[i] = static_cast<double>(F0[i]);
    t->F1[i] = static_cast<double>(F1[i]);
    for (int j = 0; j < B; j++) {
      t->Kxi[i * B + j] = static_cast<double>(Kxi[i * B + j]);
      t->iK1[i * B + j] = static_cast<double>(K1[i * B + j]);
      t->egp[i * B + j] =
          static_cast<double>((B == 1) ? 1 : lagrange(xi, B, i, static_cast<ld>(j) / order));
    }
  }
  for (int sub = 0; sub < 3; sub++)
    for (int m = 0; m < B; m++)
      for (int n = 0; n < B; n++)
        t->fgp[(sub * B + m) * B + n] =
            static_cast<double>((B == 1) ? 1 : lagrange(xi, B, m, (sub + xi[n]) / 3));
  return true;
}

double pnpm_for_order(int order) {
  static const double kCflAder[] = {1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015};
  if (order >= 0 && order < 10) return kCflAder[order];
  static const double kCorr[] = {0.18525, 0.1204125, 0.078268125, 0.05087428125, 0.0330682828125,
                                 0.021494383828125};
  if (order >= 10 && order < 16) return kCorr[order - 10] / (2 * order + 1);
  return 0.0;
}

}  // namespace exa

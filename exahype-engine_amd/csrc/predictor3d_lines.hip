// 3D nonlinear space-time predictor + volume integral, "line" decomposition.
//
// Same mathematics as the reference Picard loop
// (ExaHyPE/kernels/aderdg/generic/c/3d/spaceTimePredictorNonlinear.cpph:39-433), written as
//   G[t][s]    = sum_d sum_n dudx[s_d][n] F_d(lQi[t][s|s_d=n])        (flux divergence)
//   lQi'[t'][s]= luh[s] - dt/h * sum_t P[t'][t] G[t][s],  P = iK1 * diag(w)
// which is the reference update lQi' = W^-1 iK1 (W F0 luh - stiffness terms):
// Kxi[n][l] / w_l = dudx[l][n] and iK1 * F0 = 1 (rows of iK1 sum the F0
// weights to one), so the two agree up to rounding.
//
// Decomposition (one workgroup = CPB cells, B^3 lanes per cell):
//   lane (t, z, y) owns the x-line {(t, z, y, x) : x < B}: lQi of its B
//   points (B*V values) lives in registers.  Per Picard iteration:
//     x  : flux + x contraction entirely in registers;
//     y  : F_y is written to the cell's LDS tile, each lane reads one y-line,
//          contracts it in registers and writes the result back onto the
//          same line (lines are disjoint: no barrier between read and write),
//          the owners read their x-lines back;
//     z  : as y with z-lines;
//     t  : G is written to LDS, each lane reads one time line (all t of one
//          point), applies P and writes lQi' back onto the same line.
//   LDS traffic: 12 values per space-time point and variable per iteration,
//   no re-reads of whole lines per output point.
//   On convergence the time averages lQhi = sum_t w_t lQi and
//   lFhi_d = sum_t w_t F_d(lQi of the previous iterate) (the reference
//   averages the flux of the last iteration, :441-515) are reduced over the
//   B time lanes through LDS; extrapolation (:517-687) and the volume integral
//   (volumeIntegralNonlinear.cpph:133-223) follow as in the reference.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace exa {
namespace {

constexpr int lceil64(int n) { return (n + 63) / 64 * 64; }

template <int B>
struct LShape {
  static constexpr int V = 5;
  static constexpr int NS = B * B * B;   // spatial nodes = lanes per cell
  static constexpr int NST = NS * B;     // space-time points per cell
  static constexpr int NF = B * B;
  static constexpr int CPB = NS >= 128 ? 1 : 256 / NS;
  static constexpr int THREADS = lceil64(CPB * NS);
  static constexpr int TILE = NST * V;   // LDS values per cell
};

// Hides a uniform pointer from loop-invariant code motion so that operator
// entries are re-read through the scalar cache in each phase instead of being
// hoisted into (vector) registers for the whole kernel.
template <typename X>
__device__ __forceinline__ const X* opaque(const X* p) {
  unsigned long long a = reinterpret_cast<unsigned long long>(p);
  asm volatile("" : "+s"(a));
  // re-establish uniformity so that the loads stay scalar (s_load)
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(a));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(a >> 32));
  return reinterpret_cast<const X*>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// Index of a line's first value in an LDS array, held in a register of its
// own: the line's reads then differ from it by immediate offsets, and the
// compiler pairs them into ds_read2_b32 off this one register.  A base it can
// see through gets folded into the offsets, which overflow ds_read2's 8-bit
// field, and each pair then costs a vector add for a fresh base.  x lines (B^2
// points per step) take one per two points.
__device__ __forceinline__ int line_base(int i) {
  asm volatile("" : "+v"(i));
  return i;
}

// Makes register values opaque to the optimiser at a phase boundary.  Without
// it the flux products irho*Q[k] of the x, y, z and averaging phases are
// common subexpressions and stay live across the whole Picard iteration
// (~90 extra registers); recomputing them costs a few multiplies.
template <typename R, int N, int M>
__device__ __forceinline__ void launder(R (&a)[N][M]) {
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int j = 0; j < M; j++) asm volatile("" : "+v"(a[i][j]));
}

typedef float exa_f4 __attribute__((ext_vector_type(4)));

// out[i] = sum_n op[i][n] * L[n], i < B <= 8, on the matrix cores with
// v_mfma_f32_4x4x1_16b_f32: block b = lane / 4 computes a 4x4 outer product
// A_b (x) B_b with A_b[i] from lane 4b+i and B_b[j] from lane 4b+j and
// leaves D_b[i][j] in accumulator register i of lane 4b+j.  Every lane puts
// its own line value L[n] into B (column j = its own line) and the operator
// entry op[4h + lane%4][n] into A, so after the B rank-one steps lane 4b+j
// holds rows 4h..4h+3 of its own output line.  aop[h][n] = op[4h+lane%4][n]
// (zero for rows >= B).  Must be executed by all 64 lanes of the wave.
template <int B>
__device__ __forceinline__ void mfma_line(const float (&aop)[2][B], const float (&L)[B], float (&out)[B]) {
  exa_f4 c0 = {0.f, 0.f, 0.f, 0.f}, c1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int n = 0; n < B; n++) {
    c0 = __builtin_amdgcn_mfma_f32_4x4x1f32(aop[0][n], L[n], c0, 0, 0, 0);
    if (B > 4) c1 = __builtin_amdgcn_mfma_f32_4x4x1f32(aop[1][n], L[n], c1, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < B; i++) out[i] = i < 4 ? c0[i] : c1[i - 4];
}

// Sum of x over the 64 lanes of the wave, returned in every lane: DPP row
// shifts build each 16-lane row's sum in its lane 15, row broadcasts carry the
// row sums up to lane 63, whose value is read back as a scalar.  Must be
// executed by all 64 lanes.
__device__ __forceinline__ float wave_sum_f32(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xf, 0xf, false));  // row_shr:1
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xf, 0xf, false));  // row_shr:2
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xf, 0xe, false));  // row_shr:4
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xf, 0xc, false));  // row_shr:8
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xa, 0xf, false));  // row_bcast:15
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xc, 0xf, false));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Largest tile index (in values) any lane of the line kernel reads in the y,
// z and t phases, lanes l >= B^3 included (their (t, z, y) run past the cell).
template <int B, int LS>
constexpr int lines_max_read() {
  constexpr int V = 5;
  int mx = 0;
  for (int l = 0; l < LShape<B>::THREADS; l++) {
    const int t = l / (B * B), z = (l / B) % B, y = l % B;
    const int yb = (t * B + z) * B * LS + y * V, zb = (t * B * B + y) * LS + z * V;
    const int tb = (z * B + y) * LS + t * V;
    const int c[3] = {yb + (B - 1) * LS, zb + (B - 1) * B * LS, tb + (B - 1) * B * B * LS};
    for (int i = 0; i < 3; i++) mx = c[i] + V - 1 > mx ? c[i] + V - 1 : mx;
  }
  return mx;
}

// a[h][n] = col[h B + n]: one lane's operand columns from LDS (16-byte aligned)
template <int B>
__device__ __forceinline__ void load_op(const float* col, float (&a)[2][B]) {
#pragma unroll
  for (int hh = 0; hh < 2; hh++)
#pragma unroll
    for (int n = 0; n < B; n++) a[hh][n] = col[hh * B + n];
}

template <typename R>
__device__ __forceinline__ const R* tab_dudx(const DevTables* T);
template <>
__device__ __forceinline__ const double* tab_dudx<double>(const DevTables* T) { return T->dudx; }
template <>
__device__ __forceinline__ const float* tab_dudx<float>(const DevTables* T) { return T->dudxf; }
template <typename R>
__device__ __forceinline__ const R* tab_P(const DevTables* T);
template <>
__device__ __forceinline__ const double* tab_P<double>(const DevTables* T) { return T->Pt; }
template <>
__device__ __forceinline__ const float* tab_P<float>(const DevTables* T) { return T->Ptf; }

// 1/rho: IEEE division in FP64; the hardware reciprocal (1 ulp) in the FP32
// form, whose accuracy target is 1e-6 relative.
template <typename R>
__device__ __forceinline__ R recip(R x) {
  return R(1) / x;
}
template <>
__device__ __forceinline__ float recip<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

// Euler pressure and 1/rho at one point (EulerSolver_ADERDG.cpp:78-86).
template <typename R>
__device__ __forceinline__ void euler_prim(const R* Q, R& irho, R& p) {
  irho = recip<R>(Q[0]);
  const R j2 = Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3];
  p = (R(1.4) - R(1)) * (Q[4] - R(0.5) * irho * j2);
}

// Flux component d of the Euler flux (EulerSolver_ADERDG.cpp:88-113), in the
// reference's expression order (FP64 forms).
template <typename R>
__device__ __forceinline__ void euler_flux_dir(const R* Q, R irho, R p, int d, R* F) {
  const R qd = Q[d + 1];
  F[0] = qd;
  F[1] = irho * Q[1] * qd + (d == 0 ? p : R(0));
  F[2] = irho * Q[2] * qd + (d == 1 ? p : R(0));
  F[3] = irho * Q[3] * qd + (d == 2 ? p : R(0));
  F[4] = irho * (Q[4] + p) * qd;
}
// FP32 forms: the velocity u_d = q_d / rho is factored out (same flux, five
// multiplies instead of nine; rounding differs at the FP32 level).
template <>
__device__ __forceinline__ void euler_flux_dir<float>(const float* Q, float irho, float p, int d, float* F) {
  const float qd = Q[d + 1], ud = irho * qd;
  F[0] = qd;
  F[1] = Q[1] * ud + (d == 0 ? p : 0.f);
  F[2] = Q[2] * ud + (d == 1 ? p : 0.f);
  F[3] = Q[3] * ud + (d == 2 ? p : 0.f);
  F[4] = (Q[4] + p) * ud;
}

// Point index of the FP32 forms' per-point LDS arrays (the CERK stage fluxes,
// Fc_d, lQhi, lFhi_d): x slowest, then z, y fastest.  A time-line lane (t, z, y)
// owns point (z, y, x = t) at index (t B + z) B + y = its lane index, and the
// lines a wave reads in x, y and z start on distinct LDS banks (with the
// natural x-fastest order the x-lines' starts, 6 apart, met on the same banks).
// Strides along x, y, z: B^2, 1, B.
template <int B>
__host__ __device__ constexpr int lds_point(int z, int y, int x) {
  return (x * B + z) * B + y;
}
template <int B>
__host__ __device__ constexpr int lds_stride(int d) {
  return d == 0 ? B * B : (d == 1 ? 1 : B);
}

// At one point, from its time line q[t][v] (t ascending, FP32):
// lFhi_d = sum_t w_t F_d(q[t]), stored at af[(d NS) V + pv + v] (pv = point
// index * V), and fc_d = sum_t c_t F_d(q[t]) with c = T->cwf, returned.
// Of the 15 Euler flux components 12 are distinct: the mass fluxes are the
// momenta, the momentum fluxes are symmetric (F_x[2] = F_y[1] = q_x q_y / rho,
// F_x[3] = F_z[1], F_y[3] = F_z[2]); each distinct one is formed and summed once.
// MASS: the w- and c-weighted time sums of the momenta (the mass fluxes) are
// given (mw, mc; closed forms for a polynomial time line) instead of summed.
template <int B, bool MASS = false>
__device__ __forceinline__ void time_line_averages(const float (&q)[B][LShape<B>::V], const DevTables* __restrict__ T,
                                                   float* af, int pv, float (&fc)[3][LShape<B>::V],
                                                   const float* mw = nullptr, const float* mc = nullptr) {
  constexpr int V = 5, NS = B * B * B, NDIST = 12;
  // distinct components: q_x, q_y, q_z, xx, yy, zz, xy, xz, yz, ex, ey, ez
  float sw[NDIST], sc[NDIST];
#pragma unroll
  for (int i = 0; i < NDIST; i++) sw[i] = sc[i] = 0.f;
  constexpr int I0 = MASS ? 3 : 0;  // first summed component
#pragma unroll
  for (int t = 0; t < B; t++) {
    const float wt = T->wf[t], ct = T->cwf[t];
    const float* Q = q[t];
    float ir, pr;
    euler_prim<float>(Q, ir, pr);
    const float ux = ir * Q[1], uy = ir * Q[2], uz = ir * Q[3], e = Q[4] + pr;
    const float f[NDIST] = {Q[1],          Q[2],          Q[3],      Q[1] * ux + pr, Q[2] * uy + pr, Q[3] * uz + pr,
                            Q[2] * ux,     Q[3] * ux,     Q[3] * uy, e * ux,         e * uy,         e * uz};
#pragma unroll
    for (int i = I0; i < NDIST; i++) {
      sw[i] += f[i] * wt;
      sc[i] += f[i] * ct;
    }
  }
  if constexpr (MASS) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      sw[i] = mw[i];
      sc[i] = mc[i];
    }
  }
  // F_d[v] -> distinct component
  constexpr int map[3][V] = {{0, 3, 6, 7, 9}, {1, 6, 4, 8, 10}, {2, 7, 8, 5, 11}};
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int v = 0; v < V; v++) {
      af[d * NS * V + pv + v] = sw[map[d][v]];
      fc[d][v] = sc[map[d][v]];
    }
}

// Component v of the FP32 x flux, the same operations as euler_flux_dir<float>(.., 0, ..).
__device__ __forceinline__ float euler_flux_x_comp(const float* Q, float irho, float p, int v) {
  const float ud = irho * Q[1];
  if (v == 0) return Q[1];
  if (v == 1) return Q[1] * ud + p;
  if (v == 4) return (Q[4] + p) * ud;
  return Q[v] * ud;
}

// Post-Picard phase shared by the line and tile kernels (one cell per
// workgroup): q = previous iterate (lane (t, z, y) < B^3 holds its x-line),
// tl = final iterate [t][z][y][x][v].  Time averages lQhi = sum_t w_t lQi
// (final iterate) and lFhi_d = sum_t w_t F_d(previous iterate) (the reference
// averages the flux of the last iteration, :441-515), extrapolation
// (:517-687) and the volume integral (volumeIntegralNonlinear.cpph:133-223);
// buf is the workgroup's LDS buffer (>= 4 B^3 V doubles), tl may lie in it.
template <int B, typename R>
__device__ __forceinline__ void finish_cell(R (&q)[B][LShape<B>::V], R* tl, unsigned char* buf, int l, long cell,
                                            double dt, double h, int iters, const DevTables* __restrict__ T,
                                            double* __restrict__ lduh, double* __restrict__ lQhbnd,
                                            double* __restrict__ lFhbnd, int* __restrict__ picard,
                                            double* luh_upd) {
  using S = LShape<B>;
  constexpr int V = S::V, NS = S::NS, NF = S::NF;
  const bool lane_on = l < NS;
  const int t = l / (B * B), z = (l / B) % B, y = l % B;
  const int own = ((t * B + z) * B + y) * B * V;
  const int tbase = ((z * B + y) * B + t) * V;
  const double* w = T->w;
  // ---- time averages: lQhi = sum_t w_t lQi (final iterate),
  //      lFhi_d = sum_t w_t F_d(previous iterate) ----
  // component 0 (lQhi) is reduced straight from the tile, which holds the
  // final iterate on every x-line; components 1..3 overwrite the tile with
  // F_d of the previous iterate.
  R acc[4][V];  // reduced values of point (z, y, x = t)
#pragma unroll
  for (int c = 0; c < 4; c++) {
    launder(q);
    __syncthreads();
    if (lane_on && c > 0) {
#pragma unroll
      for (int x = 0; x < B; x++) {
        R F[V];
        R ir, pr;
        euler_prim<R>(q[x], ir, pr);
        euler_flux_dir<R>(q[x], ir, pr, c - 1, F);
#pragma unroll
        for (int v = 0; v < V; v++) tl[own + x * V + v] = F[v];
      }
    }
    __syncthreads();
    if (lane_on) {
#pragma unroll
      for (int v = 0; v < V; v++) {
        R a = R(0);
#pragma unroll
        for (int n = 0; n < B; n++) a += tl[tbase + n * NS * V + v] * static_cast<R>(w[n]);
        acc[c][v] = a;
      }
    }
  }
  __syncthreads();
  double* Qh = reinterpret_cast<double*>(buf);  // [NS][V]
  double* Fh = Qh + NS * V;                     // [3][NS][V]
  const int s = (z * B + y) * B + t;
  if (lane_on) {
#pragma unroll
    for (int v = 0; v < V; v++) {
      Qh[s * V + v] = static_cast<double>(acc[0][v]);
#pragma unroll
      for (int d = 0; d < 3; d++) Fh[(d * NS + s) * V + v] = static_cast<double>(acc[d + 1][v]);
    }
  }
  __syncthreads();
  if (lane_on) {
    // volume integral at point s = (z, y, x = t)
    const int sz = z, sy = y, sx = t;
    const double invDx = 1.0 / h;
    double du[V];
#pragma unroll
    for (int v = 0; v < V; v++) du[v] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double tw = (d == 0) ? w[sz] * w[sy] : (d == 1 ? w[sz] * w[sx] : w[sy] * w[sx]);
      const double upd = tw * invDx;
      const int c = (d == 0) ? sx : (d == 1 ? sy : sz);
      const int stride = (d == 0) ? 1 : (d == 1 ? B : B * B);
      const int base = s - c * stride;
#pragma unroll
      for (int v = 0; v < V; v++)
#pragma unroll
        for (int m = 0; m < B; m++)
          du[v] += T->Kxi[c * B + m] * Fh[(d * NS + base + m * stride) * V + v] * upd;
    }
    if (luh_upd) {
      // volume part of the solution update in place (solutionUpdate.cpph:28-67)
      const double upd = dt / (w[sz] * w[sy] * w[sx]);
#pragma unroll
      for (int v = 0; v < V; v++) luh_upd[(cell * NS + s) * V + v] += du[v] * upd;
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) lduh[(cell * NS + s) * V + v] = du[v];
    }
    // extrapolation to the six faces; lane l produces outputs l, l + NS, ...
    constexpr int NQ = 6 * NF * V;
    for (int o = l; o < NQ; o += NS) {
      const int v = o % V;
      const int fn = (o / V) % NF;
      const int f = o / (V * NF);
      const int d = f >> 1;
      const double* FC = (f & 1) ? T->F1 : T->F0;
      const int a = fn / B, b = fn % B;
      double qa = 0.0, fa = 0.0;
#pragma unroll
      for (int m = 0; m < B; m++) {
        const int node = (d == 0) ? (a * B + b) * B + m : (d == 1 ? (a * B + m) * B + b : (m * B + a) * B + b);
        qa += Qh[node * V + v] * FC[m];
        fa += Fh[(d * NS + node) * V + v] * FC[m];
      }
      lQhbnd[cell * NQ + o] = qa;
      lFhbnd[cell * NQ + o] = fa;
    }
    if (l == 0 && picard) picard[cell] = iters;
  }
}

// Post-Picard phase of the FP32 kernels, one lane per spatial node s:
// P = previous iterate, Q = final iterate, both FP32 tiles [t][s][v].
//   lQhi[s] = sum_t w_t Q[t][s],  lFhi_d[s] = sum_t w_t F_d(P[t][s])
// (the reference averages the flux of the last iteration, :441-515), then the
// volume integral (volumeIntegralNonlinear.cpph:133-223) and the extrapolation
// (:517-687) from the averages, in FP32; the update (or lduh) and the face
// arrays are written in FP64.  The flux of P is evaluated once per space-time
// point for all three directions.  avg: FP32 LDS scratch of 4 B^3 V values,
// which may overlap P (it is written after a barrier).
template <int B>
__device__ __forceinline__ void finish_from_averages(const float* aq, const float* af, int l, long cell, double dt,
                                                     double h, int iters, const DevTables* __restrict__ T,
                                                     double* __restrict__ lduh, double* __restrict__ lQhbnd,
                                                     double* __restrict__ lFhbnd, int* __restrict__ picard,
                                                     double* luh_upd, int nthreads, int f32faces);

template <int B>
__device__ __forceinline__ void finish_cell_points(const float* P, const float* Q, float* avg, int l, long cell,
                                                   double dt, double h, int iters, const DevTables* __restrict__ T,
                                                   double* __restrict__ lduh, double* __restrict__ lQhbnd,
                                                   double* __restrict__ lFhbnd, int* __restrict__ picard,
                                                   double* luh_upd, int nthreads, int f32faces) {
  constexpr int V = 5, NS = B * B * B, COLS = NS * V;
  float qh[V], fh[3][V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    qh[v] = 0.f;
    fh[0][v] = fh[1][v] = fh[2][v] = 0.f;
  }
  if (l < NS) {
#pragma unroll
    for (int t = 0; t < B; t++) {
      const float wt = T->wf[t];
      float q[V], p[V];
#pragma unroll
      for (int v = 0; v < V; v++) {
        q[v] = Q[t * COLS + l * V + v];
        p[v] = P[t * COLS + l * V + v];
      }
#pragma unroll
      for (int v = 0; v < V; v++) qh[v] += q[v] * wt;
      float ir, pr;
      euler_prim<float>(p, ir, pr);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        float F[V];
        euler_flux_dir<float>(p, ir, pr, d, F);
#pragma unroll
        for (int v = 0; v < V; v++) fh[d][v] += F[v] * wt;
      }
    }
  }
  __syncthreads();  // P may be overwritten by avg below
  float* aq = avg;             // [NS][V]
  float* af = avg + NS * V;    // [3][NS][V]
  if (l < NS) {
    const int pl = lds_point<B>(l / (B * B), (l / B) % B, l % B);  // node l in the LDS point order
#pragma unroll
    for (int v = 0; v < V; v++) {
      aq[pl * V + v] = qh[v];
#pragma unroll
      for (int d = 0; d < 3; d++) af[(d * NS + pl) * V + v] = fh[d][v];
    }
  }
  __syncthreads();
  finish_from_averages<B>(aq, af, l, cell, dt, h, iters, T, lduh, lQhbnd, lFhbnd, picard, luh_upd, nthreads,
                          f32faces);
}

// Volume integral and extrapolation of the FP32 forms from the time averages
// in LDS: aq = lQhi [NS][V], af = lFhi_d [3][NS][V] (node index (z, y, x), x
// fastest).  Lane l < NS handles node l of the volume integral; the
// extrapolation runs over (direction, face node, variable) triples.
template <int B>
__device__ __forceinline__ void finish_from_averages(const float* aq, const float* af, int l, long cell, double dt,
                                                     double h, int iters, const DevTables* __restrict__ T,
                                                     double* __restrict__ lduh, double* __restrict__ lQhbnd,
                                                     double* __restrict__ lFhbnd, int* __restrict__ picard,
                                                     double* luh_upd, int nthreads, int f32faces) {
  constexpr int V = 5, NS = B * B * B, NF = B * B;
  if (l < NS) {
    // volume integral at the node of LDS point l = (x B + z) B + y
    const int y = l % B, z = (l / B) % B, x = l / (B * B);
    const int nl = (z * B + y) * B + x;  // its index in the cell's global node order
    const float invDx = __builtin_amdgcn_rcpf(static_cast<float>(h));  // 1 ulp: FP32 forms only
    float du[V];
#pragma unroll
    for (int v = 0; v < V; v++) du[v] = 0.f;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const float tw = (d == 0) ? T->wf[z] * T->wf[y] : (d == 1 ? T->wf[z] * T->wf[x] : T->wf[y] * T->wf[x]);
      const int c = (d == 0) ? x : (d == 1 ? y : z);
      const int stride = lds_stride<B>(d) * V;  // [point][v] stride along d
      const int lb = line_base((d * NS + l) * V - c * stride);
      int pb = lb;
      float part[V];
#pragma unroll
      for (int v = 0; v < V; v++) part[v] = 0.f;
#pragma unroll
      for (int m = 0; m < B; m++) {
        const float k = T->Kxif[c * B + m];
        if (d == 0 && m > 0 && !(m & 1)) pb = line_base(lb + m * stride);  // x: a base per two points
        const int nb = d == 0 ? pb + (m & 1) * stride : lb + m * stride;
#pragma unroll
        for (int v = 0; v < V; v++) part[v] += k * af[nb + v];
      }
#pragma unroll
      for (int v = 0; v < V; v++) du[v] += part[v] * (tw * invDx);
    }
    if (luh_upd) {
      const double upd = dt * (T->iw[z] * T->iw[y] * T->iw[x]);
#pragma unroll
      for (int v = 0; v < V; v++) luh_upd[(cell * NS + nl) * V + v] += static_cast<double>(du[v]) * upd;
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) lduh[(cell * NS + nl) * V + v] = static_cast<double>(du[v]);
    }
  }
  // extrapolation: a lane takes one (variable, face node) pair and gives both
  // faces of the three directions (d unrolled: compile-time line strides and
  // face offsets; the pair is decoded once).
  // FP64 faces: the reference layout [face][node][var] (var fastest across
  // lanes); FP32 faces: [face][var][node] (node fastest), the mesh-internal
  // layout of the FP32 step whose corrector loads one variable row per step.
  constexpr int NPAIR = NF * V;
  for (int o = l; o < NPAIR; o += nthreads) {
    const int v = f32faces ? o / NF : o % V;
    const int fn = f32faces ? o % NF : o / V;
    const int a = fn / B, b = fn % B;
    const long fb = cell * 6 * NF * V + (f32faces ? v * NF + fn : fn * V + v);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      // the line through face node (a, b) along d, (a, b) = (z, y), (z, x), (y, x):
      // LDS point(m) = base + m * stride
      const int base = (d == 0) ? lds_point<B>(a, b, 0) : (d == 1 ? lds_point<B>(a, 0, b) : lds_point<B>(0, a, b));
      const int sv = lds_stride<B>(d) * V;
      const float* pq = aq + base * V + v;
      const float* pf = af + d * NS * V + base * V + v;
      float q0 = 0.f, q1 = 0.f, f0 = 0.f, f1 = 0.f;
#pragma unroll
      for (int m = 0; m < B; m++) {
        const float qv = pq[m * sv], fv = pf[m * sv];
        const float c0 = T->F0f[m], c1 = T->F1f[m];
        q0 += qv * c0;
        q1 += qv * c1;
        f0 += fv * c0;
        f1 += fv * c1;
      }
      if (f32faces) {  // the values are FP32 numbers: FP32 storage is exact
        float* qf = reinterpret_cast<float*>(lQhbnd);
        float* ff = reinterpret_cast<float*>(lFhbnd);
        qf[fb + (2 * d) * NF * V] = q0;
        qf[fb + (2 * d + 1) * NF * V] = q1;
        ff[fb + (2 * d) * NF * V] = f0;
        ff[fb + (2 * d + 1) * NF * V] = f1;
      } else {
        lQhbnd[fb + (2 * d) * NF * V] = q0;
        lQhbnd[fb + (2 * d + 1) * NF * V] = q1;
        lFhbnd[fb + (2 * d) * NF * V] = f0;
        lFhbnd[fb + (2 * d + 1) * NF * V] = f1;
      }
    }
  }
  if (l == 0 && picard) picard[cell] = iters;
}

// Picard guard of the one-iteration CERK form and the cell list of its
// recomputation.  guard = {cells listed this step, cells listed in total, list[]}.
//   ONE && CG: with guard != nullptr a cell whose relative proxy residual
//     sum (lQhi(1) - lQhi(0))^2 >= tol2 sum lQhi(1)^2 (FP64 accumulation;
//     lQhi(k) = sum_t w_t lQi(k)(t)) writes nothing (luh, faces and its Picard
//     count stay as they are) and appends itself to the list.
//   LIST (fixed-count CERK form): the workgroups loop over the listed cells.
struct GuardArgs {
  int* guard = nullptr;
  double tol2 = 0.0;
};

template <int B, typename R, bool MF, bool FIX, bool CG, bool ONE = false, bool LIST = false, bool PART = false>
// Fixed-count form: four workgroups per CU (LDS 39 KB each; the register
// budget of 128 costs a few spills, measured faster than three workgroups)
__global__ __launch_bounds__((LShape<B>::THREADS), (ONE ? 5 : (FIX ? 4 : 1))) void predictor3d_lines_kernel(
    long ncells, const double* luh, const double* __restrict__ hcell, const double* __restrict__ dtp,
    const int* __restrict__ cinst, double* __restrict__ lduh, double* __restrict__ lQhbnd,
    double* __restrict__ lFhbnd, int* __restrict__ picard, const DevTables* __restrict__ T, PdeParams P,
    int maxIter, int adaptive, double* luh_upd, int f32faces, GuardArgs G, long cell0 = 0) {
  using S = LShape<B>;
  constexpr int V = S::V, NS = S::NS;
  // x-line stride in the tile; the fixed-count form pads each line by one
  // value so that the lanes' own-line accesses (stride LS) hit distinct banks
  constexpr int LS = B * V + (FIX ? 1 : 0);
  constexpr int TILE = B * B * B * LS;
  static_assert(S::CPB == 1, "one cell per workgroup");
  static_assert(!MF || (sizeof(R) == 4 && B <= 8), "matrix-core form: FP32, B <= 8");
  static_assert(!FIX || MF, "fixed-count form: FP32 matrix-core kernel");
  static_assert(!CG || FIX, "CERK guess: fixed-count form");
  if constexpr (FIX) adaptive = 0;  // launched only for a fixed Picard count
  constexpr int NWAVE = S::THREADS / 64;
  constexpr int TILE_BYTES = TILE * int(sizeof(R));
  constexpr int AVG_BYTES = 4 * NS * V * 8;  // FP64 form: lQhi + lFhi_{x,y,z} in FP64
  // FP32 adaptive forms: a second tile for the previous iterate (point-parallel
  // finish).  Fixed Picard count (FIX): the time lanes form the averages
  // themselves, lFhi_d in iteration maxIter - 2 (from the iterate the last
  // iteration evaluates fluxes of) into af behind the tile, lQhi in the last
  // iteration; the tile then holds lQhi.
  constexpr int AF_BYTES = 3 * NS * V * 4;
  // ONE (one Picard iteration, collapsed): no sweep through the tile; the
  // buffer holds the CERK stage fluxes / Fc_d [3][NS][V] and lQhi [NS][V], then af
  static_assert(!ONE || FIX, "one collapsed iteration: fixed-count form");
  static_assert(!LIST || (FIX && CG && !ONE), "cell list: the fixed-count CERK form");
  static_assert(!PART || (ONE && CG && !LIST), "a part of the mesh: the one-iteration CERK form");
  constexpr int ONE_BYTES = 4 * NS * V * 4;
  constexpr int NEED = ONE ? ONE_BYTES + AF_BYTES
                          : (FIX ? TILE_BYTES + AF_BYTES : (sizeof(R) == 4 ? 2 * TILE_BYTES : TILE_BYTES));
  constexpr int BUF_BYTES = (FIX || NEED > AVG_BYTES) ? NEED : AVG_BYTES;
  static_assert(!MF || ONE || (lines_max_read<B, LS>() + 1) * 4 <= BUF_BYTES, "idle lanes' reads stay inside buf");
  __shared__ __align__(16) unsigned char buf[BUF_BYTES];
  __shared__ double red[NWAVE];
  [[maybe_unused]] __shared__ double red2[ONE ? NWAVE : 1];
  R* tl = reinterpret_cast<R*>(buf);

  const int l = threadIdx.x;              // (t, z, y) for l < NS
  const bool lane_on = l < NS;            // lanes NS..THREADS-1 only join barriers
  const int t = l / (B * B), z = (l / B) % B, y = l % B;

  // matrix-core operand columns, staged in LDS and loaded at each contraction
  // phase (registers are the occupancy limit): opl[0][r][h][n] = dudx[4h + r][n],
  // opl[1][r][h][n] = P[4h + r][n] (zero rows >= B); a lane uses r = l % 4
  __shared__ __align__(16) float opl[2][4][2 * B];
  if constexpr (MF) {
    if (l < 16 * B) {
      const int op = l / (8 * B), r = (l / (2 * B)) % 4, hh = (l % (2 * B)) / B, n = l % B;
      const int row = 4 * hh + r;
      const float* tab = op ? T->Ptf : T->dudxf;
      opl[op][r][hh * B + n] = row < B ? tab[row * B + n] : 0.f;
    }
    __syncthreads();
  }
  const int opr = l & 3;

  auto cell_work = [&](const long cell) {
  const double dt = dtp[cinst ? cinst[cell] : 0];
  const double h = hcell[cell];
  // operators (uniform across lanes; read through the scalar cache at use):
  //   Dm[i][j] = dudx[i][j] (derivative), Pm[i][j] = iK1[i][j] * w[j]
  const R* Dm = tab_dudx<R>(T);
  const R* Pm = tab_P<R>(T);
  const R cdt = static_cast<R>(dt / h);

  // time-line role: this lane solves in time for point (z, y, x = t)
  R u[V];
  R q[B][V];  // lQi at (t, z, y, x), x < B
  {
    const long base = (cell * NS + ((lane_on ? z : 0) * B + (lane_on ? y : 0)) * B) * V;
    if constexpr (!CG) {  // the CERK form starts from the guess in the tile
#pragma unroll
      for (int x = 0; x < B; x++)
#pragma unroll
        for (int v = 0; v < V; v++) q[x][v] = static_cast<R>(luh[base + x * V + v]);
    }
#pragma unroll
    for (int v = 0; v < V; v++) u[v] = static_cast<R>(luh[base + (lane_on ? t : 0) * V + v]);
  }
  // point (t, z, y, x) of the tile at ((t B + z) B + y) LS + x V
  const int own = ((t * B + z) * B + y) * LS;             // this lane's x-line in the tile
  const int ybase = (t * B + z) * B * LS + y * V;         // y-line (t, z, *, x = y), stride LS
  const int zbase = (t * B * B + y) * LS + z * V;         // z-line (t, *, y, x = z), stride B LS
  const int tbase = (z * B + y) * LS + t * V;             // time line of (z, y, x = t), stride B^2 LS
  [[maybe_unused]] const int pv = lds_point<B>(z, y, t) * V;  // that point in the [point][v] LDS arrays

  // FIX: the last Picard iteration is collapsed.  Its output lQi' = luh -
  // dt/h P G enters only lQhi = sum_t w_t lQi'(t), and by linearity
  //   lQhi = luh - dt/h sum_d dudx_d Fc_d,  Fc_d = sum_t c_t F_d(lQi(t)),
  //   c_t = sum_t' w_t' P[t'][t] (T->cwf)
  // with lQi the iterate that iteration evaluates fluxes of.  Its time lane
  // forms lFhi_d (into af) and Fc_d (fc, later through the tile) from that
  // iterate's time line; nfull = maxIter - 1 full sweeps remain.
  [[maybe_unused]] float* af = reinterpret_cast<float*>(buf + (ONE ? ONE_BYTES : TILE_BYTES));
  const int nfull = ONE ? 0 : (FIX ? maxIter - 1 : maxIter);
  // Fc_d [3][NS][V] at the tile's start, written once no time line is read any more
  [[maybe_unused]] auto store_fc = [&](const float (&fc)[3][V]) {
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int v = 0; v < V; v++) tl[d * NS * V + pv + v] = fc[d][v];
  };
  [[maybe_unused]] float g0[V];
  if constexpr (CG) {
    // CERK initial guess (the reference's optimised kernels,
    // KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_nonlinear_cpp.template:121-173):
    // K1 = sum_d dudx_d F_d(luh) / h, lwh = luh - dt K1, K2 the same of lwh,
    // lQi(t) = luh - dt xi_t K1 - dt xi_t^2 (K2 - K1) / 2.  Lane l works on its
    // time-line point p = (z, y, x = t) (u = luh there); the stage fluxes go
    // through the tile, which the Picard loop has not used yet.
    float* fl = tl;  // [3][NS][V], points in the LDS point order
    float K1[V], K2[V], lw[V];
#pragma unroll
    for (int v = 0; v < V; v++) lw[v] = u[v];
#pragma unroll
    for (int stage = 0; stage < 2; stage++) {
      if (lane_on) {
        float ir, pr;
        euler_prim<float>(lw, ir, pr);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          float F[V];
          euler_flux_dir<float>(lw, ir, pr, d, F);
#pragma unroll
          for (int v = 0; v < V; v++) fl[d * NS * V + pv + v] = F[v];
        }
      }
      __syncthreads();
      // K = h K_stage = sum_d dudx_d F_d; the 1/h goes into the step coefficients
      float K[V] = {0.f, 0.f, 0.f, 0.f, 0.f};
      if (lane_on) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const int c = d == 0 ? t : (d == 1 ? y : z);              // coordinate of p along d
          const int stride = lds_stride<B>(d) * V;                    // [point][v] stride along d
          // dudx[c][n] from the operator columns in LDS (opl[0][r][hh B + n] = dudx[4 hh + r][n])
          float dr[B];
#pragma unroll
          for (int n = 0; n < B; n++) dr[n] = opl[0][c & 3][(c >> 2) * B + n];
          const int lb = line_base(d * NS * V + pv - c * stride);
          int pb = lb;
#pragma unroll
          for (int n = 0; n < B; n++) {
            if (d == 0 && n > 0 && !(n & 1)) pb = line_base(lb + n * stride);  // x: a base per two points
            const int nb = d == 0 ? pb + (n & 1) * stride : lb + n * stride;
#pragma unroll
            for (int v = 0; v < V; v++) K[v] += fl[nb + v] * dr[n];
          }
        }
      }
      __syncthreads();  // fl is rewritten by the next stage / the guess
#pragma unroll
      for (int v = 0; v < V; v++) {
        if (stage == 0) {
          K1[v] = K[v];
          lw[v] = u[v] - cdt * K1[v];
        } else {
          K2[v] = K[v];
        }
      }
    }
    if constexpr (ONE) {
      // sum_t w_t xi_t = 1/2, sum_t w_t xi_t^2 = 1/3: the guess's time average
      // is luh - cdt (K1 / 2 + (K2 - K1) / 6) = luh - cdt g0
#pragma unroll
      for (int v = 0; v < V; v++) g0[v] = K1[v] * (1.f / 3.f) + K2[v] * (1.f / 6.f);
    }
    float qg[B][V];
#pragma unroll
    for (int i = 0; i < B; i++) {
      const float xt = static_cast<float>(T->x[i]);
      const float a1 = cdt * xt, a2 = 0.5f * cdt * xt * xt;
#pragma unroll
      for (int v = 0; v < V; v++) qg[i][v] = u[v] - a1 * K1[v] - a2 * (K2[v] - K1[v]);
    }
    if (lane_on) {
      if (nfull > 0) {
#pragma unroll
        for (int i = 0; i < B; i++)
#pragma unroll
          for (int v = 0; v < V; v++) tl[tbase + i * (B * B * LS) + v] = qg[i][v];
      } else {  // the collapsed iteration evaluates fluxes of the guess (stage reads are done)
        // the guess is quadratic in time, q = luh - cdt x K1 - cdt x^2 (K2 - K1) / 2, so the
        // time sums of the momenta (the mass fluxes) are closed forms: w-weighted
        // luh - cdt g0, c-weighted luh m0 - cdt K1 m1 - cdt (K2 - K1) m2 / 2 (m_k = sum_t c_t x_t^k)
        float mw[3], mc[3];
        const float m0 = T->cmom[0], m1 = T->cmom[1], m2 = T->cmom[2];
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const int v = i + 1;
          mw[i] = u[v] - cdt * (K1[v] * (1.f / 3.f) + K2[v] * (1.f / 6.f));
          mc[i] = u[v] * m0 - cdt * K1[v] * m1 - (0.5f * cdt) * (K2[v] - K1[v]) * m2;
        }
        float fc[3][V];
        time_line_averages<B, true>(qg, T, af, pv, fc, mw, mc);
        store_fc(fc);
      }
    }
    if (nfull > 0) {
      __syncthreads();
      if (lane_on) {  // this lane's x-line of the guess; the tile's next writes go to own lines only
#pragma unroll
        for (int x = 0; x < B; x++)
#pragma unroll
          for (int v = 0; v < V; v++) q[x][v] = tl[own + x * V + v];
      }
    }
  } else if constexpr (FIX) {
    if (nfull == 0 && lane_on) {  // the collapsed iteration evaluates fluxes of the initial guess
      float qi[B][V], fc[3][V];
#pragma unroll
      for (int i = 0; i < B; i++)
#pragma unroll
        for (int v = 0; v < V; v++) qi[i][v] = u[v];
      time_line_averages<B>(qi, T, af, pv, fc);
      store_fc(fc);
    }
  }

  int iters = maxIter;
  for (int iter = 0; iter < nfull; iter++) {
    R G[B][V];
    Dm = opaque(Dm);
    launder(q);
    // 1/rho and p of the line's points, once per iteration for the three flux directions
    R irl[B], prl[B];
#pragma unroll
    for (int n = 0; n < B; n++) euler_prim<R>(q[n], irl[n], prl[n]);
    // ---- x: flux + contraction in registers (streamed over the line) ----
#pragma unroll
    for (int i = 0; i < B; i++)
#pragma unroll
      for (int v = 0; v < V; v++) G[i][v] = R(0);
    if constexpr (MF) {
      float aD[2][B];
      load_op<B>(opl[0][opr], aD);
      // one flux component at a time: B operands live instead of B V
#pragma unroll
      for (int v = 0; v < V; v++) {
        float L[B], o[B];
#pragma unroll
        for (int n = 0; n < B; n++) L[n] = euler_flux_x_comp(q[n], irl[n], prl[n], v);
        mfma_line<B>(aD, L, o);
#pragma unroll
        for (int i = 0; i < B; i++) G[i][v] = o[i];
      }
    } else {
#pragma unroll
      for (int n = 0; n < B; n++) {
        R F[V];
        euler_flux_dir<R>(q[n], irl[n], prl[n], 0, F);
#pragma unroll
        for (int i = 0; i < B; i++)
#pragma unroll
          for (int v = 0; v < V; v++) G[i][v] += Dm[i * B + n] * F[v];
      }
    }
    // ---- y, z: line transposes through the tile ----
#pragma unroll
    for (int d = 1; d <= 2; d++) {
      Dm = opaque(Dm);
      launder(q);
      if (lane_on) {
#pragma unroll
        for (int x = 0; x < B; x++) {
          R F[V];
          euler_flux_dir<R>(q[x], irl[x], prl[x], d, F);
#pragma unroll
          for (int v = 0; v < V; v++) tl[own + x * V + v] = F[v];
        }
      }
      __syncthreads();
      if constexpr (MF) {
        const int lb = (d == 1) ? ybase : zbase;
        const int ls = (d == 1) ? LS : B * LS;
        float aD[2][B];
        load_op<B>(opl[0][opr], aD);
#pragma unroll
        for (int v = 0; v < V; v++) {
          float L[B], o[B];
#pragma unroll
          // lanes >= NS read inside buf too (lines_max_read, checked at compile time)
          // and discard: a matrix-core column depends only on its own lane's operand
          for (int n = 0; n < B; n++) L[n] = tl[lb + n * ls + v];
          mfma_line<B>(aD, L, o);
          if (lane_on) {
#pragma unroll
            for (int i = 0; i < B; i++) tl[lb + i * ls + v] = o[i];
          }
        }
      } else if (lane_on) {
        const int lb = (d == 1) ? ybase : zbase;
        const int ls = (d == 1) ? LS : B * LS;
        // one variable at a time; each lane writes back onto its own line only
#pragma unroll
        for (int v = 0; v < V; v++) {
          R L[B];
#pragma unroll
          for (int n = 0; n < B; n++) L[n] = tl[lb + n * ls + v];
#pragma unroll
          for (int i = 0; i < B; i++) {
            R a = R(0);
#pragma unroll
            for (int n = 0; n < B; n++) a += Dm[i * B + n] * L[n];
            tl[lb + i * ls + v] = a;
          }
        }
      }
      __syncthreads();
      if (lane_on) {
#pragma unroll
        for (int x = 0; x < B; x++)
#pragma unroll
          for (int v = 0; v < V; v++) G[x][v] += tl[own + x * V + v];
      }
      // no barrier: the next tile writes (F_z, or G) go to each lane's own
      // x-line, which no other lane touches until the next barrier
    }
    // ---- t: time solve through the tile ----
    Pm = opaque(Pm);
    if (lane_on) {
#pragma unroll
      for (int x = 0; x < B; x++)
#pragma unroll
        for (int v = 0; v < V; v++) tl[own + x * V + v] = G[x][v];
    }
    __syncthreads();
    if constexpr (MF) {
      constexpr int ts = B * B * LS;
      float aP[2][B];
      load_op<B>(opl[1][opr], aP);
      if constexpr (FIX) {
        float qn[B][V];  // the new iterate on this lane's time line
#pragma unroll
        for (int v = 0; v < V; v++) {
          float L[B], o[B];
#pragma unroll
          for (int n = 0; n < B; n++) L[n] = tl[tbase + n * ts + v];  // lanes >= NS: as above
          mfma_line<B>(aP, L, o);
#pragma unroll
          for (int i = 0; i < B; i++) qn[i][v] = u[v] - cdt * o[i];
        }
        if (iter < nfull - 1) {  // block-uniform
          if (lane_on) {
#pragma unroll
            for (int i = 0; i < B; i++)
#pragma unroll
              for (int v = 0; v < V; v++) tl[tbase + i * ts + v] = qn[i][v];
          }
        } else {
          // the iterate the collapsed iteration starts from; its flux averages
          // are formed behind the barrier, once every time line has been read
          __syncthreads();
          if (lane_on) {
            float fc[3][V];
            time_line_averages<B>(qn, T, af, pv, fc);
            store_fc(fc);
          }
        }
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) {
          float L[B], o[B];
#pragma unroll
          for (int n = 0; n < B; n++) L[n] = tl[tbase + n * ts + v];  // lanes >= NS: as above
          mfma_line<B>(aP, L, o);
          if (lane_on) {
#pragma unroll
            for (int i = 0; i < B; i++) tl[tbase + i * ts + v] = u[v] - cdt * o[i];
          }
        }
      }
    } else if (lane_on) {
      constexpr int ts = B * B * LS;
#pragma unroll
      for (int v = 0; v < V; v++) {
        R L[B];
#pragma unroll
        for (int n = 0; n < B; n++) L[n] = tl[tbase + n * ts + v];
#pragma unroll
        for (int i = 0; i < B; i++) {
          R a = R(0);
#pragma unroll
          for (int n = 0; n < B; n++) a += Pm[i * B + n] * L[n];
          tl[tbase + i * ts + v] = u[v] - cdt * a;
        }
      }
    }
    __syncthreads();
    double sq = 0.0;
    if (adaptive && lane_on) {  // the residual decides only the adaptive Picard rule
#pragma unroll
      for (int x = 0; x < B; x++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          const double dq = static_cast<double>(tl[own + x * V + v]) - static_cast<double>(q[x][v]);
          sq += dq * dq;
        }
    }
    bool last = (iter == nfull - 1);
    if (adaptive) {
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) sq += __shfl_xor(sq, m, 64);
      if ((l & 63) == 0) red[l >> 6] = sq;
      __syncthreads();
      double res = 0.0;
#pragma unroll
      for (int i = 0; i < NWAVE; i++) res += red[i];
      if (res < 1e-7 * 1e-7) {  // block-uniform decision (one cell per block)
        iters = iter;
        last = true;
      }
    }
    if (last) break;
    if (lane_on) {
#pragma unroll
      for (int x = 0; x < B; x++)
#pragma unroll
        for (int v = 0; v < V; v++) q[x][v] = tl[own + x * V + v];
    }
    // the next tile writes are to own x-lines (disjoint) behind the x phase;
    // the adaptive rule rewrites red[] next iteration
    if (adaptive) __syncthreads();
  }

  // ---- after the Picard loop: q = previous iterate, the tile holds the final one ----
  if constexpr (FIX) {
    // collapsed last iteration: Fc_d [3][NS][V] at the tile's start (stored
    // above), lQhi [NS][V] behind it
    float* fcb = tl;
    float* aq = tl + 3 * NS * V;
    [[maybe_unused]] float gsq = 0.f, qsq = 0.f;
    __syncthreads();
    if (lane_on) {
      // coordinates of p = (z, y, x = t), recomputed from pv (not kept live through the loop)
      const int pp = pv / V, py = pp % B, pz = (pp / B) % B, px = pp / (B * B);
      float dv[V] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int c = d == 0 ? px : (d == 1 ? py : pz);              // coordinate of p along d
        const int stride = lds_stride<B>(d) * V;                      // [point][v] stride along d
        float dr[B];                                                  // dudx[c][n]
#pragma unroll
        for (int n = 0; n < B; n++) dr[n] = opl[0][c & 3][(c >> 2) * B + n];
        const int lb = line_base(d * NS * V + pv - c * stride);
        int pb = lb;
#pragma unroll
        for (int n = 0; n < B; n++) {
          if (d == 0 && n > 0 && !(n & 1)) pb = line_base(lb + n * stride);  // x: a base per two points
          const int nb = d == 0 ? pb + (n & 1) * stride : lb + n * stride;
#pragma unroll
          for (int v = 0; v < V; v++) dv[v] += dr[n] * fcb[nb + v];
        }
      }
#pragma unroll
      for (int v = 0; v < V; v++) {
        const float qh = u[v] - cdt * dv[v];
        aq[pv + v] = qh;
        if constexpr (ONE && CG) {
          // lQhi(1) - lQhi(0) = cdt (g0 - dv): luh cancels, so the difference
          // carries no rounding of luh; the factor cdt^2 is applied to the sum
          const float dq = g0[v] - dv[v];
          gsq += dq * dq;
          qsq += qh * qh;
        }
      }
    }
    if constexpr (ONE && CG) {
      if (G.guard) {  // kernel-uniform branch
        // wave sums in FP32 (at most 64 x 5 terms), the workgroup sum in FP64
        const float gw = wave_sum_f32(gsq), qw = wave_sum_f32(qsq);
        if ((l & 63) == 0) {
          red[l >> 6] = static_cast<double>(gw);
          red2[l >> 6] = static_cast<double>(qw);
        }
      }
    }
    __syncthreads();  // lQhi (and the guard's wave sums) complete
    if constexpr (ONE && CG) {
      if (G.guard) {
        double res = 0.0, nrm = 0.0;
#pragma unroll
        for (int i = 0; i < NWAVE; i++) {
          res += red[i];
          nrm += red2[i];
        }
        const double c2 = static_cast<double>(cdt) * static_cast<double>(cdt);
        if (res * c2 >= G.tol2 * nrm) {  // block-uniform: leave the cell to the recomputation
          if (l == 0) G.guard[2 + atomicAdd(G.guard, 1)] = static_cast<int>(cell);
          return;
        }
      }
    }
    finish_from_averages<B>(aq, af, l, cell, dt, h, iters, T, lduh, lQhbnd, lFhbnd, picard, luh_upd, S::THREADS,
                            f32faces);
  } else if constexpr (sizeof(R) == 4) {
    float* t2 = reinterpret_cast<float*>(buf) + TILE;
    if (lane_on) {
#pragma unroll
      for (int x = 0; x < B; x++)
#pragma unroll
        for (int v = 0; v < V; v++) t2[own + x * V + v] = q[x][v];
    }
    __syncthreads();
    finish_cell_points<B>(t2, tl, t2, l, cell, dt, h, iters, T, lduh, lQhbnd, lFhbnd, picard, luh_upd,
                          S::THREADS, f32faces);
  } else {
    finish_cell<B, R>(q, tl, buf, l, cell, dt, h, iters, T, lduh, lQhbnd, lFhbnd, picard, luh_upd);
  }
  };

  if constexpr (LIST) {
    // the listed cells (guard[0] of them, at guard + 2), grid-strided
    const long nlist = G.guard[0];
    for (long i = blockIdx.x; i < nlist; i += gridDim.x) {
      cell_work(G.guard[2 + i]);
      __syncthreads();  // the next cell rewrites the LDS buffers
    }
  } else {
    // PART: the launch covers cells [cell0, cell0 + gridDim.x) (its own instantiation,
    // so that profiles tell the launches over a part of the mesh from whole-mesh ones)
    cell_work(PART ? cell0 + blockIdx.x : blockIdx.x);
  }
}


// ---------------------------------------------------------------------------
// FP32 tile form on full-rate matrix cores (v_mfma_f32_16x16x4_f32).
//
// The cell's space-time iterate lQi[t][z][y][x][v] and the flux-divergence
// accumulator G[t][z][y][x][v] live in two FP32 LDS tiles.  Every contraction
//   out[i][c] = sum_n op[i][n] X[n][c]      (op = dudx along x, y or z, P along t)
// runs over pairs of 16-column sets stacked block-diagonally:
//   rows i' = 8 set + i (i < B; rows B..7 of each set are zero),
//   k'      = B set + n (K = 2B, ceil(2B / 4) MFMA steps),
//   A[i'][k'] = op[i][n] when the two sets agree, else 0;  B[k'][j] = X[n][column j of that set].
// Lane l supplies A[l % 16][4s + l / 16] and B[4s + l / 16][l % 16] in step s and
// receives D[4 (l / 16) + r][l % 16], r < 4.  In a flux direction the columns
// are the B^3 lines along d (points (line, n)); the B values a lane supplies
// are the same points for all five variables, so the lane evaluates the Euler
// flux there once.  In the time solve the columns are the B^3 V (point,
// variable) pairs.  Per Picard iteration and cell: 3 * ceil(B^3 / 32) * 5 *
// ceil(2B / 4) + ceil(B^3 V / 32) * ceil(2B / 4) MFMAs (417 at B = 6) -- the
// line form issues 4x4x1 MFMAs at a quarter of the matrix-core rate.
// Same mathematics and Picard rule as the line form (the FP32 sums are
// ordered differently).
//
// BF16 operand form (template flag BF): the same pairs of the sum go through
// one v_mfma_f32_16x16x16_bf16 per contraction instead of NK FP32 steps.  In
// that instruction lane (g, j) = (l / 16, l % 16) supplies A[j][4g + i] and
// B[4g + i][j], i < 4; the lane puts its FP32-step values of step s = i there
// (i = NK..3 zero), so k-index 4g + i pairs the same lane values as k-index
// 4i + g of the FP32 form.  Operator entries and flux / divergence values
// are rounded to BF16 (nearest even); accumulation, the iterate in LDS, the
// flux evaluations and everything after the Picard loop stay FP32.
// ---------------------------------------------------------------------------
typedef short exa_s4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<short>(u >> 16);
}

template <int NK>
__device__ __forceinline__ exa_s4 bf16_pack(const float (&a)[NK]) {
  static_assert(NK <= 4, "one 16x16x16 step holds four k values per lane");
  exa_s4 r = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < NK; i++) r[i] = bf16_bits(a[i]);
  return r;
}

template <int B>
struct TShape {
  static constexpr int V = 5;
  static constexpr int NS = B * B * B;         // lines per direction (t and two spatial indices)
  static constexpr int TILE = NS * B * V;      // values per space-time tile
  static constexpr int COLS = NS * V;          // columns of the time solve
  static constexpr int NK = (2 * B + 3) / 4;   // MFMA steps per contraction
  static constexpr int PAIRS = (NS + 31) / 32;
  static constexpr int TPAIRS = (COLS + 31) / 32;
  static constexpr int THREADS = 256;
  static constexpr int NWAVE = THREADS / 64;
  static constexpr int TPW = (TPAIRS + NWAVE - 1) / NWAVE;  // time-solve pairs per wave
  static constexpr int PPW = (PAIRS + NWAVE - 1) / NWAVE;    // line pairs per wave and direction
};

// tile position of point n on line L of direction d (lines: t slowest, then the
// two remaining spatial indices in z, y, x order)
template <int B>
__device__ __forceinline__ int line_point(int d, int L, int n) {
  if (d == 0) return L * B + n;
  const int lo = L % B, hi = L / B;  // hi = t * B + (z or y)
  if (d == 1) return (hi * B + n) * B + lo;
  const int tt = hi / B, mid = hi % B;  // d == 2: L = (t * B + y) * B + x
  return ((tt * B + n) * B + mid) * B + lo;
}

template <int B, bool BF>
__global__ __launch_bounds__(256, 3) void predictor3d_tile_kernel(
    long ncells, const double* luh, const double* __restrict__ hcell, const double* __restrict__ dtp,
    const int* __restrict__ cinst, double* __restrict__ lduh, double* __restrict__ lQhbnd,
    double* __restrict__ lFhbnd, int* __restrict__ picard, const DevTables* __restrict__ T, int maxIter,
    int adaptive, double* luh_upd, int f32faces) {
  using S = TShape<B>;
  constexpr int V = S::V, NS = S::NS, TILE = S::TILE, COLS = S::COLS, NK = S::NK;
  static_assert(2 * B <= 4 * NK && B <= 8, "block-diagonal pair must fit 16 rows");
  constexpr int TILE_BYTES = TILE * 4;
  constexpr int AVG_BYTES = 4 * NS * V * 8;
  constexpr int BUF_BYTES = 2 * TILE_BYTES > AVG_BYTES ? 2 * TILE_BYTES : AVG_BYTES;
  __shared__ __align__(16) unsigned char buf[BUF_BYTES];
  __shared__ double red[S::NWAVE];

  const int l = threadIdx.x;
  const int wave = l >> 6, lane = l & 63;
  const int g = lane >> 4, j = lane & 15;
  const long cell = blockIdx.x;
  const double dt = dtp[cinst ? cinst[cell] : 0];
  const double h = hcell[cell];
  const float cdt = static_cast<float>(dt / h);

  // per-step operand roles of this lane
  int kset[NK], kn[NK];
  bool kok[NK];
  float aD[NK], aP[NK];
  {
    const int rset = j >> 3, ri = j & 7;
#pragma unroll
    for (int s = 0; s < NK; s++) {
      const int kp = 4 * s + g;
      kok[s] = kp < 2 * B;
      kset[s] = kok[s] ? kp / B : 0;
      kn[s] = kok[s] ? kp % B : 0;
      const bool on = kok[s] && rset == kset[s] && ri < B;
      aD[s] = on ? T->dudxf[ri * B + kn[s]] : 0.f;
      aP[s] = on ? T->Ptf[ri * B + kn[s]] : 0.f;
    }
  }
  const exa_s4 aDb = bf16_pack<NK>(aD), aPb = bf16_pack<NK>(aP);

  float* tA = reinterpret_cast<float*>(buf);
  float* tB = tA + TILE;
  // initial guess lQi(t) = luh at every time node
  const double* lc = luh + cell * NS * V;
  for (int i = l; i < TILE; i += S::THREADS) tA[i] = static_cast<float>(lc[i % COLS]);
  // luh of the time-solve columns this lane writes (output set = g >> 1)
  float u[S::TPW];
#pragma unroll
  for (int k = 0; k < S::TPW; k++) {
    const int c = 32 * (wave + S::NWAVE * k) + 16 * (g >> 1) + j;
    u[k] = c < COLS ? static_cast<float>(lc[c]) : 0.f;
  }
  __syncthreads();

  int iters = maxIter;
  for (int iter = 0; iter < maxIter; iter++) {
    // ---- G = sum_d D_d F_d(lQi): tA -> tB ----
    // The initial guess is constant in time, so in the first iteration G is
    // evaluated on the t = 0 lines only and copied to the other time slices
    // (the same operations on the same values: bitwise what the full
    // evaluation gives).  A wave's line pairs are processed together so that
    // its 5 * PPW accumulation chains interleave.
    const int nlines = iter == 0 ? B * B : NS;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      float F[S::PPW][NK][V];
#pragma unroll
      for (int q = 0; q < S::PPW; q++) {
        const int p = wave + S::NWAVE * q;
#pragma unroll
        for (int s = 0; s < NK; s++) {
          const int L = 32 * p + 16 * kset[s] + j;
          if (kok[s] && L < nlines) {
            const float* Q = tA + line_point<B>(d, L, kn[s]) * V;
            float Qv[V];
#pragma unroll
            for (int v = 0; v < V; v++) Qv[v] = Q[v];
            float ir, pr;
            euler_prim<float>(Qv, ir, pr);
            euler_flux_dir<float>(Qv, ir, pr, d, F[q][s]);
          } else {
#pragma unroll
            for (int v = 0; v < V; v++) F[q][s][v] = 0.f;
          }
        }
      }
      exa_f4 acc[S::PPW][V];
#pragma unroll
      for (int q = 0; q < S::PPW; q++)
#pragma unroll
        for (int v = 0; v < V; v++) acc[q][v] = exa_f4{0.f, 0.f, 0.f, 0.f};
      if (BF) {
#pragma unroll
        for (int q = 0; q < S::PPW; q++)
          if (32 * (wave + S::NWAVE * q) < nlines)  // wave-uniform
#pragma unroll
            for (int v = 0; v < V; v++) {
              float fv[NK];
#pragma unroll
              for (int s = 0; s < NK; s++) fv[s] = F[q][s][v];
              acc[q][v] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aDb, bf16_pack<NK>(fv), acc[q][v], 0, 0, 0);
            }
      } else {
#pragma unroll
        for (int s = 0; s < NK; s++)
#pragma unroll
          for (int q = 0; q < S::PPW; q++)
            if (32 * (wave + S::NWAVE * q) < nlines)  // wave-uniform
#pragma unroll
              for (int v = 0; v < V; v++)
                acc[q][v] = __builtin_amdgcn_mfma_f32_16x16x4f32(aD[s], F[q][s][v], acc[q][v], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < S::PPW; q++) {
        const int p = wave + S::NWAVE * q;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int ip = 4 * g + r, os = ip >> 3, oi = ip & 7;
          const int L = 32 * p + 16 * os + j;
          if (oi < B && L < nlines) {
            const int base = line_point<B>(d, L, oi) * V;
#pragma unroll
            for (int v = 0; v < V; v++) tB[base + v] = (d == 0) ? acc[q][v][r] : tB[base + v] + acc[q][v][r];
          }
        }
      }
      __syncthreads();
    }
    if (iter == 0) {  // G(t) = G(t = 0)
      for (int i = COLS + l; i < TILE; i += S::THREADS) tB[i] = tB[i % COLS];
      __syncthreads();
    }
    // ---- time solve lQi' = luh - dt/h P G, in place in tB; residual vs tA ----
    // all loads of the wave's column pairs precede its stores (in place)
    float X[S::TPW][NK];
#pragma unroll
    for (int k = 0; k < S::TPW; k++) {
      const int p = wave + S::NWAVE * k;
#pragma unroll
      for (int s = 0; s < NK; s++) {
        const int c = 32 * p + 16 * kset[s] + j;
        X[k][s] = (kok[s] && c < COLS) ? tB[kn[s] * COLS + c] : 0.f;
      }
    }
    exa_f4 tacc[S::TPW];
#pragma unroll
    for (int k = 0; k < S::TPW; k++) tacc[k] = exa_f4{0.f, 0.f, 0.f, 0.f};
    if (BF) {
#pragma unroll
      for (int k = 0; k < S::TPW; k++)
        if (wave + S::NWAVE * k < S::TPAIRS)  // wave-uniform
          tacc[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aPb, bf16_pack<NK>(X[k]), tacc[k], 0, 0, 0);
    } else {
#pragma unroll
      for (int s = 0; s < NK; s++)
#pragma unroll
        for (int k = 0; k < S::TPW; k++)
          if (wave + S::NWAVE * k < S::TPAIRS)  // wave-uniform
            tacc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(aP[s], X[k][s], tacc[k], 0, 0, 0);
    }
    double sq = 0.0;
#pragma unroll
    for (int k = 0; k < S::TPW; k++) {
      const int p = wave + S::NWAVE * k;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ip = 4 * g + r, os = ip >> 3, oi = ip & 7;
        const int c = 32 * p + 16 * os + j;
        if (oi < B && c < COLS) {
          const int idx = oi * COLS + c;
          const float nv = u[k] - cdt * tacc[k][r];
          if (adaptive) {  // the residual decides only the adaptive Picard rule
            const double dq = static_cast<double>(nv) - static_cast<double>(tA[idx]);
            sq += dq * dq;
          }
          tB[idx] = nv;
        }
      }
    }
    __syncthreads();
    bool last = (iter == maxIter - 1);
    if (adaptive) {
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) sq += __shfl_xor(sq, m, 64);
      if (lane == 0) red[wave] = sq;
      __syncthreads();
      double res = 0.0;
#pragma unroll
      for (int i = 0; i < S::NWAVE; i++) res += red[i];
      if (res < 1e-7 * 1e-7) {  // block-uniform decision (one cell per block)
        iters = iter;
        last = true;
      }
      __syncthreads();  // red is rewritten next iteration
    }
    if (last) break;
    float* tmp = tA;
    tA = tB;
    tB = tmp;
  }

  // ---- tB = final iterate, tA = previous iterate ----
  finish_cell_points<B>(tA, tB, tA, l, cell, dt, h, iters, T, lduh, lQhbnd, lFhbnd, picard, luh_upd, S::THREADS,
                        f32faces);
}

}  // namespace

bool lines_supported(const KernelCfg& k) {
  return k.dim == 3 && k.nvar == 5 && k.npar == 0 && k.order + 1 == 6 && k.pde.id == kPdeEuler;
}

int launch_predictor3d_tile(const KernelCfg& k, hipStream_t st, long n, const double* luh, const double* h,
                            const double* dt, double* lduh, double* qb, double* fb, int* pic, const int* cinst,
                            double* luh_upd, bool f32faces) {
  constexpr int B = 6;
  if (k.precision == 0) return -1;  // FP32 form only
  const int adaptive = k.picard_max <= 0;
  const int maxIter = adaptive ? 2 * (k.order + 1) : k.picard_max;
  if (k.precision == 2)  // BF16 operands
    hipLaunchKernelGGL((predictor3d_tile_kernel<B, true>), dim3(n), dim3(TShape<B>::THREADS), 0, st, n, luh, h, dt,
                       cinst, lduh, qb, fb, pic, k.tab, maxIter, adaptive, luh_upd, f32faces ? 1 : 0);
  else
    hipLaunchKernelGGL((predictor3d_tile_kernel<B, false>), dim3(n), dim3(TShape<B>::THREADS), 0, st, n, luh, h, dt,
                       cinst, lduh, qb, fb, pic, k.tab, maxIter, adaptive, luh_upd, f32faces ? 1 : 0);
  return static_cast<int>(hipGetLastError());
}

// guard[0] -> guard[1] (cells recomputed in total), guard[0] = 0
__global__ void guard_reset_kernel(int* guard) {
  if (threadIdx.x == 0) {
    guard[1] += guard[0];
    guard[0] = 0;
  }
}

int launch_predictor3d_lines(const KernelCfg& k, hipStream_t st, long n, const double* luh, const double* h,
                             const double* dt, double* lduh, double* qb, double* fb, int* pic, bool mfma,
                             const int* cinst, double* luh_upd, bool f32faces, int* guard, long cell0,
                             long cell_end) {
  constexpr int B = 6;
  using S = LShape<B>;
  static_assert(S::CPB == 1, "one cell per workgroup");
  // cells [cell0, cell_end): the one-iteration CERK form only; the others cover all n cells
  const bool range = cell0 != 0 || (cell_end >= 0 && cell_end != n);
  if (cell_end < 0) cell_end = n;
  if (cell0 < 0 || cell0 >= cell_end || cell_end > n) return -1;
  const long blocks = range ? cell_end - cell0 : (n + S::CPB - 1) / S::CPB;
  const int adaptive = k.picard_max <= 0;
  const int maxIter = adaptive ? 2 * (k.order + 1) : k.picard_max;
  const int f32 = (f32faces && k.precision != 0) ? 1 : 0;  // FP32 faces exist only in the FP32 forms
  const dim3 grid(blocks), block(S::THREADS);
  const GuardArgs none;
  if (k.precision == 0) {
    if (range || k.guess) return -1;  // the FP64 line kernel has no CERK guess (the node-column kernel has)
    hipLaunchKernelGGL((predictor3d_lines_kernel<B, double, false, false, false>), grid, block, 0, st, n, luh, h, dt,
                       cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, 0, none);
  } else {
    if (k.guess && !(mfma && !adaptive)) return -1;  // CERK: fixed-count matrix-core form only
    if (range && !(mfma && !adaptive && maxIter == 1 && k.guess)) return -1;
    if (mfma && !adaptive && maxIter == 1 && k.guess) {  // one collapsed iteration: no sweep code, smaller LDS
      // Picard guard: cells whose proxy residual reaches the tolerance are
      // recomputed with the fixed count k.guard_recount (listed cells only)
      GuardArgs g;
      if (guard && k.guard_tol > 0) {
        g.guard = guard;
        g.tol2 = k.guard_tol * k.guard_tol;
      }
      if (range)
        hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, true, true, true, false, true>), grid, block, 0,
                           st, n, luh, h, dt, cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32,
                           g, cell0);
      else
        hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, true, true, true>), grid, block, 0, st, n, luh,
                           h, dt, cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32, g);
      if (g.guard) {
        // persistent grid over the list: four workgroups per CU of the fixed-count
        // kernel (a part of the mesh: one per CU, the list is usually empty)
        const long lmax = range ? 256 : 1024;
        const long lgrid = blocks < lmax ? blocks : lmax;
        const int rc = k.guard_recount > 0 ? k.guard_recount : (B - 3 > 1 ? B - 3 : 1);
        hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, true, true, false, true>), dim3(lgrid), block, 0,
                           st, n, luh, h, dt, cinst, lduh, qb, fb, pic, k.tab, k.pde, rc, adaptive, luh_upd, f32, g);
        hipLaunchKernelGGL(guard_reset_kernel, dim3(1), dim3(64), 0, st, guard);
      }
    } else if (mfma && !adaptive && maxIter == 1)
      hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, true, false, true>), grid, block, 0, st, n, luh, h,
                         dt, cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32, none);
    else if (mfma && !adaptive && k.guess)
      hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, true, true>), grid, block, 0, st, n, luh, h, dt,
                         cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32, none);
    else if (mfma && !adaptive)
      hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, true, false>), grid, block, 0, st, n, luh, h, dt,
                         cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32, none);
    else if (mfma)
      hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, true, false, false>), grid, block, 0, st, n, luh, h, dt,
                         cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32, none);
    else
      hipLaunchKernelGGL((predictor3d_lines_kernel<B, float, false, false, false>), grid, block, 0, st, n, luh, h, dt,
                         cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, luh_upd, f32, none);
  }
  return static_cast<int>(hipGetLastError());
}

bool line_kernels_serve(const KernelCfg& k) {
  if (!lines_supported(k) || k.dialect != 0) return false;
  return !k.guess || (k.precision != 0 && k.picard_max > 0 && (k.pred_variant == 0 || k.pred_variant == 3));
}

bool predictor_supports_range(const KernelCfg& k) {
  return line_kernels_serve(k) && k.precision == 1 && k.picard_max == 1 && k.guess &&
         (k.pred_variant == 0 || k.pred_variant == 3);
}

bool predictor_uses_guard(const KernelCfg& k) { return predictor_supports_range(k) && k.guard_tol > 0; }

bool predictor_f32_faces_supported(const KernelCfg& k) {
  return line_kernels_serve(k) && k.precision != 0 && k.pred_variant != 1;
}

}  // namespace exa

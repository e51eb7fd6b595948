// HIP kernels of the nonlinear ADER-DG cell update for CDNA4 (gfx950).
//
// Reference path (ExaHyPE/kernels/aderdg/generic/c/{2d,3d}):
//   predictor_kernel  : spaceTimePredictorNonlinear.cpph (Picard loop :39-433,
//                       time averaging :441-515, extrapolation :517-687) fused with
//                       volumeIntegralNonlinear.cpph (time-averaged branch :133-223)
//   riemann_kernel    : riemannSolverNonlinear.cpph:37-128 + RiemannSolverUtils.h:152-177
//   corrector_kernel  : faceIntegralNonlinear.cpph:21-103 (faces 0..2D-1 in order,
//                       ADERDGSolver.cpp:1207-1221) + solutionUpdate.cpph:28-67 +
//                       stableTimeStepSize.cpph:14-62 with the global min reduction
//
// Work decomposition (MI355X-first, not the reference's loop nests):
//   * predictor: one workgroup holds CPB whole cells; one thread owns one
//     spatial node s of one cell and keeps that node's whole space-time
//     column lQi[s][0..B)[0..nData) plus its Picard accumulators in VGPRs.
//     Per time slice the node's flux is evaluated in registers and published
//     to a double-buffered LDS slab; each thread then reads the B-point lines
//     through its node along x/y/z (stiffness contraction), so the only LDS
//     traffic is the flux slab.  lQi/rhs/lFi never touch HBM: per cell-step
//     the kernel reads luh once and writes lduh, lQhbnd, lFhbnd once.
//   * Riemann: one wavefront per face, lanes over face nodes; the face
//     averages are wave reductions (DPP shuffles).
//   * corrector: thread per (cell, node); the per-cell CFL minimum is reduced
//     in LDS and merged with one 64-bit atomicMin per workgroup on the IEEE
//     bit pattern (exact and order independent for positive doubles), into
//     the instance's dt_next or, on the uniform mesh, into one of 1024 slots
//     on separate cache lines that dt_reduce_kernel merges afterwards.
//   * CERK initial guess (optimised-kernel dialect, or chosen per context):
//     two flux divergences through the flux slab before the Picard loop
//     (predictor_tile<..., CG = true>).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "engine.h"

namespace exa {
namespace {

constexpr int ceil64(int n) { return (n + 63) / 64 * 64; }
constexpr int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

template <int D, int B, int V, int NP>
struct Shape {
  static constexpr int ND = V + NP;
  static constexpr int NS = ipow(B, D);      // spatial nodes per cell
  static constexpr int NF = ipow(B, D - 1);  // nodes per face
  static constexpr int CPB = NS >= 128 ? 1 : 256 / NS;
  static constexpr int THREADS = ceil64(CPB * NS);
  static constexpr int NSP = next_pow2(NS);
};

template <typename R>
__device__ __forceinline__ R tabK(const DevTables* T, int i);
template <>
__device__ __forceinline__ double tabK<double>(const DevTables* T, int i) { return T->Kxi[i]; }
template <>
__device__ __forceinline__ float tabK<float>(const DevTables* T, int i) { return T->Kxif[i]; }
template <typename R>
__device__ __forceinline__ R tabI(const DevTables* T, int i);
template <>
__device__ __forceinline__ double tabI<double>(const DevTables* T, int i) { return T->iK1[i]; }
template <>
__device__ __forceinline__ float tabI<float>(const DevTables* T, int i) { return T->iK1f[i]; }

template <typename R>
__device__ __forceinline__ R tabD(const DevTables* T, int i);
template <>
__device__ __forceinline__ double tabD<double>(const DevTables* T, int i) { return T->dudx[i]; }
template <>
__device__ __forceinline__ float tabD<float>(const DevTables* T, int i) { return T->dudxf[i]; }

// Coordinates (z, y, x) of spatial node s; z = 0 in 2D.
template <int D, int B>
__device__ __forceinline__ void node_zyx(int s, int& z, int& y, int& x) {
  x = s % B;
  y = (s / B) % B;
  z = (D == 3) ? s / (B * B) : 0;
}

// Product of the quadrature weights transverse to direction d, in the order
// the reference multiplies them (x: w_z*w_y, y: w_z*w_x, z: w_y*w_x).
template <int D>
__device__ __forceinline__ double transverse_weight(const double* w, int d, int z, int y, int x) {
  if (D == 3) {
    if (d == 0) return w[z] * w[y];
    if (d == 1) return w[z] * w[x];
    return w[y] * w[x];
  }
  return d == 0 ? w[y] : w[x];
}

// Node index along direction d of node s, and the index of s's projection
// on the face orthogonal to d (x-faces (z,y), y-faces (z,x), z-faces (y,x)).
template <int D, int B>
__device__ __forceinline__ int coord(int d, int z, int y, int x) {
  return d == 0 ? x : (d == 1 ? y : z);
}
template <int D, int B>
__device__ __forceinline__ int face_node(int d, int z, int y, int x) {
  if (D == 3) return d == 0 ? z * B + y : (d == 1 ? z * B + x : y * B + x);
  return d == 0 ? y : x;
}
template <int D, int B>
__device__ __forceinline__ int stride_of(int d) {
  return d == 0 ? 1 : (d == 1 ? B : B * B);
}

// Volume integral (volumeIntegralNonlinear.cpph:133-223; lduh zeroed, then x, y,
// z contributions) and extrapolation (spaceTimePredictorNonlinear.cpph:517-687)
// of node s of one cell from its time averages in LDS: Qh [node][ND], Fh
// [d][node][V] (FP64).  du_add (nullable): W-weighted extra term of the node
// (the time-averaged source).  Shared by the node-column and slab predictors.
template <int D, int B, int V, int NP>
__device__ __forceinline__ void node_volume_and_faces(
    const double* Qh, const double* Fh, int s, long cell, const double (&ta)[D], const double (&tb)[D],
    const int (&cc)[D], const int (&base)[D], const double (&invDx)[D], double W, double dt, int iters,
    const double* du_add, const DevTables* __restrict__ T, double* __restrict__ lduh, double* __restrict__ lQhbnd,
    double* __restrict__ lFhbnd, int* __restrict__ picard, double* luh_upd) {
  using S = Shape<D, B, V, NP>;
  constexpr int ND = S::ND, NS = S::NS, NF = S::NF;
  // ---- volume integral (lduh zeroed, then x, y, z contributions) ----
  {
    double du[V];
#pragma unroll
    for (int v = 0; v < V; v++) du[v] = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const double upd = (D == 3 ? ta[d] * tb[d] : ta[d]) * invDx[d];
      const int c = cc[d];
#pragma unroll
      for (int m = 0; m < B; m++) {
        // Kxi[c][m] times the weight factor once per (d, m): one FMA per flux value
        const double k = T->Kxi[c * B + m] * upd;
#pragma unroll
        for (int v = 0; v < V; v++) du[v] += k * Fh[(d * NS + base[d] + m * stride_of<D, B>(d)) * V + v];
      }
    }
    if (du_add) {  // source / NCP: lduh += W lShi (volumeIntegralNonlinear.cpph:204-222)
#pragma unroll
      for (int v = 0; v < V; v++) du[v] += W * du_add[v];
    }
    if (luh_upd) {
      // volume part of the solution update in place (solutionUpdate.cpph:28-67: dt / w)
      const double upd = dt / W;
#pragma unroll
      for (int v = 0; v < V; v++) luh_upd[(cell * NS + s) * ND + v] += du[v] * upd;
    } else {
#pragma unroll
      for (int v = 0; v < V; v++) lduh[(cell * NS + s) * V + v] = du[v];
    }
  }

  // ---- extrapolation to the 2D faces (faces zeroed, then FCoeff sums) ----
  if constexpr (D == 3) {
    // a lane takes one (face node, value) pair and gives both faces of the
    // three directions: the pair is decoded once, the line strides are
    // compile-time constants; outputs in the reference layout [face][node][value]
    constexpr int NPAIR = NF * ND;
    for (int pr = s; pr < NPAIR; pr += NS) {
      const int v = pr % ND, fn = pr / ND;
      const int a = fn / B, b = fn % B;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        // the line through face node (a, b) along d: node n0 + l stride, (a, b) = (z, y), (z, x), (y, x)
        const int n0 = d == 0 ? (a * B + b) * B : (d == 1 ? a * B * B + b : a * B + b);
        constexpr int st[3] = {1, B, B * B};
        double q0 = 0.0, q1 = 0.0, f0 = 0.0, f1 = 0.0;
#pragma unroll
        for (int l = 0; l < B; l++) {
          const int node = n0 + l * st[d];
          const double qv = Qh[node * ND + v];
          q0 += qv * T->F0[l];
          q1 += qv * T->F1[l];
          if (v < V) {
            const double fv = Fh[(d * NS + node) * V + v];
            f0 += fv * T->F0[l];
            f1 += fv * T->F1[l];
          }
        }
        lQhbnd[cell * (2 * D * NF * ND) + ((2 * d) * NF + fn) * ND + v] = q0;
        lQhbnd[cell * (2 * D * NF * ND) + ((2 * d + 1) * NF + fn) * ND + v] = q1;
        if (v < V) {
          lFhbnd[cell * (2 * D * NF * V) + ((2 * d) * NF + fn) * V + v] = f0;
          lFhbnd[cell * (2 * D * NF * V) + ((2 * d + 1) * NF + fn) * V + v] = f1;
        }
      }
    }
  } else {
    constexpr int NQ = 2 * D * NF * ND;
    for (int o = s; o < NQ; o += NS) {
      const int v = o % ND;
      const int fn = (o / ND) % NF;
      const int f = o / (ND * NF);
      const int d = f >> 1;
      const double* FC = (f & 1) ? T->F1 : T->F0;
      // node of the face line at parameter l
      int a = fn / B, b = fn % B;
      if (D == 2) { a = 0; b = fn; }
      double acc = 0.0;
      const bool skip = (D == 2) && (d == 0) && (v >= V);  // 2D x-faces carry no parameters
      if (!skip) {
#pragma unroll
        for (int l = 0; l < B; l++) {
          int nz, ny, nx;
          if (D == 3) {
            if (d == 0) { nz = a; ny = b; nx = l; }
            else if (d == 1) { nz = a; ny = l; nx = b; }
            else { nz = l; ny = a; nx = b; }
          } else {
            nz = 0;
            if (d == 0) { ny = b; nx = l; }
            else { ny = l; nx = b; }
          }
          const int node = (D == 3) ? (nz * B + ny) * B + nx : ny * B + nx;
          acc += Qh[node * ND + v] * FC[l];
        }
      }
      lQhbnd[cell * (2 * D * NF * ND) + o] = acc;
    }
    constexpr int NFL = 2 * D * NF * V;
    for (int o = s; o < NFL; o += NS) {
      const int v = o % V;
      const int fn = (o / V) % NF;
      const int f = o / (V * NF);
      const int d = f >> 1;
      const double* FC = (f & 1) ? T->F1 : T->F0;
      int a = fn / B, b = fn % B;
      if (D == 2) { a = 0; b = fn; }
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < B; l++) {
        int nz, ny, nx;
        if (D == 3) {
          if (d == 0) { nz = a; ny = b; nx = l; }
          else if (d == 1) { nz = a; ny = l; nx = b; }
          else { nz = l; ny = a; nx = b; }
        } else {
          nz = 0;
          if (d == 0) { ny = b; nx = l; }
          else { ny = l; nx = b; }
        }
        const int node = (D == 3) ? (nz * B + ny) * B + nx : ny * B + nx;
        acc += Fh[(d * NS + node) * V + v] * FC[l];
      }
      lFhbnd[cell * (2 * D * NF * V) + o] = acc;
    }
  }
  if (s == 0 && picard) picard[cell] = iters;
}

// The 3D Euler flux has 12 distinct components: the mass fluxes are the
// momenta (slots 0-2), the momentum fluxes are symmetric (diagonal 3-5,
// xy, xz, yz 6-8), the energy fluxes 9-11.  euler_slot(d, v) = slot of F_d[v];
// the first occurrence (d smallest) of a slot is its canonical component.
__host__ __device__ constexpr int euler_slot(int d, int v) {
  return v == 0 ? d : (v == 4 ? 9 + d : (v == d + 1 ? 3 + d : 4 + d + v));
}
__host__ __device__ constexpr bool euler_canonical(int d, int v) { return !(v >= 1 && v <= 3 && v < d + 1); }

// Node weight W (reference: w[z] w[y] w[x], 2D w[y] w[x]), the two transverse
// weights per direction in the reference's multiplication order, the node's
// coordinate along d and the first node of its line along d.
template <int D, int B>
__device__ __forceinline__ void node_geometry(int s, const double* w, double& W, double (&ta)[D], double (&tb)[D],
                                              int (&cc)[D], int (&base)[D]) {
  int z, y, x;
  node_zyx<D, B>(s, z, y, x);
  W = (D == 3) ? w[z] * w[y] * w[x] : w[y] * w[x];
#pragma unroll
  for (int d = 0; d < D; d++) {
    if (D == 3) {
      ta[d] = d == 2 ? w[y] : w[z];
      tb[d] = d == 0 ? w[y] : w[x];
    } else {
      ta[d] = d == 0 ? w[y] : w[x];
      tb[d] = 1.0;
    }
    cc[d] = coord<D, B>(d, z, y, x);
    base[d] = s - cc[d] * stride_of<D, B>(d);
  }
}

// ---------------------------------------------------------------------------
// Space-time predictor + volume integral
// ---------------------------------------------------------------------------
// One tile (CPB cells) of the predictor; called by predictor_kernel for
// tile blk.  Threads of invalid cells return at the end, after the last
// barrier of the tile.
//
// NCS: the PDE has a source and / or a non-conservative product
// (spaceTimePredictorNonlinear.cpph:104-176, 323-367): per time slice the
// node also publishes its state, reads the B-point lines through it for
// gradQ = dudx lQi / dx, adds w_t W dt (S - B(Q) gradQ) to its rhs and
// accumulates the time-averaged S, which the volume integral adds as W lShi
// (volumeIntegralNonlinear.cpph:204-222).
template <int D, int B, int V, int NP, typename R, bool CG, bool NCS = false, bool EUS = false>
__device__ __forceinline__ void predictor_tile(
    long blk, long ncells, const double* luh, const double* __restrict__ hcell,
    const double* __restrict__ dtp, const int* __restrict__ cinst, double* __restrict__ lduh,
    double* __restrict__ lQhbnd, double* __restrict__ lFhbnd, int* __restrict__ picard,
    const DevTables* __restrict__ T, PdeParams P, int maxIter, int adaptive, double* luh_upd, int dialect,
    int hdim) {
  using S = Shape<D, B, V, NP>;
  constexpr int ND = S::ND, NS = S::NS, CPB = S::CPB, NSP = S::NSP, THREADS_ = S::THREADS;
  static_assert(!NCS || V == kUserNVar, "user PDE hooks: the user PDE's variables");
  static_assert(CPB > 1 || THREADS_ / 64 <= CPB * NSP, "wave sums fit in red");
  // EUS (3D Euler, FP64): a slab holds the 12 distinct flux components per
  // node, in slot order, and the time-averaged flux accumulates the canonical
  // ones in registers (the reference's F_y[1] = rho^-1 q_y q_x, F_z[1], F_z[2]
  // are taken from F_x[2], F_x[3], F_y[3]: the same products in another order)
  static_assert(!EUS || (D == 3 && V == 5 && !NCS), "Euler flux slots: 3D, five variables");
  // slab values per node (EUS: 12 slots padded to 13, an odd stride in doubles,
  // so that the nodes a wave reads spread over the LDS banks)
  constexpr int SW = EUS ? 13 : D * V;
  constexpr int SLAB = NS * SW + (NCS ? NS * V : 0);  // one flux (+ state) slab, per cell
  constexpr int SLAB_BYTES = 2 * CPB * SLAB * int(sizeof(R));
  constexpr int END_BYTES = CPB * NS * (ND + D * V) * 8;  // time averages, FP64
  constexpr int BUF_BYTES = SLAB_BYTES > END_BYTES ? SLAB_BYTES : END_BYTES;
  // slab position of F_d[v] of node n; `keep` is false for the flux components
  // EUS stores once under another (d, v)
  auto sidx = [](int n, int d, int v) { return EUS ? n * SW + euler_slot(d, v) : (n * D + d) * V + v; };
  auto keep = [](int d, int v) { return !EUS || euler_canonical(d, v); };
  __shared__ __align__(16) unsigned char buf[BUF_BYTES];
  __shared__ double red[CPB * NSP];

  const int tid = threadIdx.x;
  const int cib = tid / NS;
  const int s = tid - cib * NS;
  const bool inslot = cib < CPB;            // thread maps to a cell slot of this block
  const long cell = blk * CPB + cib;
  const bool valid = inslot && (cell < ncells);
  const int cslot = inslot ? cib : 0;       // only dereferenced by in-slot threads

  int z, y, x;
  node_zyx<D, B>(s, z, y, x);
  const double* w = T->w;
  const double dt = dtp[(cinst && valid) ? cinst[cell] : 0];
  // 1/dx per direction: hcell[cell] (cubic cells, hdim = 0) or hcell[cell][d] (hdim = D)
  double invDx[D];
#pragma unroll
  for (int d = 0; d < D; d++) invDx[d] = 1.0 / (valid ? hcell[hdim ? cell * hdim + d : cell] : 1.0);

  double u[ND];
  R q[B][ND];
  if (valid) {
#pragma unroll
    for (int v = 0; v < ND; v++) u[v] = luh[(cell * NS + s) * ND + v];
  } else {
#pragma unroll
    for (int v = 0; v < ND; v++) u[v] = (v == 0) ? 1.0 : 0.0;
  }
#pragma unroll
  for (int t = 0; t < B; t++)
#pragma unroll
    for (int v = 0; v < ND; v++) q[t][v] = static_cast<R>(u[v]);

  // node weight (reference: w[z]*w[y]*w[x], 2D w[y]*w[x]) and the two
  // transverse weights per direction in the reference's multiplication order
  double W, ta[D], tb[D];
  int cc[D], base[D];
  node_geometry<D, B>(s, w, W, ta, tb, cc, base);
  const double iweight = 1.0 / W;
  // Kxi in LDS: a node reads its stiffness column Kxi[n][c_d] per use instead
  // of holding D B entries in registers (they would cost the second wave per SIMD)
  __shared__ R kxs[B * B];
  if (tid < B * B) kxs[tid] = tabK<R>(T, tid);  // read after the first barrier below

  // CERK initial guess of the optimised kernels
  // (KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_nonlinear_cpp.template:121-173,
  // templates/subtemplates/RK_loop.template:12-40): K1 = sum_d dudx_d F_d(luh) / dx,
  // lwh = luh - dt K1, K2 = the same of lwh, lQi(t) = luh - dt xi_t K1 - dt xi_t^2 (K2 - K1) / 2
  if constexpr (CG) {
    R* slab = reinterpret_cast<R*>(buf) + cslot * SLAB;
    R K1[V], K2[V], lw[ND];
#pragma unroll
    for (int v = 0; v < ND; v++) lw[v] = q[0][v];
#pragma unroll
    for (int stage = 0; stage < 2; stage++) {
      if (valid) {
        R F[D][V];
        pde_flux<D, R>(P, lw, F);
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int v = 0; v < V; v++)
            if (keep(d, v)) slab[sidx(s, d, v)] = F[d][v];
      }
      __syncthreads();
      R K[V];
#pragma unroll
      for (int v = 0; v < V; v++) K[v] = R(0);
      if (valid) {
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            R a = R(0);
#pragma unroll
            for (int n = 0; n < B; n++)
              a += slab[sidx(base[d] + n * stride_of<D, B>(d), d, v)] *
                   (static_cast<R>(invDx[d]) * tabD<R>(T, cc[d] * B + n));
            K[v] += a;
          }
      }
      __syncthreads();  // slab rewritten by the next stage / the Picard loop
#pragma unroll
      for (int v = 0; v < V; v++) {
        if (stage == 0) {
          K1[v] = K[v];
          lw[v] = static_cast<R>(u[v]) - static_cast<R>(dt) * K1[v];
        } else {
          K2[v] = K[v];
        }
      }
    }
    const R rdt = static_cast<R>(dt);
#pragma unroll
    for (int t = 0; t < B; t++) {
      const R xt = static_cast<R>(T->x[t]);
#pragma unroll
      for (int v = 0; v < V; v++)
        q[t][v] = static_cast<R>(u[v]) - (rdt * xt * K1[v]) - (R(0.5) * rdt * xt * xt * (K2[v] - K1[v]));
    }
  } else if constexpr (NCS) {
    // trivial guess: hide that the B slices start equal, or the compiler keeps
    // the copies folded and its schedule of the first iteration spills
#pragma unroll
    for (int t = 0; t < B; t++)
#pragma unroll
      for (int v = 0; v < ND; v++) asm volatile("" : "+v"(q[t][v]));
  }

  // time-averaged flux of the last executed iteration: in registers, or (FH_LDS,
  // FP64 3D order 5, trivial guess) in LDS [d][node][v], where its 30 registers
  // would cost spills at two waves per SIMD (64.6 -> 60.0 ms at 81^3; the CERK
  // form does not spill and is faster with registers); same operations
  constexpr bool FH_LDS = sizeof(R) == 8 && D == 3 && B == 6 && CPB == 1 && !NCS && !CG && !EUS;
  __shared__ double fhs[FH_LDS ? D * NS * V : 1];
  [[maybe_unused]] R fh[FH_LDS ? 1 : D][V];
  [[maybe_unused]] R sh[V];  // NCS: time-averaged source - ncp of the last executed iteration
#pragma unroll
  for (int d = 0; d < (FH_LDS ? 1 : D); d++)
#pragma unroll
    for (int v = 0; v < V; v++) fh[d][v] = R(0);
  if constexpr (NCS) {
#pragma unroll
    for (int v = 0; v < V; v++) sh[v] = R(0);
  }
  bool active = valid;
  int iters = maxIter;

  // Reduction of the squared update of iteration iter over the cell and the
  // adaptive rule (stop below 1e-7); true when no cell of the block iterates further.
  auto finished = [&](int iter, double sq) -> bool {
    if (!adaptive) {
      __syncthreads();
      return false;
    }
    double res;
    if constexpr (CPB == 1) {
      // one cell per workgroup: butterfly sum in each wave (every lane ends
      // with the same bits: a + b == b + a), then the wave sums in order --
      // deterministic, one barrier instead of the tree's log2(NSP) + 1
      double v = sq;  // 0 on lanes past the cell
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
      if ((tid & 63) == 0) red[tid >> 6] = v;
      __syncthreads();
      res = 0.0;
#pragma unroll
      for (int i = 0; i < THREADS_ / 64; i++) res += red[i];
    } else {
      // deterministic per-cell tree reduction of the squared update
      if (inslot) {
        red[cib * NSP + s] = sq;
        if (s < NSP - NS) red[cib * NSP + NS + s] = 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int stride = NSP / 2; stride > 0; stride >>= 1) {
        if (inslot && s < stride) red[cib * NSP + s] += red[cib * NSP + s + stride];
        __syncthreads();
      }
      res = inslot ? red[cib * NSP] : 0.0;
    }
    const double tol = 1e-7;
    if (active && res < tol * tol) {
      active = false;
      // reference return conventions; optimised kernels: min(iter + 1, MaxIterations)
      iters = dialect ? (iter + 1 < maxIter ? iter + 1 : maxIter) : ((D == 3) ? iter : maxIter + 2);
    }
    // one cell per workgroup: every thread holds the same res, so the decision
    // is uniform without a second barrier (red is rewritten only after the
    // next iteration's time-slice barriers)
    if constexpr (CPB == 1) return !(blk < ncells && res >= tol * tol);  // the block's cell is blk
    return !__syncthreads_or(active);
  };

  // The time solve lQi_new[t'] = sum_t iK1[t'][t] rhs[t] / W: the 1/W is folded
  // into the right-hand side's factors -- the stiffness factor of the node per
  // direction, (ta tb dt / dx)_d / W (a time slice scales it by w_t), and the
  // F0 term's W F0[t] / W
  R cd[D];
#pragma unroll
  for (int d = 0; d < D; d++) cd[d] = static_cast<R>(ta[d] * tb[d] * dt * invDx[d] * iweight);
  int first = 0;
  if constexpr (!CG && !NCS) {
    // Trivial guess: the B time slices of the first iterate all equal luh, so
    // iteration 0 forms their flux, its time average and the stiffness sum once:
    // rhs[t] = W F0[t] luh - w[t] sum_d sum_n (ta tb dt / dx)_d Kxi[n][c_d] F_d(luh).
    // Its squared update is taken against luh, so the B copies of luh are never held.
    if (maxIter > 0) {
      R acc[B][V];  // time slice 0 assigns, the others accumulate (active lanes only)
      R* slab = reinterpret_cast<R*>(buf) + cslot * SLAB;
      if (active) {
        R F[D][V];
        pde_flux<D, R>(P, q[0], F);
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            R a = R(0);
#pragma unroll
            for (int t = 0; t < B; t++) a += F[d][v] * static_cast<R>(w[t]);
            if constexpr (FH_LDS)
              fhs[(d * NS + s) * V + v] = a;
            else if (keep(d, v))
              fh[d][v] = a;
            if (keep(d, v)) slab[sidx(s, d, v)] = F[d][v];
          }
      }
      __syncthreads();
      double sq = 0.0;
      if (active) {
        R g[V];
#pragma unroll
        for (int v = 0; v < V; v++) g[v] = R(0);
#pragma unroll
        for (int d = 0; d < D; d++) {
          const R upd = cd[d];
#pragma unroll
          for (int n = 0; n < B; n++) {
            const R ukc = upd * kxs[n * B + cc[d]];
#pragma unroll
            for (int v = 0; v < V; v++) g[v] += ukc * slab[sidx(base[d] + n * stride_of<D, B>(d), d, v)];
          }
        }
#pragma unroll
        for (int t = 0; t < B; t++) {
          const R rw = static_cast<R>(W * T->F0[t] * iweight);
          const R wt = static_cast<R>(w[t]);
#pragma unroll
          for (int v = 0; v < V; v++) {
            const R r = rw * static_cast<R>(u[v]) - wt * g[v];
#pragma unroll
            for (int tp = 0; tp < B; tp++) {
              if (t == 0)
                acc[tp][v] = r * tabI<R>(T, tp * B + t);
              else
                acc[tp][v] += r * tabI<R>(T, tp * B + t);
            }
          }
        }
#pragma unroll
        for (int t = 0; t < B; t++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            const double dq = static_cast<double>(acc[t][v]) - static_cast<double>(static_cast<R>(u[v]));
            sq += dq * dq;
            q[t][v] = acc[t][v];
          }
      }
      first = finished(0, sq) ? maxIter : 1;
    }
  }

  for (int iter = first; iter < maxIter; iter++) {
    // the new iterate and the time-averaged flux of this iteration: time slice 0
    // assigns, the others accumulate (active lanes only)
    R acc[B][V];
    if constexpr (NCS) {
      if (active) {
#pragma unroll
        for (int v = 0; v < V; v++) sh[v] = R(0);
      }
    }

#pragma unroll
    for (int t = 0; t < B; t++) {
      R* slab = reinterpret_cast<R*>(buf) + ((t & 1) * CPB + cslot) * SLAB;
      if (active) {
        R F[D][V];
        pde_flux<D, R>(P, q[t], F);
        const R wt = static_cast<R>(w[t]);
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            if constexpr (FH_LDS) {
              if (t == 0)
                fhs[(d * NS + s) * V + v] = F[d][v] * wt;  // own node: no other lane touches it
              else
                fhs[(d * NS + s) * V + v] += F[d][v] * wt;
            } else if (keep(d, v)) {
              if (t == 0)
                fh[d][v] = F[d][v] * wt;
              else
                fh[d][v] += F[d][v] * wt;
            }
            if (keep(d, v)) slab[sidx(s, d, v)] = F[d][v];
          }
        if constexpr (NCS) {
#pragma unroll
          for (int v = 0; v < V; v++) slab[NS * SW + s * V + v] = q[t][v];
        }
      }
      __syncthreads();
      if (active) {
        // rhs0 / W = F0[t] luh (as W F0[t] / W), then minus the x, y, z stiffness terms
        const R rw = static_cast<R>(W * T->F0[t] * iweight);
        R r[V];
#pragma unroll
        for (int v = 0; v < V; v++) r[v] = rw * static_cast<R>(u[v]);
#pragma unroll
        for (int d = 0; d < D; d++) {
          const R upd = static_cast<R>(w[t]) * cd[d];
#pragma unroll
          for (int n = 0; n < B; n++) {
            // Kxi[n][c_d] times the update factor, once per (d, n): one FMA per
            // flux value instead of a multiply and an FMA
            const R ukc = upd * kxs[n * B + cc[d]];
#pragma unroll
            for (int v = 0; v < V; v++) {
              const R f = slab[sidx(base[d] + n * stride_of<D, B>(d), d, v)];
              r[v] -= ukc * f;
            }
          }
        }
        if constexpr (NCS) {
          // gradQ of time slice t (:113-176), then rhs += w_t W dt (S - ncp) (:323-367)
          const R* qs = slab + NS * SW;
          R gq[D][V];
#pragma unroll
          for (int d = 0; d < D; d++)
#pragma unroll
            for (int v = 0; v < V; v++) {
              R g = R(0);
#pragma unroll
              for (int n = 0; n < B; n++)
                g += static_cast<R>(invDx[d]) * qs[(base[d] + n * stride_of<D, B>(d)) * V + v] *
                     tabD<R>(T, cc[d] * B + n);
              gq[d][v] = g;
            }
          R Sv[V];
          pde_ncs<D, R>(P, q[t], gq, Sv);
          const R us = static_cast<R>(w[t] * W * dt * iweight);
          const R wt = static_cast<R>(w[t]);
#pragma unroll
          for (int v = 0; v < V; v++) {
            r[v] += us * Sv[v];
            sh[v] += Sv[v] * wt;
          }
        }
        // (K1)^-1 in time: lQi_new[t'] += rhs[t] / W * iK1[t'][t] (r holds rhs / W)
#pragma unroll
        for (int tp = 0; tp < B; tp++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            if (t == 0)
              acc[tp][v] = r[v] * tabI<R>(T, tp * B + t);
            else
              acc[tp][v] += r[v] * tabI<R>(T, tp * B + t);
          }
      }
    }
    // residual and update of lQi (variables only; parameters are carried)
    double sq = 0.0;
    if (active) {
#pragma unroll
      for (int t = 0; t < B; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          const double dq = static_cast<double>(acc[t][v]) - static_cast<double>(q[t][v]);
          sq += dq * dq;
          q[t][v] = acc[t][v];
        }
    }
    if (finished(iter, sq)) break;
  }

  // ---- time averages -> LDS (FP64) ----
  double* Qh = reinterpret_cast<double*>(buf) + cslot * NS * (ND + D * V);
  double* Fh = FH_LDS ? fhs : Qh + NS * ND;  // [d][node][v]
  __syncthreads();  // slabs no longer read
  if (valid) {
#pragma unroll
    for (int v = 0; v < ND; v++) {
      double a = 0.0;
#pragma unroll
      for (int t = 0; t < B; t++) a += static_cast<double>(q[t][v]) * w[t];
      Qh[s * ND + v] = a;
    }
    if constexpr (!FH_LDS) {
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          // EUS: the component stored once is read from its canonical (d, v)
          const int dd = keep(d, v) ? d : v - 1, vv = keep(d, v) ? v : d + 1;
          Fh[(d * NS + s) * V + v] = static_cast<double>(fh[dd][vv]);
        }
    }
  }
  __syncthreads();
  if (!valid) return;

  // ---- volume integral and extrapolation ----
  double shd[V];
  if constexpr (NCS) {
#pragma unroll
    for (int v = 0; v < V; v++) shd[v] = static_cast<double>(sh[v]);
  }
  // the per-node constants again, from a node index the optimiser cannot
  // equate with s: the loop's copies end with the loop instead of being held
  // (spilled at two waves per SIMD) across it
  int s2 = s;
  asm volatile("" : "+v"(s2));
  double W2, ta2[D], tb2[D], invDx2[D];
  int cc2[D], base2[D];
  node_geometry<D, B>(s2, w, W2, ta2, tb2, cc2, base2);
#pragma unroll
  for (int d = 0; d < D; d++) invDx2[d] = 1.0 / hcell[hdim ? cell * hdim + d : cell];
  const double dt2 = dtp[cinst ? cinst[cell] : 0];
  node_volume_and_faces<D, B, V, NP>(Qh, Fh, s2, cell, ta2, tb2, cc2, base2, invDx2, W2, dt2, iters,
                                     NCS ? shd : nullptr, T, lduh, lQhbnd, lFhbnd, picard, luh_upd);
}

// Predictor + volume integral, one workgroup per tile of CPB cells.  Two
// waves per SIMD (at most 256 registers per lane) for the conservative forms;
// the source / NCP form keeps the extra state of its gradient and source terms
// and runs at one.
template <int D, int B, int V, int NP, typename R, bool CG, bool NCS = false, bool EUS = false>
__global__ __launch_bounds__((Shape<D, B, V, NP>::THREADS), NCS ? 1 : 2) void predictor_kernel(
    long ncells, long ntiles, const double* luh, const double* __restrict__ hcell,
    const double* __restrict__ dtp, const int* __restrict__ cinst, double* __restrict__ lduh,
    double* __restrict__ lQhbnd, double* __restrict__ lFhbnd, int* __restrict__ picard,
    const DevTables* __restrict__ T, PdeParams P, int maxIter, int adaptive, double* luh_upd, int dialect,
    int hdim) {
  predictor_tile<D, B, V, NP, R, CG, NCS, EUS>(blockIdx.x, ncells, luh, hcell, dtp, cinst, lduh, lQhbnd, lFhbnd,
                                               picard, T, P, maxIter, adaptive, luh_upd, dialect, hdim);
}

// Persistent form: a grid smaller than ntiles loops over the tiles.
template <int D, int B, int V, int NP, typename R, bool CG, bool EUS = false>
__global__ __launch_bounds__((Shape<D, B, V, NP>::THREADS), 2) void predictor_persistent_kernel(
    long ncells, long ntiles, const double* luh, const double* __restrict__ hcell,
    const double* __restrict__ dtp, const int* __restrict__ cinst, double* __restrict__ lduh,
    double* __restrict__ lQhbnd, double* __restrict__ lFhbnd, int* __restrict__ picard,
    const DevTables* __restrict__ T, PdeParams P, int maxIter, int adaptive, double* luh_upd, int dialect,
    int hdim) {
  for (long blk = blockIdx.x; blk < ntiles; blk += gridDim.x) {
    predictor_tile<D, B, V, NP, R, CG, false, EUS>(blk, ncells, luh, hcell, dtp, cinst, lduh, lQhbnd, lFhbnd,
                                                   picard, T, P, maxIter, adaptive, luh_upd, dialect, hdim);
    __syncthreads();  // LDS of this tile fully consumed before the next one
  }
}

// ---------------------------------------------------------------------------
// Rusanov flux, one wavefront per face
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Face entry (x, y, d, w): left cell x (its face 2d+1), right cell y (its
// face 2d); with ghost faces given, w & 2 (w & 1): the left (right) side is
// ghost face x (y) instead.
template <int D, int B, int V, int NP>
__global__ __launch_bounds__(256) void riemann_kernel(long nfaces, const int4* __restrict__ faces,
                                                      const double* __restrict__ lQhbnd,
                                                      double* __restrict__ lFhbnd, const double* __restrict__ gq,
                                                      double* __restrict__ gf, const DevTables* __restrict__ T,
                                                      PdeParams P) {
  using S = Shape<D, B, V, NP>;
  constexpr int ND = S::ND, NF = S::NF;
  const int lane = threadIdx.x & 63;
  const long f = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (f >= nfaces) return;  // uniform per wavefront; no block barriers below
  const int4 fc = faces[f];
  const int d = fc.z;
  const bool gl = gq && (fc.w & 2) != 0, gr = gq && (fc.w & 1) != 0;  // w is ignored without ghost faces
  const double* QL = gl ? gq + (long)fc.x * NF * ND : lQhbnd + ((long)fc.x * 2 * D + 2 * d + 1) * NF * ND;
  const double* QR = gr ? gq + (long)fc.y * NF * ND : lQhbnd + ((long)fc.y * 2 * D + 2 * d) * NF * ND;
  double* FL = gl ? gf + (long)fc.x * NF * V : lFhbnd + ((long)fc.x * 2 * D + 2 * d + 1) * NF * V;
  double* FR = gr ? gf + (long)fc.y * NF * V : lFhbnd + ((long)fc.y * 2 * D + 2 * d) * NF * V;
  const double* w = T->w;
  double aL[ND], aR[ND];
#pragma unroll
  for (int k = 0; k < ND; k++) aL[k] = aR[k] = 0.0;
  for (int j = lane; j < NF; j += 64) {
    const double wj = (D == 3) ? w[j / B] * w[j % B] : w[j];
#pragma unroll
    for (int k = 0; k < ND; k++) {
      aL[k] += wj * QL[j * ND + k];
      aR[k] += wj * QR[j * ND + k];
    }
  }
#pragma unroll
  for (int k = 0; k < ND; k++) {
    aL[k] = wave_sum(aL[k]);
    aR[k] = wave_sum(aR[k]);
  }
  const double smax = fmax(pde_max_abs_eig<D, V>(P, aL, d), pde_max_abs_eig<D, V>(P, aR, d));
  if constexpr (V == kUserNVar) {
    if (pde_has_ncp(P)) {
      // fluctuations (riemannSolverNonlinear.cpph:95-120): the NCP of the node
      // average with gradQ[d] = QR - QL; FR = F - ncp/2, FL = F + ncp/2
      for (int j = lane; j < NF; j += 64) {
        double gq[D][V], qa[V], ncp[V];
#pragma unroll
        for (int e = 0; e < D; e++)
#pragma unroll
          for (int k = 0; k < V; k++) gq[e][k] = 0.0;
#pragma unroll
        for (int k = 0; k < V; k++) {
          gq[d][k] = QR[j * ND + k] - QL[j * ND + k];
          qa[k] = 0.5 * (QR[j * ND + k] + QL[j * ND + k]);
        }
        ExaUserPde::ncp<D, double>(P.u, qa, gq, ncp);
#pragma unroll
        for (int k = 0; k < V; k++) {
          const double fl = FL[j * V + k], fr = FR[j * V + k];
          const double ql = QL[j * ND + k], qr = QR[j * ND + k];
          const double flux = 0.5 * (fr + fl) - 0.5 * smax * (qr - ql);
          FR[j * V + k] = flux - 0.5 * ncp[k];
          FL[j * V + k] = flux + 0.5 * ncp[k];
        }
      }
      return;
    }
  }
  for (int j = lane; j < NF; j += 64) {
#pragma unroll
    for (int k = 0; k < V; k++) {
      const double fl = FL[j * V + k], fr = FR[j * V + k];
      const double ql = QL[j * ND + k], qr = QR[j * ND + k];
      const double flux = 0.5 * (fr + fl) - 0.5 * smax * (qr - ql);
      FL[j * V + k] = flux;
      FR[j * V + k] = flux;
    }
  }
}

// ---------------------------------------------------------------------------
// Face integrals + solution update + CFL
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_min_pos_double(double* addr, double v) {
  // positive IEEE doubles order like their bit patterns
  atomicMin(reinterpret_cast<unsigned long long*>(addr), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// A node's stable dt is CFL*PNPM / den with den = sum_d max_v |lambda_v(Q, d)| / dx
// (stableTimeStepSize.cpph:14-62); node_cfl_den returns den.  For Euler the
// direction-independent 1/rho, p and c are computed once; the per-direction
// maximum keeps pde_max_abs_eig's expressions (same bits).
template <int D, int B, int V, int NP>
__device__ __forceinline__ double node_cfl_den(const double* __restrict__ Q, const PdeParams& P,
                                               const double (&invDx)[D]) {
  double den = 0.0;
  if constexpr (V == 5) {
    if (P.id == kPdeEuler) {
      constexpr double gamma = 1.4;
      const double irho = 1. / Q[0];
      const double j2 = (D == 3) ? Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3] : Q[1] * Q[1] + Q[2] * Q[2];
      const double p = (gamma - 1) * (Q[4] - 0.5 * irho * j2);
      const double c = sqrt(gamma * p * irho);
#pragma unroll
      for (int d = 0; d < D; d++) {
        const double u_n = Q[d + 1] * irho;
        const double l0 = fabs(u_n - c), l1 = fabs(u_n), l4 = fabs(u_n + c);
        den += fmax(fmax(l0, l1), l4) * invDx[d];
      }
      return den;
    }
  }
#pragma unroll
  for (int d = 0; d < D; d++) den += pde_max_abs_eig<D, V>(P, Q, d) * invDx[d];
  return den;
}

template <int D, int B, int V, int NP>
__device__ __forceinline__ double node_cfl_dt(const double* __restrict__ Q, const PdeParams& P,
                                              const double (&invDx)[D], double cflpnpm) {
  return cflpnpm / node_cfl_den<D, B, V, NP>(Q, P, invDx);
}

// Workgroup maximum (fmax: NaN lanes are ignored, as fmin ignores them in block_min).
template <int THREADS>
__device__ __forceinline__ double block_max(double v, double* sred) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sred[wid] = v;
  __syncthreads();
  double r = sred[0];
#pragma unroll
  for (int i = 1; i < THREADS / 64; i++) r = fmax(r, sred[i]);
  return r;
}

template <int THREADS>
__device__ __forceinline__ double block_min(double v, double* sred) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sred[wid] = v;
  __syncthreads();
  double r = sred[0];
#pragma unroll
  for (int i = 1; i < THREADS / 64; i++) r = fmin(r, sred[i]);
  return r;
}

template <int D, int B, int V, int NP>
__device__ __forceinline__ void corrector_tile(
    long blk, long ncells, double* __restrict__ luh, const double* __restrict__ lduh,
    const double* __restrict__ lFhbnd, const double* __restrict__ hcell, const double* __restrict__ dtp,
    const int* __restrict__ cinst, double* __restrict__ dt_next, const DevTables* __restrict__ T, PdeParams P,
    double cflpnpm, int do_update, double* sred, double* __restrict__ slots, int hdim) {
  using S = Shape<D, B, V, NP>;
  constexpr int ND = S::ND, NS = S::NS, NF = S::NF, CPB = S::CPB, THREADS = S::THREADS;
  const int tid = threadIdx.x;
  const int cib = tid / NS;
  const int s = tid - cib * NS;
  const long cell = blk * CPB + cib;
  const bool valid = (cib < CPB) && (cell < ncells);
  const int inst = (cinst && valid) ? cinst[cell] : 0;
  double mydt = DBL_MAX;
  if (valid) {
    int z, y, x;
    node_zyx<D, B>(s, z, y, x);
    const double* w = T->w;
    double invDx[D];  // hcell[cell] (cubic cells, hdim = 0) or hcell[cell][d] (hdim = D)
#pragma unroll
    for (int d = 0; d < D; d++) invDx[d] = 1.0 / hcell[hdim ? cell * hdim + d : cell];
    double Q[ND];
#pragma unroll
    for (int v = 0; v < ND; v++) Q[v] = luh[(cell * NS + s) * ND + v];
    if (do_update) {
      const double dt = dtp[inst];
      double du[V];
#pragma unroll
      for (int v = 0; v < V; v++) du[v] = lduh[(cell * NS + s) * V + v];
#pragma unroll
      for (int d = 0; d < D; d++) {
        const int c = coord<D, B>(d, z, y, x);
        const int fn = face_node<D, B>(d, z, y, x);
        const double tw = transverse_weight<D>(w, d, z, y, x);
#pragma unroll
        for (int o = 0; o < 2; o++) {
          const double scaling = invDx[d] * (2.0 * o - 1.0);
          const double sw = tw * scaling;
          const double fcoef = o ? T->F1[c] : T->F0[c];
          const double* F = lFhbnd + ((cell * 2 * D + 2 * d + o) * NF + fn) * V;
#pragma unroll
          for (int v = 0; v < V; v++) du[v] -= sw * fcoef * F[v];
        }
      }
      const double Wn = (D == 3) ? w[z] * w[y] * w[x] : w[y] * w[x];
      const double upd = dt / Wn;
#pragma unroll
      for (int v = 0; v < V; v++) {
        Q[v] = Q[v] + du[v] * upd;
        luh[(cell * NS + s) * ND + v] = Q[v];
      }
    }
    mydt = node_cfl_dt<D, B, V, NP>(Q, P, invDx, cflpnpm);
  }
  // cells of one instance are contiguous: a tile whose first and last cell
  // share the instance reduces once, otherwise every cell merges its own dt
  const long first = blk * CPB, last = (first + CPB - 1 < ncells) ? first + CPB - 1 : ncells - 1;
  const bool uniform = !cinst || cinst[first] == cinst[last];
  if (uniform) {
    const double bmin = block_min<THREADS>(mydt, sred);
    double* dst = (slots && !cinst) ? slots + (blk % kDtSlots) * kDtSlotStride : dt_next + (cinst ? cinst[first] : 0);
    if (tid == 0 && bmin < DBL_MAX) atomic_min_pos_double(dst, bmin);
  } else if (valid && mydt < DBL_MAX) {
    atomic_min_pos_double(dt_next + inst, mydt);
  }
}

template <int D, int B, int V, int NP>
__global__ __launch_bounds__((Shape<D, B, V, NP>::THREADS)) void corrector_kernel(
    long ncells, long ntiles, double* __restrict__ luh, const double* __restrict__ lduh,
    const double* __restrict__ lFhbnd, const double* __restrict__ hcell, const double* __restrict__ dtp,
    const int* __restrict__ cinst, double* __restrict__ dt_next, const DevTables* __restrict__ T, PdeParams P,
    double cflpnpm, int do_update, double* __restrict__ slots, int hdim) {
  __shared__ double sred[Shape<D, B, V, NP>::THREADS / 64];
  corrector_tile<D, B, V, NP>(blockIdx.x, ncells, luh, lduh, lFhbnd, hcell, dtp, cinst, dt_next, T, P, cflpnpm,
                              do_update, sred, slots, hdim);
}

template <int D, int B, int V, int NP>
__global__ __launch_bounds__((Shape<D, B, V, NP>::THREADS)) void corrector_persistent_kernel(
    long ncells, long ntiles, double* __restrict__ luh, const double* __restrict__ lduh,
    const double* __restrict__ lFhbnd, const double* __restrict__ hcell, const double* __restrict__ dtp,
    const int* __restrict__ cinst, double* __restrict__ dt_next, const DevTables* __restrict__ T, PdeParams P,
    double cflpnpm, int do_update, double* __restrict__ slots, int hdim) {
  __shared__ double sred[Shape<D, B, V, NP>::THREADS / 64];
  for (long blk = blockIdx.x; blk < ntiles; blk += gridDim.x) {
    corrector_tile<D, B, V, NP>(blk, ncells, luh, lduh, lFhbnd, hcell, dtp, cinst, dt_next, T, P, cflpnpm,
                                do_update, sred, slots, hdim);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused Riemann + face integrals + update + CFL for a uniform periodic n^D
// mesh (cells x fastest).  Every cell solves the Rusanov problems of its own
// 2D faces from both sides' extrapolated data (riemannSolverNonlinear.cpph:37-128)
// and keeps the flux; the two cells of a face evaluate it with the same
// operations in the same order, so both see the same flux bits (conservation
// as with one solve per face).  The volume part of the update is already in
// luh (predictor with LaunchOpts::luh_update).
//   phase 1: w_fn * Q of both sides per face node -> LDS
//   phase 2: face averages, one thread per (face, side, variable), fixed order
//   phase 3: flux per face node -> LDS
//   phase 4: face integrals + update + stable dt per node
// ---------------------------------------------------------------------------
template <int D, int B, int V, int NP>
struct FusedShape {
  using S = Shape<D, B, V, NP>;
  static constexpr int NTASK = 2 * D * S::NF;  // face nodes per cell
  static constexpr int RED = NTASK * 2 * S::ND;  // partial products per cell
  static constexpr int AVG = 2 * D * 2 * S::ND;  // averages per cell
  static constexpr int NQ = S::NF >= 16 ? 4 : 1;  // partial sums of a face average (large faces)
  static constexpr int PART = NQ > 1 ? AVG * NQ : 0;  // partial sums per cell
  static constexpr int SMAX = 2 * D;                   // Rusanov speed per face
  static constexpr int LDS_BYTES = S::CPB * (RED + AVG + PART + SMAX) * 8;
  static constexpr bool ok = LDS_BYTES <= 60 * 1024;
};

// Integer coordinates of a cell of a uniform n^D mesh (x fastest).
struct CellXYZ {
  long ix, iy, iz;
};
template <int D>
__device__ __forceinline__ CellXYZ cell_xyz(long cell, int n) {
  return {cell % n, (cell / n) % n, (D == 3) ? cell / (long(n) * n) : 0};
}

// Periodic neighbour of the cell at c in direction d (dir = +-1): the
// coordinate steps with a wrap at the domain ends, no division.
template <int D>
__device__ __forceinline__ long periodic_neighbour(const CellXYZ& c, int n, int d, int dir) {
  auto step = [&](long i) { return (i + dir == n) ? 0 : (i + dir < 0 ? n - 1 : i + dir); };
  const long ix = d == 0 ? step(c.ix) : c.ix, iy = d == 1 ? step(c.iy) : c.iy, iz = d == 2 ? step(c.iz) : c.iz;
  return (D == 3) ? (iz * n + iy) * n + ix : iy * n + ix;
}

// Ghost face of face f = 2d + o of the cell at c when that face lies on a
// non-periodic domain face (bit f of bcmask), else -1.  Index f * n^(D-1) +
// transverse cell index: (iz, iy) for d = 0, (iz, ix) for d = 1, (iy, ix) for
// d = 2; in 2D iy for d = 0, ix for d = 1.
template <int D>
__device__ __forceinline__ long ghost_face(const CellXYZ& c, int n, int f, unsigned bcmask) {
  if (!((bcmask >> f) & 1u)) return -1;
  const int d = f >> 1;
  const long id = (d == 0) ? c.ix : (d == 1 ? c.iy : c.iz);
  if (id != ((f & 1) ? n - 1 : 0)) return -1;
  long tr;
  if (D == 3)
    tr = (d == 0) ? c.iz * n + c.iy : (d == 1 ? c.iz * n + c.ix : c.iy * n + c.ix);
  else
    tr = (d == 0) ? c.iy : c.ix;
  return long(f) * ((D == 3) ? long(n) * n : long(n)) + tr;
}

// Value k of face node fn in the face block at p (NF nodes, M values per node):
// the reference layout [node][value] for FP64 faces, [value][node] for the
// FP32 faces of the FP32 step (coalesced per value across a face's nodes).
template <typename TF>
__device__ __forceinline__ double face_val(const TF* p, int fn, int k, int NF, int M) {
  if constexpr (sizeof(TF) == 4)
    return static_cast<double>(p[k * NF + fn]);
  else
    return p[fn * M + k];
}

template <int D, int B, int V, int NP, typename TF>
__global__ __launch_bounds__((Shape<D, B, V, NP>::THREADS)) void corrector_riemann_kernel(
    long ncells, int n, double* __restrict__ luh, const TF* __restrict__ lQhbnd,
    const TF* __restrict__ lFhbnd, const double* __restrict__ hcell, const double* __restrict__ dtp,
    double* __restrict__ dt_next, const DevTables* __restrict__ T, PdeParams P, double cflpnpm,
    const TF* __restrict__ gq, const TF* __restrict__ gf, unsigned bcmask, double* __restrict__ slots,
    long cell0) {
  using S = Shape<D, B, V, NP>;
  using FS = FusedShape<D, B, V, NP>;
  constexpr int ND = S::ND, NS = S::NS, NF = S::NF, CPB = S::CPB, THREADS = S::THREADS;
  constexpr int NTASK = FS::NTASK;
  // GRID3 (3D, one cell per workgroup): the grid is (n, n, planes) and block
  // (x, y, z) is the cell (x, y, cell0 + z) -- cell0 is then the first z plane;
  // otherwise the launch covers cells [cell0, ncells), CPB per workgroup
  constexpr bool GRID3 = D == 3 && CPB == 1;
  __shared__ double sred[THREADS / 64];
  __shared__ double sbuf[CPB * (FS::RED + FS::AVG + FS::PART + FS::SMAX)];
  const int tid = threadIdx.x;
  const int cib = tid / NS;
  const int s = tid - cib * NS;
  // the cell's coordinates: GRID3 from the block index (no division)
  const long cg = cell0 + long(blockIdx.x) * CPB + cib;
  const CellXYZ cxyz = GRID3 ? CellXYZ{long(blockIdx.x), long(blockIdx.y), cell0 + long(blockIdx.z)}
                             : cell_xyz<D>(cg, n);
  const long cell = GRID3 ? (cxyz.iz * n + cxyz.iy) * n + cxyz.ix : cg;
  const bool valid = (cib < CPB) && (cell < ncells);
  double* red = sbuf + (valid ? cib : 0) * (FS::RED + FS::AVG + FS::PART + FS::SMAX);  // [task][side][ND]; later [task][V]
  double* avg = red + FS::RED;                                                         // [face][side][ND]
  double* part = avg + FS::AVG;                                                        // [face][side][ND][quarter]
  double* sm = part + FS::PART;                                                        // [face]
  const double* w = T->w;
  // One face-node task per thread (3D order 5, 2D order >= 3): the thread
  // loads both sides' states and fluxes once and keeps them for phase 3.
  constexpr bool ONE = NTASK <= NS;
  // kept in the face storage type (FP32 faces: half the registers), widened at use
  TF rQo[ONE ? ND : 1], rQn[ONE ? ND : 1], rFo[ONE ? V : 1], rFn[ONE ? V : 1];
  // the node's solution, loaded up front so that its latency overlaps the face phases
  double Q[ND];
  if (valid) {
#pragma unroll
    for (int v = 0; v < ND; v++) Q[v] = luh[(cell * NS + s) * ND + v];
  }
  // phase 1
  if constexpr (ONE) {
    if (valid && s < NTASK) {
      const int task = s;
      const int f = task / NF, fn = task - f * NF;
      // the other side of face f: face block nb * 2D + (f ^ 1) of the neighbour
      // (ob >= 0) or ghost block g (ob = -1 - g); GRID3: the cell is the
      // workgroup's, so the six are uniform and a lane selects its face's
      auto other = [&](int ff) -> long {
        const long g = ghost_face<D>(cxyz, n, ff, bcmask);
        return g >= 0 ? -1 - g : periodic_neighbour<D>(cxyz, n, ff >> 1, (ff & 1) ? 1 : -1) * 2 * D + (ff ^ 1);
      };
      long ob;
      if constexpr (GRID3) {
        ob = other(0);
#pragma unroll
        for (int ff = 1; ff < 2 * D; ff++) {
          const long b = other(ff);
          ob = f == ff ? b : ob;
        }
      } else {
        ob = other(f);
      }
      // face blocks (NF nodes each) of this cell and of the neighbour or ghost
      const TF* qo = lQhbnd + (cell * 2 * D + f) * NF * ND;
      const TF* qn = ob < 0 ? gq + (-1 - ob) * NF * ND : lQhbnd + ob * NF * ND;
      const TF* fo = lFhbnd + (cell * 2 * D + f) * NF * V;
      const TF* fnb = ob < 0 ? gf + (-1 - ob) * NF * V : lFhbnd + ob * NF * V;
      // this cell's side and the other side (the left state is the other
      // side's for o = 0, this cell's for o = 1)
#pragma unroll
      for (int k = 0; k < ND; k++) {
        rQo[k] = static_cast<TF>(face_val(qo, fn, k, NF, ND));
        rQn[k] = static_cast<TF>(face_val(qn, fn, k, NF, ND));
      }
#pragma unroll
      for (int k = 0; k < V; k++) {
        rFo[k] = static_cast<TF>(face_val(fo, fn, k, NF, V));
        rFn[k] = static_cast<TF>(face_val(fnb, fn, k, NF, V));
      }
      // the face averages' partial products per side; the Rusanov speed is the
      // larger of the two sides' (fmax: the side order does not matter)
      const double wj = (D == 3) ? w[fn / B] * w[fn % B] : w[fn];
#pragma unroll
      for (int k = 0; k < ND; k++) {
        red[(task * 2 + 0) * ND + k] = wj * static_cast<double>(rQo[k]);
        red[(task * 2 + 1) * ND + k] = wj * static_cast<double>(rQn[k]);
      }
    }
  } else if (valid) {
    for (int task = s; task < NTASK; task += NS) {
      const int f = task / NF, fn = task - f * NF, o = f & 1;
      const long g = ghost_face<D>(cxyz, n, f, bcmask);
      const TF* qo = lQhbnd + (cell * 2 * D + f) * NF * ND;
      const TF* qn = g >= 0 ? gq + g * NF * ND
                            : lQhbnd + (periodic_neighbour<D>(cxyz, n, f >> 1, o ? 1 : -1) * 2 * D + (f ^ 1)) * NF * ND;
      const TF* QL = o ? qo : qn;
      const TF* QR = o ? qn : qo;
      const double wj = (D == 3) ? w[fn / B] * w[fn % B] : w[fn];
#pragma unroll
      for (int k = 0; k < ND; k++) {
        red[(task * 2 + 0) * ND + k] = wj * face_val(QL, fn, k, NF, ND);
        red[(task * 2 + 1) * ND + k] = wj * face_val(QR, fn, k, NF, ND);
      }
    }
  }
  __syncthreads();
  // phase 2: face averages in a fixed order that depends only on the face
  // (quarter sums over consecutive face nodes, then ((p0 + p1) + p2) + p3)
  constexpr int CH = (NF + FS::NQ - 1) / FS::NQ;
  if constexpr (FS::NQ == 1) {
    if (valid) {
      for (int q = s; q < FS::AVG; q += NS) {
        const int f = q / (2 * ND), side = (q / ND) & 1, k = q % ND;
        double acc = 0.0;
        for (int fn = 0; fn < NF; fn++) acc += red[((f * NF + fn) * 2 + side) * ND + k];
        avg[q] = acc;
      }
    }
    __syncthreads();
  } else {
  if (valid) {
    for (int q = s; q < FS::PART; q += NS) {
      const int item = q / FS::NQ, quarter = q - item * FS::NQ;
      const int f = item / (2 * ND), side = (item / ND) & 1, k = item % ND;
      const int lo = quarter * CH, hi = lo + CH < NF ? lo + CH : NF;
      double acc = 0.0;
      for (int fn = lo; fn < hi; fn++) acc += red[((f * NF + fn) * 2 + side) * ND + k];
      part[q] = acc;
    }
  }
  __syncthreads();
  if (valid) {
    for (int q = s; q < FS::AVG; q += NS) {
      const double* pq = part + q * FS::NQ;
      double acc = pq[0];
#pragma unroll
      for (int i = 1; i < FS::NQ; i++) acc += pq[i];
      avg[q] = acc;  // q = (f * 2 + side) * ND + k
    }
  }
  __syncthreads();
  }
  // Rusanov speed of each face from its two averages (riemannSolverNonlinear.cpph:51-72).
  // Even NS: lane s < 4D takes side s & 1 of face s >> 1 and the pair's maximum
  // comes from a lane swap (tid and tid ^ 1 are the cell's lanes s and s ^ 1)
  if constexpr (NS % 2 == 0) {
    if (valid && s < 4 * D) {
      const double e = pde_max_abs_eig<D, V>(P, avg + s * ND, s >> 2);
      const double m = fmax(e, __shfl_xor(e, 1, 64));
      if (!(s & 1)) sm[s >> 1] = m;
    }
  } else if (valid && s < 2 * D) {
    sm[s] = fmax(pde_max_abs_eig<D, V>(P, avg + (s * 2 + 0) * ND, s >> 1),
                 pde_max_abs_eig<D, V>(P, avg + (s * 2 + 1) * ND, s >> 1));
  }
  __syncthreads();
  // phase 3: fluxes into red[task][V] (red's partial products were consumed
  // in phase 2; phase 3 reads only sm and the face data)
  if constexpr (ONE) {
    if (valid && s < NTASK) {
      const int task = s;
      const int f = task / NF;
      // 0.5 (FR + FL) - 0.5 smax (QR - QL) with (L, R) = (other, own) for o = 0 and
      // (own, other) for o = 1: FR + FL is the sum of both sides either way and
      // QR - QL = +-(Qother - Qown) exactly, so the sign moves to the factor and
      // both cells of a face form the same bits
      const double hs = (f & 1) ? 0.5 * sm[f] : -(0.5 * sm[f]);
#pragma unroll
      for (int k = 0; k < V; k++) {
        const double fw = rFo[k], fx = rFn[k], qw = rQo[k], qx = rQn[k];
        red[task * V + k] = 0.5 * (fw + fx) - hs * (qx - qw);
      }
    }
  } else if (valid) {
    for (int task = s; task < NTASK; task += NS) {
      const int f = task / NF, fn = task - f * NF, d = f >> 1, o = f & 1;
      const long g = ghost_face<D>(cxyz, n, f, bcmask);
      const long nb = g >= 0 ? 0 : periodic_neighbour<D>(cxyz, n, d, o ? 1 : -1);
      const TF* qo = lQhbnd + (cell * 2 * D + f) * NF * ND;
      const TF* qn = g >= 0 ? gq + g * NF * ND : lQhbnd + (nb * 2 * D + (f ^ 1)) * NF * ND;
      const TF* fo = lFhbnd + (cell * 2 * D + f) * NF * V;
      const TF* fnb = g >= 0 ? gf + g * NF * V : lFhbnd + (nb * 2 * D + (f ^ 1)) * NF * V;
      const TF* QL = o ? qo : qn;
      const TF* QR = o ? qn : qo;
      const TF* FL = o ? fo : fnb;
      const TF* FR = o ? fnb : fo;
      const double smax = sm[f];
#pragma unroll
      for (int k = 0; k < V; k++) {
        const double fr = face_val(FR, fn, k, NF, V), fl = face_val(FL, fn, k, NF, V);
        const double qr = face_val(QR, fn, k, NF, ND), ql = face_val(QL, fn, k, NF, ND);
        red[task * V + k] = 0.5 * (fr + fl) - 0.5 * smax * (qr - ql);
      }
    }
  }
  __syncthreads();
  // phase 4
  double myden = 0.0;  // lanes past the cell: no constraint
  if (valid) {
    int z, y, x;
    node_zyx<D, B>(s, z, y, x);
    const double h = hcell[cell];
    const double dt = *dtp;
    double du[V];
#pragma unroll
    for (int v = 0; v < V; v++) du[v] = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int c = coord<D, B>(d, z, y, x);
      const int fn = face_node<D, B>(d, z, y, x);
      const double tw = transverse_weight<D>(w, d, z, y, x);
#pragma unroll
      for (int o = 0; o < 2; o++) {
        const double scaling = 1.0 / h * (2.0 * o - 1.0);
        const double sw = tw * scaling;
        const double fcoef = o ? T->F1[c] : T->F0[c];
        const double* F = red + ((2 * d + o) * NF + fn) * V;
#pragma unroll
        for (int v = 0; v < V; v++) du[v] -= sw * fcoef * F[v];
      }
    }
    const double Wn = (D == 3) ? w[z] * w[y] * w[x] : w[y] * w[x];
    const double upd = dt / Wn;
#pragma unroll
    for (int v = 0; v < V; v++) {
      Q[v] = Q[v] + du[v] * upd;
      luh[(cell * NS + s) * ND + v] = Q[v];
    }
    double invDx[D];
#pragma unroll
    for (int d = 0; d < D; d++) invDx[d] = 1.0 / h;
    myden = node_cfl_den<D, B, V, NP>(Q, P, invDx);
  }
  // the cell's stable dt: min over nodes of cflpnpm / den = cflpnpm / max den
  // (correctly rounded division is monotone in den > 0: same bits, one division)
  const double bmax = block_max<THREADS>(myden, sred);
  const long lin = long(blockIdx.x) + long(gridDim.x) * (long(blockIdx.y) + long(gridDim.y) * blockIdx.z);
  double* dst = slots ? slots + (lin % kDtSlots) * kDtSlotStride : dt_next;
  if (tid == 0 && bmax > 0.0) {
    const double bmin = cflpnpm / bmax;
    if (bmin < DBL_MAX) atomic_min_pos_double(dst, bmin);
  }
}

// Single-cell helpers used by the per-cell compatibility entry points.
template <int D, int B, int V, int NP>
__global__ void face_integral_kernel(double* __restrict__ lduh, const double* __restrict__ F, int d, int o,
                                     double invdx, const DevTables* __restrict__ T) {
  using S = Shape<D, B, V, NP>;
  constexpr int NS = S::NS;
  const int s = threadIdx.x;
  if (s >= NS) return;
  int z, y, x;
  node_zyx<D, B>(s, z, y, x);
  const int c = coord<D, B>(d, z, y, x);
  const int fn = face_node<D, B>(d, z, y, x);
  const double scaling = invdx * (2.0 * o - 1.0);
  const double sw = transverse_weight<D>(T->w, d, z, y, x) * scaling;
  const double fcoef = o ? T->F1[c] : T->F0[c];
#pragma unroll
  for (int v = 0; v < V; v++) lduh[s * V + v] -= sw * fcoef * F[fn * V + v];
}

template <int D, int B, int V, int NP>
__global__ void solution_update_kernel(double* __restrict__ luh, const double* __restrict__ luhOld,
                                       const double* __restrict__ lduh, const double* __restrict__ dtp,
                                       const DevTables* __restrict__ T) {
  using S = Shape<D, B, V, NP>;
  constexpr int NS = S::NS, ND = S::ND;
  const int s = threadIdx.x;
  if (s >= NS) return;
  int z, y, x;
  node_zyx<D, B>(s, z, y, x);
  const double* w = T->w;
  const double Wn = (D == 3) ? w[z] * w[y] * w[x] : w[y] * w[x];
  const double upd = *dtp / Wn;
#pragma unroll
  for (int v = 0; v < V; v++) luh[s * ND + v] = luhOld[s * ND + v] + lduh[s * V + v] * upd;
}

// state[i] = {t, dt, dt_next, -} per instance; strided layout: t[ninst], dt[ninst], dt_next[ninst]
__global__ void advance_time_kernel(double* state, int ninst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ninst) {
    state[i] += state[ninst + i];
    state[ninst + i] = state[2 * ninst + i];
    state[2 * ninst + i] = DBL_MAX;
  }
}

__global__ void fill_kernel(double* p, long n, double value) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// ---------------------------------------------------------------------------
// Dispatch over the compiled (D, B, V, NP) shapes
// ---------------------------------------------------------------------------
template <int D, int B, int V, int NP>
struct Launch {
  using S = Shape<D, B, V, NP>;
  static long grid_of(long tiles, const LaunchOpts& o) { return (o.grid > 0 && o.grid < tiles) ? o.grid : tiles; }
  static int predictor(const KernelCfg& k, hipStream_t st, long n, const double* luh, const double* h,
                       const double* dt, double* lduh, double* qb, double* fb, int* pic, const LaunchOpts& o) {
    const long tiles = (n + S::CPB - 1) / S::CPB;
    const long grid = grid_of(tiles, o);
    // Picard rule: generic (adaptive or fixed count) or the optimised kernels'
    // (fusedSPTVI_nonlinear_cpp.template:177, tolerance exit always on)
    const int adaptive = k.dialect ? 1 : k.picard_max <= 0;
    const int maxIter = k.dialect ? (k.guess ? (B - 3 <= 1 ? 1 : B - 3) : 2 * B + 1)
                                  : (adaptive ? 2 * (k.order + 1) : k.picard_max);
    const int dl = k.dialect;
    const bool cerk = k.guess != 0;
    // the CERK form is compiled without the persistent variant (one workgroup per tile)
    const bool persistent = grid < tiles && !cerk;
    if (pde_has_ncs(k.pde)) {
      // source / non-conservative terms: FP64, trivial guess, one workgroup per tile, 5 variables
      if constexpr (V == kUserNVar && NP == 0) {
        if (k.precision != 0 || cerk) return -1;
        hipLaunchKernelGGL((predictor_kernel<D, B, V, NP, double, false, true>), dim3(tiles), dim3(S::THREADS), 0,
                           st, n, tiles, luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive,
                           o.luh_update, dl, o.hdim);
        return static_cast<int>(hipGetLastError());
      } else {
        return -1;
      }
    }
    if (k.precision == 0)
      launch_predictor_form<double>(st, cerk, persistent, grid, tiles, n, luh, h, dt, lduh, qb, fb, pic, k,
                                    maxIter, adaptive, o, dl);
    else
      launch_predictor_form<float>(st, cerk, persistent, grid, tiles, n, luh, h, dt, lduh, qb, fb, pic, k,
                                   maxIter, adaptive, o, dl);
    return static_cast<int>(hipGetLastError());
  }
  template <typename R>
  static void launch_predictor_form(hipStream_t st, bool cerk, bool persistent, long grid, long tiles, long n,
                                    const double* luh, const double* h, const double* dt, double* lduh, double* qb,
                                    double* fb, int* pic, const KernelCfg& k, int maxIter, int adaptive,
                                    const LaunchOpts& o, int dl) {
    // 3D Euler of orders 4 and 5 in FP64: the slab of the 12 distinct flux components
    if constexpr (D == 3 && (B == 5 || B == 6) && V == 5 && NP == 0 && sizeof(R) == 8) {
      if (k.pde.id == kPdeEuler) {
        if (persistent)
          hipLaunchKernelGGL((predictor_persistent_kernel<D, B, V, NP, R, false, true>), dim3(grid),
                             dim3(S::THREADS), 0, st, n, tiles, luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde,
                             maxIter, adaptive, o.luh_update, dl, o.hdim);
        else if (cerk)
          hipLaunchKernelGGL((predictor_kernel<D, B, V, NP, R, true, false, true>), dim3(tiles), dim3(S::THREADS), 0,
                             st, n, tiles, luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive,
                             o.luh_update, dl, o.hdim);
        else
          hipLaunchKernelGGL((predictor_kernel<D, B, V, NP, R, false, false, true>), dim3(tiles), dim3(S::THREADS), 0,
                             st, n, tiles, luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive,
                             o.luh_update, dl, o.hdim);
        return;
      }
    }
    if (cerk)
      hipLaunchKernelGGL((predictor_kernel<D, B, V, NP, R, true>), dim3(tiles), dim3(S::THREADS), 0, st, n, tiles,
                         luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, o.luh_update, dl, o.hdim);
    else if (persistent)
      hipLaunchKernelGGL((predictor_persistent_kernel<D, B, V, NP, R, false>), dim3(grid), dim3(S::THREADS), 0, st,
                         n, tiles, luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive,
                         o.luh_update, dl, o.hdim);
    else
      hipLaunchKernelGGL((predictor_kernel<D, B, V, NP, R, false>), dim3(tiles), dim3(S::THREADS), 0, st, n, tiles,
                         luh, h, dt, o.cinst, lduh, qb, fb, pic, k.tab, k.pde, maxIter, adaptive, o.luh_update, dl, o.hdim);
  }
  static int riemann(const KernelCfg& k, hipStream_t st, long nf, const int* faces, const double* qb,
                     double* fb, const LaunchOpts& o) {
    // riemann_kernel is compiled for at most 256 threads (4 waves) per workgroup
    const int waves = (o.riemann_waves >= 1 && o.riemann_waves <= 4) ? o.riemann_waves : 4;
    const long blocks = (nf + waves - 1) / waves;
    hipLaunchKernelGGL((riemann_kernel<D, B, V, NP>), dim3(blocks), dim3(64 * waves), 0, st, nf,
                       reinterpret_cast<const int4*>(faces), qb, fb, o.ghost_q, o.ghost_f, k.tab, k.pde);
    return static_cast<int>(hipGetLastError());
  }
  static int corrector(const KernelCfg& k, hipStream_t st, long n, double* luh, double* lduh, const double* fb,
                       const double* h, const double* dt, double* dt_next, int do_update, const LaunchOpts& o) {
    const long tiles = (n + S::CPB - 1) / S::CPB;
    const long grid = grid_of(tiles, o);
    if (grid < tiles)
      hipLaunchKernelGGL((corrector_persistent_kernel<D, B, V, NP>), dim3(grid), dim3(S::THREADS), 0, st, n, tiles,
                         luh, lduh, fb, h, dt, o.cinst, dt_next, k.tab, k.pde, k.cfl * k.pnpm, do_update, o.dt_slots, o.hdim);
    else
      hipLaunchKernelGGL((corrector_kernel<D, B, V, NP>), dim3(grid), dim3(S::THREADS), 0, st, n, tiles, luh, lduh,
                         fb, h, dt, o.cinst, dt_next, k.tab, k.pde, k.cfl * k.pnpm, do_update, o.dt_slots, o.hdim);
    return static_cast<int>(hipGetLastError());
  }
  static int corrector_riemann(const KernelCfg& k, hipStream_t st, long n, int nside, double* luh,
                               const double* qb, const double* fb, const double* h, const double* dt,
                               double* dt_next, const LaunchOpts& o) {
    if constexpr (FusedShape<D, B, V, NP>::ok) {
      if (pde_has_ncp(k.pde)) return -1;  // the fused corrector solves conservative Rusanov problems only
      if (n <= 0) return 0;
      // cells [c0, c1) of the n cells; one cell per workgroup in 3D: whole z planes,
      // grid (nside, nside, planes)
      const long c0 = o.cell0, c1 = o.cell_end < 0 ? n : o.cell_end;
      if (c0 < 0 || c0 >= c1 || c1 > n) return -1;
      constexpr bool GRID3 = D == 3 && S::CPB == 1;
      const long plane = long(nside) * nside;
      if (GRID3 && (c0 % plane || c1 % plane)) return -1;
      const dim3 grid = GRID3 ? dim3(nside, nside, unsigned((c1 - c0) / plane))
                              : dim3(unsigned((c1 - c0 + S::CPB - 1) / S::CPB));
      const long first = GRID3 ? c0 / plane : c0;
      if (o.faces_f32) {
        if constexpr (D == 3 && B == 6 && V == 5 && NP == 0) {  // the FP32 line predictor's shape
          hipLaunchKernelGGL((corrector_riemann_kernel<D, B, V, NP, float>), grid, dim3(S::THREADS), 0, st, c1, nside,
                             luh, reinterpret_cast<const float*>(qb), reinterpret_cast<const float*>(fb), h, dt,
                             dt_next, k.tab, k.pde, k.cfl * k.pnpm, reinterpret_cast<const float*>(o.ghost_q),
                             reinterpret_cast<const float*>(o.ghost_f), o.bcmask, o.dt_slots, first);
          return static_cast<int>(hipGetLastError());
        } else {
          return -1;
        }
      }
      hipLaunchKernelGGL((corrector_riemann_kernel<D, B, V, NP, double>), grid, dim3(S::THREADS), 0, st, c1, nside,
                         luh, qb, fb, h, dt, dt_next, k.tab, k.pde, k.cfl * k.pnpm, o.ghost_q, o.ghost_f, o.bcmask,
                         o.dt_slots, first);
      return static_cast<int>(hipGetLastError());
    } else {
      return -1;
    }
  }
  static int face_integral(const KernelCfg& k, hipStream_t st, double* lduh, const double* F, int d, int o,
                           double invdx) {
    hipLaunchKernelGGL((face_integral_kernel<D, B, V, NP>), dim3(1), dim3(ceil64(S::NS)), 0, st, lduh, F, d, o,
                       invdx, k.tab);
    return static_cast<int>(hipGetLastError());
  }
  static int solution_update(const KernelCfg& k, hipStream_t st, double* luh, const double* luhOld,
                             const double* lduh, const double* dt) {
    hipLaunchKernelGGL((solution_update_kernel<D, B, V, NP>), dim3(1), dim3(ceil64(S::NS)), 0, st, luh, luhOld,
                       lduh, dt, k.tab);
    return static_cast<int>(hipGetLastError());
  }
};

struct Entry {
  int D, B, V, NP;
  int (*predictor)(const KernelCfg&, hipStream_t, long, const double*, const double*, const double*, double*,
                   double*, double*, int*, const LaunchOpts&);
  int (*riemann)(const KernelCfg&, hipStream_t, long, const int*, const double*, double*, const LaunchOpts&);
  int (*corrector)(const KernelCfg&, hipStream_t, long, double*, double*, const double*, const double*,
                   const double*, double*, int, const LaunchOpts&);
  int (*corrector_riemann)(const KernelCfg&, hipStream_t, long, int, double*, const double*, const double*,
                           const double*, const double*, double*, const LaunchOpts&);
  int (*face_integral)(const KernelCfg&, hipStream_t, double*, const double*, int, int, double);
  int (*solution_update)(const KernelCfg&, hipStream_t, double*, const double*, const double*, const double*);
};

#define EXA_ENTRY(D, B, V, NP)                                                                      \
  Entry {                                                                                           \
    D, B, V, NP, &Launch<D, B, V, NP>::predictor, &Launch<D, B, V, NP>::riemann,                    \
        &Launch<D, B, V, NP>::corrector, &Launch<D, B, V, NP>::corrector_riemann,                   \
        &Launch<D, B, V, NP>::face_integral,                                                        \
        &Launch<D, B, V, NP>::solution_update                                                       \
  }

// Compiled shapes: 5 variables (Euler, advection, the default user PDE) for
// orders 1..6 in 2D and 1..5 in 3D, plus the reference unit-test shape (Order
// 3, one parameter); other variable counts for linear advection (1 and 3
// variables) and for a user PDE that declares kNVar (orders 3 and 4; shapes
// that coincide with the list above are found there first).
const Entry kEntries[] = {
    EXA_ENTRY(2, 2, 5, 0), EXA_ENTRY(2, 3, 5, 0), EXA_ENTRY(2, 4, 5, 0), EXA_ENTRY(2, 5, 5, 0),
    EXA_ENTRY(2, 6, 5, 0), EXA_ENTRY(2, 7, 5, 0), EXA_ENTRY(3, 2, 5, 0), EXA_ENTRY(3, 3, 5, 0),
    EXA_ENTRY(3, 4, 5, 0), EXA_ENTRY(3, 5, 5, 0), EXA_ENTRY(3, 6, 5, 0), EXA_ENTRY(2, 4, 5, 1),
    EXA_ENTRY(3, 4, 5, 1),
    EXA_ENTRY(2, 4, 1, 0), EXA_ENTRY(3, 4, 1, 0), EXA_ENTRY(3, 3, 3, 0),
    EXA_ENTRY(2, 4, kUserNVar, 0), EXA_ENTRY(2, 5, kUserNVar, 0), EXA_ENTRY(3, 4, kUserNVar, 0),
    EXA_ENTRY(3, 5, kUserNVar, 0),
};

const Entry* find_entry(const KernelCfg& k) {
  for (const Entry& e : kEntries)
    if (e.D == k.dim && e.B == k.order + 1 && e.V == k.nvar && e.NP == k.npar) return &e;
  return nullptr;
}

}  // namespace

bool shape_supported(int dim, int order, int nvar, int npar) {
  for (const Entry& e : kEntries)
    if (e.D == dim && e.B == order + 1 && e.V == nvar && e.NP == npar) return true;
  return false;
}

int launch_predictor(const KernelCfg& k, hipStream_t s, long n, const double* luh, const double* h,
                     const double* dt, double* lduh, double* qb, double* fb, int* pic, const LaunchOpts& o) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  if (n <= 0) return 0;
  // a part of the cells: the one-iteration FP32 CERK line kernel only
  if ((o.cell0 != 0 || o.cell_end >= 0) && (!predictor_supports_range(k) || o.hdim)) return -1;
  // BF16 operands: the tile kernel (3D Euler order 5, trivial guess) only
  if (k.precision == 2) {
    if (!line_kernels_serve(k) || k.guess || o.hdim) return -1;
    return launch_predictor3d_tile(k, s, n, luh, h, dt, lduh, qb, fb, pic, o.cinst, o.luh_update, o.faces_f32);
  }
  // auto: FP32 -> line kernel on the matrix cores; FP64 -> node-column kernel
  // (non-cubic cells: node-column kernel)
  if (line_kernels_serve(k) && !o.hdim) {
    const int v = k.pred_variant;
    const bool fp32 = k.precision != 0;
    if ((v == 0 && fp32) || v == 3)
      return launch_predictor3d_lines(k, s, n, luh, h, dt, lduh, qb, fb, pic, fp32, o.cinst, o.luh_update,
                                      o.faces_f32, o.guard, o.cell0, o.cell_end);
    if (v == 2)
      return launch_predictor3d_lines(k, s, n, luh, h, dt, lduh, qb, fb, pic, false, o.cinst, o.luh_update,
                                      o.faces_f32);
    if (v == 4 && fp32)
      return launch_predictor3d_tile(k, s, n, luh, h, dt, lduh, qb, fb, pic, o.cinst, o.luh_update, o.faces_f32);
  }
  if (o.faces_f32) return -1;  // the generic kernels write FP64 faces
  return e->predictor(k, s, n, luh, h, dt, lduh, qb, fb, pic, o);
}

int launch_riemann(const KernelCfg& k, hipStream_t s, long nf, const int* faces, const double* qb, double* fb,
                   const LaunchOpts& o) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  if (nf <= 0) return 0;
  return e->riemann(k, s, nf, faces, qb, fb, o);
}

int launch_corrector(const KernelCfg& k, hipStream_t s, long n, double* luh, double* lduh, const double* fb,
                     const double* h, const double* dt, double* dt_next, const LaunchOpts& o) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  if (n <= 0) return 0;
  return e->corrector(k, s, n, luh, lduh, fb, h, dt, dt_next, 1, o);
}

int launch_corrector_riemann(const KernelCfg& k, hipStream_t s, long n, int nside, double* luh,
                             const double* qb, const double* fb, const double* h, const double* dt,
                             double* dt_next, const LaunchOpts& o) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  if (n <= 0) return 0;
  return e->corrector_riemann(k, s, n, nside, luh, qb, fb, h, dt, dt_next, o);
}

bool corrector_riemann_supported(const KernelCfg& k) {
  const Entry* e = find_entry(k);
  // n = 0 launches nothing; -1 marks a shape whose fused kernel is not compiled
  return e && e->corrector_riemann(k, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                   LaunchOpts()) != -1;
}

int launch_stable_dt(const KernelCfg& k, hipStream_t s, long n, const double* luh, const double* h,
                     double* dt_out, const LaunchOpts& o) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  if (n <= 0) return 0;
  return e->corrector(k, s, n, const_cast<double*>(luh), nullptr, nullptr, h, nullptr, dt_out, 0, o);
}

int launch_face_integral(const KernelCfg& k, hipStream_t s, double* lduh, const double* F, int d, int o,
                         double invdx) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  return e->face_integral(k, s, lduh, F, d, o, invdx);
}

int launch_solution_update(const KernelCfg& k, hipStream_t s, double* luh, const double* luhOld,
                           const double* lduh, const double* dt) {
  const Entry* e = find_entry(k);
  if (!e) return -1;
  return e->solution_update(k, s, luh, luhOld, lduh, dt);
}

namespace {
__global__ __launch_bounds__(256) void dt_reduce_kernel(double* slots, double* dt_next) {
  __shared__ double sred[4];
  double m = DBL_MAX;
  for (int i = threadIdx.x; i < kDtSlots; i += 256) {
    m = fmin(m, slots[i * kDtSlotStride]);
    slots[i * kDtSlotStride] = DBL_MAX;  // ready for the next step
  }
  const double bmin = block_min<256>(m, sred);
  if (threadIdx.x == 0 && bmin < DBL_MAX) atomic_min_pos_double(dt_next, bmin);
}

__global__ __launch_bounds__(256) void dt_slots_init_kernel(double* slots) {
  for (int i = threadIdx.x; i < kDtSlots; i += 256) slots[i * kDtSlotStride] = DBL_MAX;
}
}  // namespace

int launch_dt_reduce(hipStream_t s, double* slots, double* dt_next) {
  hipLaunchKernelGGL(dt_reduce_kernel, dim3(1), dim3(256), 0, s, slots, dt_next);
  return static_cast<int>(hipGetLastError());
}

int launch_dt_slots_init(hipStream_t s, double* slots) {
  hipLaunchKernelGGL(dt_slots_init_kernel, dim3(1), dim3(256), 0, s, slots);
  return static_cast<int>(hipGetLastError());
}

int launch_advance_time(hipStream_t s, double* state, int ninst) {
  hipLaunchKernelGGL(advance_time_kernel, dim3((ninst + 255) / 256), dim3(256), 0, s, state, ninst);
  return static_cast<int>(hipGetLastError());
}

int launch_fill(hipStream_t s, double* p, long n, double value) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, value);
  return static_cast<int>(hipGetLastError());
}

}  // namespace exa

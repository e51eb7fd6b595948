// AMR projections between a cell of level l and its 3^D children
// (reference: ExaHyPE/kernels/aderdg/generic/c/{2d,3d}/amrRoutines.cpph):
//   prolongation  fine[m] = sum_n coarse[n] prod_k fgp[s_k][n_k][m_k]               (:23-41)
//   restriction   coarse[m] += sum_n prod_k (w[n_k] fgp[s_k][m_k][n_k] / w[m_k]) fine[n] / 3^K
//                                                                                    (:145-170)
// with K = D-1 axes for face data and K = D for volume data, node index
// first axis fastest, fgp[s][n][m] = phi_n((s + x_m) / 3).
//
// On the static two-level mesh (amr_host.cpp) a coarse leaf next to a
// refined cell owns 3^(D-1) virtual children per such face: their face data
// are the coarse face prolongated (face_prolong_kernel, the reference's
// prolongateFaceDataToVirtualCell), the Riemann solver runs on the fine
// subfaces, and face_restrict_kernel sums the children's restricted
// fluctuations into the coarse face (restrictToTopMostParent + the
// faceIntegral glue with levelDelta = 1), which the corrector then
// integrates with the coarse cell size.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace exa {
namespace {

template <int K, int B>
__device__ __forceinline__ double prolong_weight(const double* __restrict__ fgp, int s0, int s1, int n, int m) {
  if (K == 1) return fgp[(s0 * B + n) * B + m];
  const int n1 = n % B, n2 = n / B, m1 = m % B, m2 = m / B;
  if (K == 2) return fgp[(s0 * B + n1) * B + m1] * fgp[(s1 * B + n2) * B + m2];
  return 0.0;
}

template <int K, int B>
__device__ __forceinline__ double restrict_weight(const double* __restrict__ fgp, const double* __restrict__ w, int s0,
                                                  int s1, int s2, int n, int m) {
  // prod_k w[n_k] fgp[s_k][m_k][n_k] / w[m_k]
  double r = 1.0;
  const int sk[3] = {s0, s1, s2};
  int nn = n, mm = m;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int a = nn % B, b = mm % B;
    r *= w[a] * fgp[(sk[k] * B + b) * B + a] / w[b];
    nn /= B;
    mm /= B;
  }
  return r;
}

template <int D, int B>
__global__ __launch_bounds__(256) void face_prolong_kernel(long n, const int4* __restrict__ ent, double* __restrict__ qb,
                                                           double* __restrict__ fb, const DevTables* __restrict__ T,
                                                           int V, int ND) {
  constexpr int K = D - 1;
  constexpr int NF = (D == 3) ? B * B : B;
  const long e = long(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= n) return;  // per wavefront, no barriers
  const int4 q = ent[e];
  const long C = q.x, Tc = q.w;
  const int f = q.y, s0 = q.z % 3, s1 = q.z / 3;
  const double* cq = qb + ((C * 2 * D + f) * NF) * ND;
  const double* cf = fb + ((C * 2 * D + f) * NF) * V;
  double* tq = qb + ((Tc * 2 * D + f) * NF) * ND;
  double* tf = fb + ((Tc * 2 * D + f) * NF) * V;
  for (int o = lane; o < NF * ND; o += 64) {
    const int m = o / ND, v = o - m * ND;
    double acc = 0.0;
    for (int nn = 0; nn < NF; nn++) acc += cq[nn * ND + v] * prolong_weight<K, B>(T->fgp, s0, s1, nn, m);
    tq[o] = acc;
  }
  for (int o = lane; o < NF * V; o += 64) {
    const int m = o / V, v = o - m * V;
    double acc = 0.0;
    for (int nn = 0; nn < NF; nn++) acc += cf[nn * V + v] * prolong_weight<K, B>(T->fgp, s0, s1, nn, m);
    tf[o] = acc;
  }
}

template <int D, int B>
__global__ __launch_bounds__(256) void face_restrict_kernel(long n, const int4* __restrict__ ent, double* __restrict__ fb,
                                                            const DevTables* __restrict__ T, int V) {
  constexpr int K = D - 1;
  constexpr int NF = (D == 3) ? B * B : B;
  constexpr int NSUB = (D == 3) ? 9 : 3;
  const long e = long(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= n) return;
  const int4 q = ent[e];
  const long C = q.x, T0 = q.z;
  const int f = q.y;
  double* cf = fb + ((C * 2 * D + f) * NF) * V;
  for (int o = lane; o < NF * V; o += 64) {
    const int m = o / V, v = o - m * V;
    double acc = 0.0;
    for (int sub = 0; sub < NSUB; sub++) {
      const double* sf = fb + (((T0 + sub) * 2 * D + f) * NF) * V;
      double part = 0.0;
      for (int nn = 0; nn < NF; nn++)
        part += restrict_weight<K, B>(T->fgp, T->w, sub % 3, sub / 3, 0, nn, m) * sf[nn * V + v];
      acc += part / double(NSUB);
    }
    cf[o] = acc;
  }
}

template <int K, int B>
__global__ __launch_bounds__(256) void project_kernel(int nv, const double* __restrict__ in, double* __restrict__ out,
                                                      int s0, int s1, int s2, int mode,
                                                      const DevTables* __restrict__ T) {
  int NK = 1;
#pragma unroll
  for (int k = 0; k < K; k++) NK *= B;
  const double scale = (K == 1) ? 3.0 : (K == 2 ? 9.0 : 27.0);
  for (int o = threadIdx.x; o < NK * nv; o += blockDim.x) {
    const int m = o / nv, v = o - m * nv;
    double acc = 0.0;
    for (int nn = 0; nn < NK; nn++) {
      double wgt;
      if (mode == 0) {
        wgt = 1.0;
        int a = nn, b = m;
        const int sk[3] = {s0, s1, s2};
        for (int k = 0; k < K; k++) {
          wgt *= T->fgp[(sk[k] * B + a % B) * B + b % B];
          a /= B;
          b /= B;
        }
      } else {
        wgt = restrict_weight<K, B>(T->fgp, T->w, s0, s1, s2, nn, m);
      }
      acc += wgt * in[nn * nv + v];
    }
    if (mode == 0)
      out[o] = acc;
    else
      out[o] += acc / scale;
  }
}

template <int B>
int dispatch_faces(const KernelCfg& k, hipStream_t s, long n, const int* ent, double* qb, double* fb, bool prolong) {
  const unsigned blocks = static_cast<unsigned>((n + 3) / 4);
  const int ND = k.nvar + k.npar;
  if (k.dim == 3) {
    if (prolong)
      face_prolong_kernel<3, B><<<blocks, 256, 0, s>>>(n, reinterpret_cast<const int4*>(ent), qb, fb, k.tab, k.nvar, ND);
    else
      face_restrict_kernel<3, B><<<blocks, 256, 0, s>>>(n, reinterpret_cast<const int4*>(ent), fb, k.tab, k.nvar);
  } else {
    if (prolong)
      face_prolong_kernel<2, B><<<blocks, 256, 0, s>>>(n, reinterpret_cast<const int4*>(ent), qb, fb, k.tab, k.nvar, ND);
    else
      face_restrict_kernel<2, B><<<blocks, 256, 0, s>>>(n, reinterpret_cast<const int4*>(ent), fb, k.tab, k.nvar);
  }
  return static_cast<int>(hipGetLastError());
}

int faces(const KernelCfg& k, hipStream_t s, long n, const int* ent, double* qb, double* fb, bool prolong) {
  if (n <= 0) return 0;
  switch (k.order + 1) {
    case 2: return dispatch_faces<2>(k, s, n, ent, qb, fb, prolong);
    case 3: return dispatch_faces<3>(k, s, n, ent, qb, fb, prolong);
    case 4: return dispatch_faces<4>(k, s, n, ent, qb, fb, prolong);
    case 5: return dispatch_faces<5>(k, s, n, ent, qb, fb, prolong);
    case 6: return dispatch_faces<6>(k, s, n, ent, qb, fb, prolong);
    case 7: return dispatch_faces<7>(k, s, n, ent, qb, fb, prolong);
    default: return -1;
  }
}

template <int K>
int project_k(const KernelCfg& k, hipStream_t s, int nv, const double* in, double* out, int s0, int s1, int s2,
              int mode) {
  switch (k.order + 1) {
    case 2: project_kernel<K, 2><<<1, 256, 0, s>>>(nv, in, out, s0, s1, s2, mode, k.tab); break;
    case 3: project_kernel<K, 3><<<1, 256, 0, s>>>(nv, in, out, s0, s1, s2, mode, k.tab); break;
    case 4: project_kernel<K, 4><<<1, 256, 0, s>>>(nv, in, out, s0, s1, s2, mode, k.tab); break;
    case 5: project_kernel<K, 5><<<1, 256, 0, s>>>(nv, in, out, s0, s1, s2, mode, k.tab); break;
    case 6: project_kernel<K, 6><<<1, 256, 0, s>>>(nv, in, out, s0, s1, s2, mode, k.tab); break;
    case 7: project_kernel<K, 7><<<1, 256, 0, s>>>(nv, in, out, s0, s1, s2, mode, k.tab); break;
    default: return -1;
  }
  return static_cast<int>(hipGetLastError());
}

}  // namespace

int launch_face_prolong(const KernelCfg& k, hipStream_t s, long n, const int* entries, double* qb, double* fb) {
  return faces(k, s, n, entries, qb, fb, true);
}

int launch_face_restrict(const KernelCfg& k, hipStream_t s, long n, const int* entries, double* fb) {
  return faces(k, s, n, entries, nullptr, fb, false);
}

int launch_face_project(const KernelCfg& k, hipStream_t s, int nv, const double* in, double* out, int s0, int s1,
                        int mode) {
  return k.dim == 3 ? project_k<2>(k, s, nv, in, out, s0, s1, 0, mode) : project_k<1>(k, s, nv, in, out, s0, 0, 0, mode);
}

int launch_volume_project(const KernelCfg& k, hipStream_t s, int nv, const double* in, double* out, int s0, int s1,
                          int s2, int mode) {
  return k.dim == 3 ? project_k<3>(k, s, nv, in, out, s0, s1, s2, mode) : project_k<2>(k, s, nv, in, out, s0, s1, 0, mode);
}

}  // namespace exa

// Internal declarations shared by the C-ABI layer, the mesh driver and the
// HIP kernels.  Not part of the public interface (include/exahype_amd.h).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "pde.h"
#include "tables.h"

namespace exa {

// Operator tables as stored on the device (one block per context).
struct DevTables {
  double w[kMaxBasis];
  double x[kMaxBasis];
  double F0[kMaxBasis];
  double F1[kMaxBasis];
  double Kxi[kMaxBasis * kMaxBasis];
  double iK1[kMaxBasis * kMaxBasis];
  double dudx[kMaxBasis * kMaxBasis];
  // FP32 copies for the reduced-precision predictor
  float wf[kMaxBasis];
  float F0f[kMaxBasis];
  float F1f[kMaxBasis];
  float Kxif[kMaxBasis * kMaxBasis];
  float iK1f[kMaxBasis * kMaxBasis];
  // derived operators of the line-decomposed predictor
  float dudxf[kMaxBasis * kMaxBasis];
  double Pt[kMaxBasis * kMaxBasis];  // Pt[i][j] = iK1[i][j] * w[j]
  float Ptf[kMaxBasis * kMaxBasis];
  float cwf[kMaxBasis];  // c[j] = sum_i w[i] Pt[i][j]: time average of one Picard update
  float cmom[3];         // sum_j c[j] x[j]^k, k = 0, 1, 2 (c-weighted time moments)
  double fgp[3 * kMaxBasis * kMaxBasis];  // fineGridProjector[sub][n][m] = phi_n((sub + x_m) / 3), [(sub*B+n)*B+m]
  double iw[kMaxBasis];                   // 1 / w[i]
};

struct KernelCfg {
  int dim, order, nvar, npar;
  int precision;   // exa_precision
  int picard_max;  // <= 0 adaptive
  double cfl, pnpm;
  PdeParams pde;
  const DevTables* tab;  // device pointer
  int pred_variant;      // 0 auto, 1 node-column, 2 line (VALU), 3 line (FP32 matrix cores),
                         // 4 tile (FP32, full-rate matrix cores)
  int dialect;           // 0 generic Picard rule; 1 the optimised kernels' rule (MaxIterations
                         // 2B+1, or max(B-3, 1) with the CERK guess; tolerance exit always on;
                         // count min(iter+1, Max)); runs on the node-column kernel
  int guess;             // initial guess: 0 luh at every time node; 1 CERK (node-column kernel,
                         // and the FP32 matrix-core line kernel with a fixed Picard count)
  // Picard guard of the one-iteration FP32 CERK form: a cell whose relative
  // proxy residual |lQhi(1) - lQhi(0)| / |lQhi(1)| reaches guard_tol (<= 0:
  // off) is recomputed with guard_recount fixed iterations (0: max(B - 3, 1),
  // the optimised kernels' CERK cap)
  double guard_tol;
  int guard_recount;
  // linear path: 1 = the optimised kernels' split Cauchy-Kowalevski predictor
  // (one derivative level kept, B levels in the time average, material matrix
  // on the flux part only); the other linear kernels stay generic
  int split_ck;
};

// Launch options of the batched kernels.
//   cinst          per-cell instance index (nullable): dt[cinst[c]], dt_next[cinst[c]];
//                  cells of one instance must be contiguous
//   grid           > 0: persistent launch with this many workgroups looping over the
//                  cell tiles (predictor, corrector); 0: one workgroup per tile
//   riemann_waves  faces (wavefronts) per Riemann workgroup, 1..4 (the kernel's
//                  256-thread launch bound; other values fall back to 4)
//   luh_update     predictor only: add the volume part of the update to this luh
//                  (the predictor's own input) instead of writing lduh
//   ghost_q/_f     Riemann kernels: outside face data of boundary faces
//                  ([g][NF][ND], [g][NF][V]); a face-list entry with w & 2
//                  (w & 1) takes its left (right) side from ghost face x (y)
//   bcmask         fused corrector: bit f set = domain face f is not periodic;
//                  its cells read ghost face f * n^(D-1) + transverse index
struct LaunchOpts {
  const int* cinst = nullptr;
  long grid = 0;
  int riemann_waves = 4;
  double* luh_update = nullptr;
  const double* ghost_q = nullptr;
  double* ghost_f = nullptr;
  unsigned bcmask = 0;
  // FP32 predictors and the fused corrector: extrapolated faces stored as FP32
  // (the FP32 forms compute them in FP32, so the storage is exact)
  bool faces_f32 = false;
  // correctors without cinst: workgroup minima of the stable dt go to one of
  // kDtSlots slots (separate cache lines) instead of all to dt_next; a
  // launch_dt_reduce then merges the slots into dt_next and resets them
  double* dt_slots = nullptr;
  // predictor, one-iteration FP32 CERK form: Picard guard buffer of 2 + ncells
  // ints {listed this launch (0 on entry), listed in total, list}; nullptr: no guard
  int* guard = nullptr;
  // node-column predictor and corrector: 0 = one edge length per cell (h[cell]),
  // D = one per direction (h[cell][d]); the other kernels take cubic cells only
  int hdim = 0;
  // the one-iteration FP32 CERK line predictor and the fused corrector: the
  // launch covers cells [cell0, cell_end) of the n cells (cell_end < 0: n);
  // the other kernels cover all cells and return -1 for a part
  long cell0 = 0;
  long cell_end = -1;
};

// Slots of the spread dt reduction: one device-scope atomic per workgroup on a
// single address serialises at that address (~10 ns each, 0.5 M workgroups at
// 81^3); spread over 1024 cache lines the atomics proceed in parallel.
constexpr int kDtSlots = 1024;
constexpr int kDtSlotStride = 16;  // doubles: one 128-byte line per slot
// slots [kDtSlots * kDtSlotStride] -> dt_next = min(dt_next, min slots); slots reset to DBL_MAX
int launch_dt_reduce(hipStream_t s, double* slots, double* dt_next);
int launch_dt_slots_init(hipStream_t s, double* slots);

// Boundary kinds per domain face f = 2d + o (exa_boundary_kind).
enum BcKind { kBcPeriodic = 0, kBcWall = 1, kBcDirichletEntropyWave = 2, kBcExchange = 3 };
struct BcKinds {
  int k[6];
};

// Launchers (return hipError_t as int; 0 on success, -1 for an unsupported
// (dim, order, nvar, npar, precision) combination).
int launch_predictor(const KernelCfg& k, hipStream_t s, long ncells, const double* luh, const double* h,
                     const double* dt, double* lduh, double* lQhbnd, double* lFhbnd, int* picard,
                     const LaunchOpts& o = LaunchOpts());
int launch_riemann(const KernelCfg& k, hipStream_t s, long nfaces, const int* faces, const double* lQhbnd,
                   double* lFhbnd, const LaunchOpts& o = LaunchOpts());
int launch_corrector(const KernelCfg& k, hipStream_t s, long ncells, double* luh, double* lduh,
                     const double* lFhbnd, const double* h, const double* dt, double* dt_next,
                     const LaunchOpts& o = LaunchOpts());
int launch_stable_dt(const KernelCfg& k, hipStream_t s, long ncells, const double* luh, const double* h,
                     double* dt_out, const LaunchOpts& o = LaunchOpts());
// Fused Riemann + face integrals + update + CFL on a uniform periodic mesh of
// nside cells per direction (cells x fastest); luh already holds the volume
// part of the update.  -1 when the shape's LDS need exceeds the kernel budget.
int launch_corrector_riemann(const KernelCfg& k, hipStream_t s, long ncells, int nside, double* luh,
                             const double* lQhbnd, const double* lFhbnd, const double* h, const double* dt,
                             double* dt_next, const LaunchOpts& o = LaunchOpts());
bool corrector_riemann_supported(const KernelCfg& k);
// Single-cell helper kernels for the compatibility entry points.
int launch_face_integral(const KernelCfg& k, hipStream_t s, double* lduh, const double* lFface, int direction,
                         int orientation, double invdx_scaled);
int launch_solution_update(const KernelCfg& k, hipStream_t s, double* luh, const double* luhOld,
                           const double* lduh, const double* dt);
// Line-decomposed 3D predictor (predictor3d_lines.hip) for the shapes it covers.
bool lines_supported(const KernelCfg& k);
// The 3D order-5 line/tile predictors serve k: generic Picard rule, and with the
// CERK guess only the FP32 matrix-core line kernel with a fixed Picard count.
bool line_kernels_serve(const KernelCfg& k);
int launch_predictor3d_lines(const KernelCfg& k, hipStream_t s, long ncells, const double* luh, const double* h,
                             const double* dt, double* lduh, double* lQhbnd, double* lFhbnd, int* picard,
                             bool mfma, const int* cinst = nullptr, double* luh_update = nullptr,
                             bool f32faces = false, int* guard = nullptr, long cell0 = 0, long cell_end = -1);
// FP32 tile form of the same predictor (3D Euler, order 5): space-time iterate
// in LDS, contractions on 16x16x4 matrix-core steps; -1 for FP64.
int launch_predictor3d_tile(const KernelCfg& k, hipStream_t s, long ncells, const double* luh, const double* h,
                            const double* dt, double* lduh, double* lQhbnd, double* lFhbnd, int* picard,
                            const int* cinst = nullptr, double* luh_update = nullptr, bool f32faces = false);
// The predictor selected by k writes FP32 faces on request (LaunchOpts::faces_f32).
bool predictor_f32_faces_supported(const KernelCfg& k);
// The predictor selected by k runs the Picard guard when given LaunchOpts::guard
// (the one-iteration FP32 CERK line kernel with guard_tol > 0).
bool predictor_uses_guard(const KernelCfg& k);
// The predictor selected by k covers a part of the cells (LaunchOpts::cell0,
// cell_end): the one-iteration FP32 CERK line kernel.
bool predictor_supports_range(const KernelCfg& k);
// Time bookkeeping of the device step per instance, state = t[ninst], dt[ninst],
// dt_next[ninst]: t += dt; dt = dt_next; dt_next = DBL_MAX.
int launch_advance_time(hipStream_t s, double* state, int ninst = 1);

// Boundary conditions (Euler): outside face data of every boundary face of a
// uniform n^D mesh (ghost layout of LaunchOpts::ghost_q) at the step's t, dt
// read from state[0], state[1]; or of one face of one cell.
// Point probes: values [npts][nData] of the solution of a uniform n^D mesh at
// points [npts][dim] inside the domain (tensor-product Lagrange interpolation).
int launch_probe(const KernelCfg& k, hipStream_t s, int n, double h, const double* luh, int npts, const double* pts,
                 double* out);
// Global observables of a uniform mesh: part holds observables_blocks(ncells)
// * 3 * nData doubles, out 2 + 3 nData ([dgcells, fvcells, integral, min, max]).
int observables_blocks(long ncells);
int launch_observables(const KernelCfg& k, hipStream_t s, long ncells, double h, const double* luh, double* part,
                       double* out);
bool boundary_supported(const KernelCfg& k);
int launch_boundary_mesh(const KernelCfg& k, hipStream_t s, int n, double h, const BcKinds& bk, const double* qb,
                         double* gq, double* gf, const double* state);
int launch_boundary_face(const KernelCfg& k, hipStream_t s, int kind, const double* centre, const double* dx,
                         double t, double dt, int d, int o, const double* qin, double* gq, double* gf);
// Halo message of domain face f of a uniform n^D mesh: the extrapolated
// predictor faces [n^(D-1)][NF][ND] and fluctuation faces [n^(D-1)][NF][V] of
// the cells on that face, in transverse-index order (the ghost-face order).
int launch_halo_pack(const KernelCfg& k, hipStream_t s, int n, int f, const double* qb, const double* fb,
                     double* msg_q, double* msg_f);
// Initial condition EXA_IC_DENSITY_WAVE on a uniform n^D mesh (initial.hip).
int launch_density_wave(hipStream_t s, long ncells, int n, int D, int B, int ND, double h, const double* d_xi,
                        const double* d_prm, double* luh);
int launch_entropy_wave(hipStream_t s, long ncells, int n, int D, int B, int ND, double h, const double* d_xi,
                        double t, double* luh);
int launch_fill(hipStream_t s, double* p, long n, double value);

// ---- per-cell linear (acoustic) kernels (linear_cell.hip) ------------------
enum { kLinPredictor = 0, kLinRiemann, kLinFace, kLinUpdate, kLinDt, kLinPredictorSplit };
struct LinearCellArgs {
  int op = 0;
  const DevTables* tab = nullptr;
  const double* luh = nullptr;  // [B^3][6]
  double* luh_out = nullptr;    // update: luh in place
  double dx[3] = {1, 1, 1};
  double dt = 0, cflpnpm = 0;
  const double* sg = nullptr;    // point source: spatial factor [B^3]
  const double* srow = nullptr;  // point source: {sigma0, s_1..s_{B-1}, sum_t w_t f_t};
                                 // split predictor: {sigma0, s_0..s_{B-1}, sum_t w_t f_t}
  double* lduh = nullptr;        // [B^3][4]
  double* lQb = nullptr;         // [6][B^2][6]
  double* lFb = nullptr;         // [6][B^2][4]
  double *FL = nullptr, *FR = nullptr;
  const double *QL = nullptr, *QR = nullptr;
  const double* Fface = nullptr;  // [B^2][4]
  int direction = 0, orientation = 0;
  double* out = nullptr;
};
int launch_linear_cell(int order, const LinearCellArgs& a, hipStream_t s);

// ---- AMR projections (amr.hip), single level, 2D or 3D ---------------------
// Face prolongation: per entry {coarse cell, face, subface s0 + 3 s1, target
// cell}: lQhbnd / lFhbnd of the target's face = the coarse face prolongated
// (overwrite).  Face restriction: per entry {coarse cell, face, first of
// 3^(D-1) consecutive source cells in subface order, -}: the coarse cell's
// lFhbnd on that face = sum of the sources' restricted fluxes (overwrite).
int launch_face_prolong(const KernelCfg& k, hipStream_t s, long n, const int* entries, double* qb, double* fb);
int launch_face_restrict(const KernelCfg& k, hipStream_t s, long n, const int* entries, double* fb);
// Single-level projections of explicit arrays (compatibility entry points):
// face arrays [B^(D-1)][nv], volume arrays [B^D][nv]; sub[k] in 0..2.
// mode 0: out = P(in) (prolongation); mode 1: out += R(in) (restriction).
int launch_face_project(const KernelCfg& k, hipStream_t s, int nv, const double* in, double* out, int s0, int s1,
                        int mode);
int launch_volume_project(const KernelCfg& k, hipStream_t s, int nv, const double* in, double* out, int s0, int s1,
                          int s2, int mode);

// ---- linear acoustics (acoustic.hip) ---------------------------------------
// Device arrays of the acoustic mesh: u[cell][node][4] (p,u,v,w), mat[cell][2]
// (rho, c), qb[cell][6][B*B][4] (time-averaged variables on the faces),
// adjoint cotangents ubar (like u), fbar and qbbar (like qb), gc[cell].
struct AcousticArgs {
  long ncells = 0;
  int n = 0;  // cells per direction (periodic n^3 mesh)
  double h = 0, dt = 0;
  const DevTables* tab = nullptr;
  double* u = nullptr;
  const double* mat = nullptr;
  double* qb = nullptr;
  long src_cell = -1;              // cell holding the point source, -1: none
  const double* src_g = nullptr;   // [B^3] spatial factor of the source moments
  const double* src_tab = nullptr; // [steps][B] time factors per step ([steps][B + 1] split)
  const int* step_ptr = nullptr;   // device step counter
  int step_override = -1;          // >= 0: use this step instead of *step_ptr
  int nrec = 0;
  const int* rcell = nullptr;      // [nrec]
  const double* rw = nullptr;      // [nrec][B^3] interpolation weights
  double* rec = nullptr;           // [steps][nrec]
  double* ubar = nullptr;
  double* fbar = nullptr;
  double* qbbar = nullptr;
  double* gc = nullptr;
  int pred_variant = 0;  // forward predictor: 0 x-line form, 1 node form
  int split_ck = 0;      // split Cauchy-Kowalevski predictor (x-line form only)
};
enum { kAcPredictUpdate = 0, kAcPredictFaces, kAcCorrector, kAcFbar, kAcRiemannAdj, kAcPredictAdj, kAcReceivers };
bool acoustic_order_supported(int order);
int launch_acoustic(int which, int order, const AcousticArgs& a, hipStream_t s);
int launch_acoustic_receivers_adj(int order, const AcousticArgs& a, const double* res, hipStream_t s);
int launch_acoustic_advance(int* step, hipStream_t s);

void set_error(const std::string& msg);

}  // namespace exa

// Host driver of the linear acoustic mesh and its discrete adjoint
// (exa_acoustic_* of include/exahype_amd.h).
//
// Forward step (order of the reference's fused linear ADER-DG step,
// ExaHyPE/exahype/solvers/ADERDGSolver.cpp:887-947 prediction, 1764-1823
// Riemann, 1190-1256 face integrals + update):
//   ac_predict (update)  predictor + volume integral + volume part of the update
//   ac_corrector         Riemann + face integrals + face part of the update
//   ac_receivers         p at the receivers
// Point source: getPointSources / deltaDistribution
// (generic/c/3d/deltaDistribution.cpph:37-146) for a pressure source with a
// Ricker time function.  The moments are separable, PSi[n][node] =
// f(t_n) * g[node] with g = phi(x_ref) / (W_node h^3), so the device needs g
// and, per step, B numbers: sigma_0 = dt sum_t w_t f(t_t) (added to level 0
// and removed by the volume integral, spaceTimePredictorLinear.cpph:111-131,482-503),
// f(t) (level 1, :133-140) and the time derivatives of the interpolant of f
// at the time nodes (levels 2..B-1, :141-178).
//
// Adjoint with checkpointing: states are stored every `every` steps on the
// forward sweep; the backward sweep recomputes each segment from its
// checkpoint, then runs the transposed step backwards through it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exahype_amd.h"
#include "engine.h"
#include "trace.h"

namespace exa {
const KernelCfg& ctx_cfg(const exa_ctx* c);
}

struct exa_acoustic {
  exa::KernelCfg k{};
  exa::Tables tab{};
  int n = 0, B = 0, N = 0, F = 0;
  long nc = 0;
  double h = 0, dt = 0;
  int step = 0;  // host mirror of the device step counter
  hipStream_t stream = nullptr;
  double *d_u = nullptr, *d_mat = nullptr, *d_qb = nullptr;
  double *d_ubar = nullptr, *d_fbar = nullptr, *d_qbbar = nullptr, *d_gc = nullptr;
  int* d_step = nullptr;
  std::vector<double> mat;  // host copy of (rho, c) per cell
  // source
  bool has_src = false;
  double src_x[3] = {0, 0, 0}, amp = 0, f0 = 0, t0 = 0;
  long src_cell = -1;
  double* d_src_g = nullptr;
  double* d_src_tab = nullptr;
  int src_tab_steps = 0;  // steps covered by d_src_tab
  // receivers
  int nrec = 0;
  int* d_rcell = nullptr;
  double* d_rw = nullptr;
  double* d_rec = nullptr;
  long rec_cap = 0;  // doubles
  double* d_res = nullptr;
  long res_cap = 0;
  double ms[3] = {0, 0, 0};
  // per-kernel timing of exa_acoustic_run (HIP events on the launch stream)
  bool timing = false;
  std::vector<hipEvent_t> ev;
  int timed_steps = 0;
  int pred_variant = 0;  // forward predictor: 0 x-line form, 1 node form
};

namespace {

int fail(hipError_t e, const char* what) {
  exa::set_error(std::string(what) + ": " + hipGetErrorString(e));
  return EXA_ERR_HIP;
}

int argerr(const char* what) {
  exa::set_error(what);
  return EXA_ERR_ARG;
}

#define HIP_OK(expr, what)                       \
  do {                                           \
    hipError_t _e = (expr);                      \
    if (_e != hipSuccess) return fail(_e, what); \
  } while (0)

#define RC_OK(expr)          \
  do {                       \
    const int _rc = (expr);  \
    if (_rc) return _rc;     \
  } while (0)

int status(int rc, const char* what) {
  if (rc == -1) {
    exa::set_error(std::string(what) + ": order not compiled into this library");
    return EXA_ERR_UNSUPPORTED;
  }
  if (rc != 0) return fail(static_cast<hipError_t>(rc), what);
  return EXA_OK;
}

double lagrange(const double* x, int B, int k, double s) {
  double r = 1.0;
  for (int m = 0; m < B; m++)
    if (m != k) r *= (s - x[m]) / (x[k] - x[m]);
  return r;
}

double ricker(double t, double f0, double t0) {
  const double pi = 3.14159265358979323846;
  const double a = pi * pi * f0 * f0 * (t - t0) * (t - t0);
  return (1.0 - 2.0 * a) * std::exp(-a);
}

long locate(const exa_acoustic* ac, const double* p, double* xi) {
  int ic[3];
  for (int d = 0; d < 3; d++) {
    ic[d] = static_cast<int>(std::floor(p[d] / ac->h));
    if (ic[d] < 0) ic[d] = 0;
    if (ic[d] >= ac->n) ic[d] = ac->n - 1;
    if (xi) xi[d] = (p[d] - ic[d] * ac->h) / ac->h;
  }
  return (long(ic[2]) * ac->n + ic[1]) * ac->n + ic[0];
}

int grow(double** p, long* cap, long need) {
  if (*cap >= need) return EXA_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIP_OK(hipMalloc(p, sizeof(double) * need), "hipMalloc");
  *cap = need;
  return EXA_OK;
}

// Time factors of the source for steps [0, steps).
int ensure_source_table(exa_acoustic* ac, int steps) {
  if (!ac->has_src || steps <= ac->src_tab_steps) return EXA_OK;
  const int B = ac->B;
  const double* w = ac->tab.w;
  const double* x = ac->tab.x;
  const double* dudx = ac->tab.dudx;
  const double* F0 = ac->tab.F0;
  const int cap = steps + 64;
  const int SW = ac->k.split_ck ? B + 1 : B;  // row width
  std::vector<double> tabh(size_t(cap) * SW);
  std::vector<double> fn(B + 1), psd(B), tmp(B);
  for (int k = 0; k < cap; k++) {
    const double t = k * ac->dt;
    for (int m = 0; m < B + 1; m++) {
      const double tm = (m == 0) ? t : ac->dt * x[m - 1] + t;
      fn[m] = ac->amp * ricker(tm, ac->f0, ac->t0);
    }
    double sum = 0.0;
    for (int m = 1; m < B + 1; m++) sum += w[m - 1] * fn[m];
    double* row = &tabh[size_t(k) * SW];
    if (ac->k.split_ck) {
      // split predictor (fusedSPTVI_linear_split_ck_cpp.template:179-199, 313-331):
      // {dt sum_t w_t f_t, s_0 .. s_{B-1}}, s_k = derPS_k . f at the Gauss time
      // nodes, derPS_0 = FLCoeff, derPS_{k+1}[j] = sum_i derPS_k[i] dudx[i][j]
      row[0] = ac->dt * sum;
      std::vector<double> der(F0, F0 + B), nxt(B);
      for (int q = 0; q < B; q++) {
        double v = 0.0;
        for (int j = 0; j < B; j++) v += der[j] * fn[j + 1];
        row[1 + q] = v;
        for (int j = 0; j < B; j++) {
          double acc = 0.0;
          for (int i = 0; i < B; i++) acc += der[i] * dudx[i * B + j];
          nxt[j] = acc;
        }
        der = nxt;
      }
      continue;
    }
    row[0] = ac->dt * sum;
    if (B > 1) row[1] = fn[0];
    for (int m = 0; m < B; m++) psd[m] = fn[m + 1];
    const double invDt = 1.0 / ac->dt;
    for (int i = 2; i < B; i++) {
      for (int tt = 0; tt < B; tt++) {
        double s = 0.0;
        for (int q = 0; q < B; q++) s += psd[q] * dudx[tt * B + q] * invDt;
        tmp[tt] = s;
      }
      double v = 0.0;
      for (int tt = 0; tt < B; tt++) v += tmp[tt] * F0[tt];
      row[i] = v;
      psd = tmp;
    }
  }
  if (ac->d_src_tab) (void)hipFree(ac->d_src_tab);
  ac->d_src_tab = nullptr;
  ac->src_tab_steps = 0;
  HIP_OK(hipMalloc(&ac->d_src_tab, sizeof(double) * tabh.size()), "hipMalloc(source table)");
  HIP_OK(hipMemcpy(ac->d_src_tab, tabh.data(), sizeof(double) * tabh.size(), hipMemcpyHostToDevice),
         "upload source table");
  ac->src_tab_steps = cap;
  return EXA_OK;
}

exa::AcousticArgs args(exa_acoustic* ac, double* u) {
  exa::AcousticArgs a;
  a.ncells = ac->nc;
  a.n = ac->n;
  a.h = ac->h;
  a.dt = ac->dt;
  a.tab = ac->k.tab;
  a.u = u;
  a.mat = ac->d_mat;
  a.qb = ac->d_qb;
  a.src_cell = ac->has_src ? ac->src_cell : -1;
  a.src_g = ac->d_src_g;
  a.src_tab = ac->d_src_tab;
  a.step_ptr = ac->d_step;
  a.nrec = ac->nrec;
  a.rcell = ac->d_rcell;
  a.rw = ac->d_rw;
  a.ubar = ac->d_ubar;
  a.fbar = ac->d_fbar;
  a.qbbar = ac->d_qbbar;
  a.gc = ac->d_gc;
  a.pred_variant = ac->pred_variant;
  a.split_ck = ac->k.split_ck;
  return a;
}

// one forward step on u (step index `step`, or the device counter if < 0);
// receivers recorded into rec[(step - rec_base) * nrec + r] when rec != nullptr
int enqueue_forward(exa_acoustic* ac, double* u, int step, double* rec, hipEvent_t* ev = nullptr) {
  exa::TraceRange tr("exa.acoustic.step");
  exa::AcousticArgs a = args(ac, u);
  a.step_override = step;
  const int o = ac->k.order;
  if (ev) HIP_OK(hipEventRecord(ev[0], ac->stream), "event");
  RC_OK(status(exa::launch_acoustic(exa::kAcPredictUpdate, o, a, ac->stream), "acoustic predictor"));
  if (ev) HIP_OK(hipEventRecord(ev[1], ac->stream), "event");
  RC_OK(status(exa::launch_acoustic(exa::kAcCorrector, o, a, ac->stream), "acoustic corrector"));
  if (ev) HIP_OK(hipEventRecord(ev[2], ac->stream), "event");
  if (rec && ac->nrec > 0) {
    a.rec = rec;
    RC_OK(status(exa::launch_acoustic(exa::kAcReceivers, o, a, ac->stream), "acoustic receivers"));
  }
  return EXA_OK;
}

void free_acoustic(exa_acoustic* ac) {
  double* bufs[] = {ac->d_u,     ac->d_mat,   ac->d_qb,     ac->d_ubar,  ac->d_fbar, ac->d_qbbar,
                    ac->d_gc,    ac->d_src_g, ac->d_src_tab, ac->d_rw,   ac->d_rec,  ac->d_res};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (ac->d_step) (void)hipFree(ac->d_step);
  if (ac->d_rcell) (void)hipFree(ac->d_rcell);
  for (hipEvent_t e : ac->ev) (void)hipEventDestroy(e);
  if (ac->stream) (void)hipStreamDestroy(ac->stream);
  delete ac;
}

int upload_material(exa_acoustic* ac) {
  HIP_OK(hipMemcpyAsync(ac->d_mat, ac->mat.data(), sizeof(double) * ac->mat.size(), hipMemcpyHostToDevice,
                        ac->stream),
         "upload material");
  HIP_OK(hipStreamSynchronize(ac->stream), "sync");
  return EXA_OK;
}

int set_step(exa_acoustic* ac, int step) {
  ac->step = step;
  HIP_OK(hipMemcpyAsync(ac->d_step, &ac->step, sizeof(int), hipMemcpyHostToDevice, ac->stream), "set step");
  HIP_OK(hipStreamSynchronize(ac->stream), "sync");
  return EXA_OK;
}

}  // namespace

extern "C" {

int exa_acoustic_create(exa_acoustic** out, exa_ctx* ctx, int n, double h) {
  if (!out || !ctx || n < 1 || !(h > 0)) return argerr("exa_acoustic_create: invalid argument");
  *out = nullptr;
  const exa::KernelCfg& k = exa::ctx_cfg(ctx);
  if (k.pde.id != EXA_PDE_ACOUSTIC || k.dim != 3 || k.nvar != 4 || k.npar != 2 || k.precision != EXA_PREC_FP64) {
    exa::set_error("exa_acoustic_create: context must be 3D EXA_PDE_ACOUSTIC, 4 variables, 2 parameters, FP64");
    return EXA_ERR_ARG;
  }
  if (!exa::acoustic_order_supported(k.order)) {
    exa::set_error("exa_acoustic_create: order not compiled into this library");
    return EXA_ERR_UNSUPPORTED;
  }
  exa_acoustic* ac = new exa_acoustic;
  ac->k = k;
  exa::build_tables(k.order, &ac->tab);
  ac->n = n;
  ac->h = h;
  ac->B = k.order + 1;
  ac->N = ac->B * ac->B * ac->B;
  ac->F = ac->B * ac->B;
  ac->nc = long(n) * n * n;
  ac->mat.assign(size_t(2) * ac->nc, 1.0);
  const size_t su = size_t(ac->nc) * ac->N * 4, sq = size_t(ac->nc) * 6 * ac->F * 4;
  hipError_t e = hipStreamCreateWithFlags(&ac->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&ac->d_u, sizeof(double) * su);
  if (e == hipSuccess) e = hipMalloc(&ac->d_mat, sizeof(double) * 2 * ac->nc);
  if (e == hipSuccess) e = hipMalloc(&ac->d_qb, sizeof(double) * sq);
  if (e == hipSuccess) e = hipMalloc(&ac->d_step, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&ac->d_src_g, sizeof(double) * ac->N);
  if (e == hipSuccess) e = hipMemsetAsync(ac->d_u, 0, sizeof(double) * su, ac->stream);
  if (e != hipSuccess) {
    free_acoustic(ac);
    return fail(e, "exa_acoustic_create");
  }
  int rc = upload_material(ac);
  if (!rc) rc = set_step(ac, 0);
  if (!rc) rc = exa_acoustic_set_dt(ac, 0.0);
  if (rc) {
    free_acoustic(ac);
    return rc;
  }
  *out = ac;
  return EXA_OK;
}

int exa_acoustic_destroy(exa_acoustic* ac) {
  if (!ac) return EXA_OK;
  if (ac->stream) (void)hipStreamSynchronize(ac->stream);
  free_acoustic(ac);
  return EXA_OK;
}

long exa_acoustic_ncells(const exa_acoustic* ac) { return ac ? ac->nc : 0; }

int exa_acoustic_set_solution(exa_acoustic* ac, const double* luh) {
  if (!ac || !luh) return argerr("exa_acoustic_set_solution: NULL argument");
  const int N = ac->N;
  std::vector<double> u(size_t(ac->nc) * N * 4);
  std::vector<double> mat(size_t(2) * ac->nc);
  for (long c = 0; c < ac->nc; c++) {
    const double* q = luh + size_t(c) * N * 6;
    const double rho = q[4], cc = q[5];
    for (int nd = 0; nd < N; nd++) {
      const double* qn = q + size_t(nd) * 6;
      if (qn[4] != rho || qn[5] != cc) {
        exa::set_error("exa_acoustic_set_solution: rho and c must be constant within a cell");
        return EXA_ERR_UNSUPPORTED;
      }
      for (int v = 0; v < 4; v++) u[(size_t(c) * N + nd) * 4 + v] = qn[v];
    }
    if (!(rho > 0)) return argerr("exa_acoustic_set_solution: rho must be positive");
    mat[2 * c] = rho;
    mat[2 * c + 1] = cc;
  }
  ac->mat = mat;
  HIP_OK(hipMemcpyAsync(ac->d_u, u.data(), sizeof(double) * u.size(), hipMemcpyHostToDevice, ac->stream),
         "set_solution");
  return upload_material(ac);
}

int exa_acoustic_get_solution(exa_acoustic* ac, double* luh) {
  if (!ac || !luh) return argerr("exa_acoustic_get_solution: NULL argument");
  const int N = ac->N;
  std::vector<double> u(size_t(ac->nc) * N * 4);
  HIP_OK(hipMemcpyAsync(u.data(), ac->d_u, sizeof(double) * u.size(), hipMemcpyDeviceToHost, ac->stream),
         "get_solution");
  HIP_OK(hipStreamSynchronize(ac->stream), "sync");
  for (long c = 0; c < ac->nc; c++)
    for (int nd = 0; nd < N; nd++) {
      double* q = luh + (size_t(c) * N + nd) * 6;
      for (int v = 0; v < 4; v++) q[v] = u[(size_t(c) * N + nd) * 4 + v];
      q[4] = ac->mat[2 * c];
      q[5] = ac->mat[2 * c + 1];
    }
  return EXA_OK;
}

int exa_acoustic_set_material(exa_acoustic* ac, const double* rho, const double* c) {
  if (!ac) return argerr("exa_acoustic_set_material: NULL acoustic");
  for (long i = 0; i < ac->nc; i++) {
    if (rho) {
      if (!(rho[i] > 0)) return argerr("exa_acoustic_set_material: rho must be positive");
      ac->mat[2 * i] = rho[i];
    }
    if (c) ac->mat[2 * i + 1] = c[i];
  }
  return upload_material(ac);
}

int exa_acoustic_get_material(exa_acoustic* ac, double* rho, double* c) {
  if (!ac) return argerr("exa_acoustic_get_material: NULL acoustic");
  for (long i = 0; i < ac->nc; i++) {
    if (rho) rho[i] = ac->mat[2 * i];
    if (c) c[i] = ac->mat[2 * i + 1];
  }
  return EXA_OK;
}

int exa_acoustic_set_dt(exa_acoustic* ac, double dt) {
  if (!ac) return argerr("exa_acoustic_set_dt: NULL acoustic");
  if (!(dt > 0)) {
    double cmax = 0.0;
    for (long i = 0; i < ac->nc; i++) cmax = std::fmax(cmax, std::fabs(ac->mat[2 * i + 1]));
    if (!(cmax > 0)) return argerr("exa_acoustic_set_dt: wave speed is zero everywhere");
    dt = ac->k.cfl * ac->k.pnpm / (3.0 * (cmax / ac->h));
  }
  ac->dt = dt;
  ac->src_tab_steps = 0;  // the source table depends on dt
  return EXA_OK;
}

int exa_acoustic_set_step(exa_acoustic* ac, int step) {
  if (!ac || step < 0) return argerr("exa_acoustic_set_step: invalid argument");
  return set_step(ac, step);
}

int exa_acoustic_time(exa_acoustic* ac, int* step, double* t, double* dt) {
  if (!ac) return argerr("exa_acoustic_time: NULL acoustic");
  if (step) *step = ac->step;
  if (t) *t = ac->step * ac->dt;
  if (dt) *dt = ac->dt;
  return EXA_OK;
}

int exa_acoustic_set_source(exa_acoustic* ac, const double* x, double amp, double f0, double t0) {
  if (!ac) return argerr("exa_acoustic_set_source: NULL acoustic");
  if (amp == 0.0 || !x) {
    ac->has_src = false;
    ac->src_cell = -1;
    return EXA_OK;
  }
  const double L = ac->n * ac->h;
  for (int d = 0; d < 3; d++)
    if (!(x[d] >= 0 && x[d] <= L)) return argerr("exa_acoustic_set_source: source outside the domain");
  ac->has_src = true;
  for (int d = 0; d < 3; d++) ac->src_x[d] = x[d];
  ac->amp = amp;
  ac->f0 = f0;
  ac->t0 = t0;
  ac->src_cell = locate(ac, x, nullptr);
  const long c = ac->src_cell;
  const long ix = c % ac->n, iy = (c / ac->n) % ac->n, iz = c / (long(ac->n) * ac->n);
  const double centre[3] = {(ix + 0.5) * ac->h, (iy + 0.5) * ac->h, (iz + 0.5) * ac->h};
  double xr[3];
  for (int d = 0; d < 3; d++) xr[d] = (x[d] - centre[d]) / ac->h + 0.5;
  const int B = ac->B;
  const double* w = ac->tab.w;
  std::vector<double> g(ac->N);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int kk = 0; kk < B; kk++) {
        const double basis = lagrange(ac->tab.x, B, kk, xr[0]) * lagrange(ac->tab.x, B, j, xr[1]) *
                             lagrange(ac->tab.x, B, i, xr[2]);
        const double winv = 1.0 / (w[kk] * ac->h * w[j] * ac->h * w[i] * ac->h);
        g[(i * B + j) * B + kk] = winv * basis;
      }
  HIP_OK(hipMemcpy(ac->d_src_g, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice), "upload source");
  ac->src_tab_steps = 0;
  return EXA_OK;
}

int exa_acoustic_set_receivers(exa_acoustic* ac, int nrec, const double* xr) {
  if (!ac || nrec < 0 || (nrec > 0 && !xr)) return argerr("exa_acoustic_set_receivers: invalid argument");
  if (ac->d_rcell) (void)hipFree(ac->d_rcell);
  if (ac->d_rw) (void)hipFree(ac->d_rw);
  ac->d_rcell = nullptr;
  ac->d_rw = nullptr;
  ac->nrec = 0;
  if (nrec == 0) return EXA_OK;
  const int B = ac->B, N = ac->N;
  std::vector<int> cells(nrec);
  std::vector<double> wts(size_t(nrec) * N);
  for (int r = 0; r < nrec; r++) {
    double xi[3];
    cells[r] = static_cast<int>(locate(ac, xr + 3 * r, xi));
    double phi[3][exa::kMaxBasis];
    for (int d = 0; d < 3; d++)
      for (int kk = 0; kk < B; kk++) phi[d][kk] = lagrange(ac->tab.x, B, kk, xi[d]);
    for (int z = 0; z < B; z++)
      for (int y = 0; y < B; y++)
        for (int x = 0; x < B; x++) wts[size_t(r) * N + (z * B + y) * B + x] = phi[0][x] * phi[1][y] * phi[2][z];
  }
  HIP_OK(hipMalloc(&ac->d_rcell, sizeof(int) * nrec), "hipMalloc(receivers)");
  HIP_OK(hipMalloc(&ac->d_rw, sizeof(double) * wts.size()), "hipMalloc(receivers)");
  HIP_OK(hipMemcpy(ac->d_rcell, cells.data(), sizeof(int) * nrec, hipMemcpyHostToDevice), "upload receivers");
  HIP_OK(hipMemcpy(ac->d_rw, wts.data(), sizeof(double) * wts.size(), hipMemcpyHostToDevice), "upload receivers");
  ac->nrec = nrec;
  return EXA_OK;
}

int exa_acoustic_run(exa_acoustic* ac, int nsteps, int use_graph, double* rec) {
  if (!ac || nsteps < 0) return argerr("exa_acoustic_run: invalid argument");
  if (nsteps == 0) return EXA_OK;
  const int s0 = ac->step;
  RC_OK(ensure_source_table(ac, s0 + nsteps));
  RC_OK(grow(&ac->d_rec, &ac->rec_cap, long(nsteps) * (ac->nrec > 0 ? ac->nrec : 1)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_OK(hipEventCreate(&e0), "event");
  HIP_OK(hipEventCreate(&e1), "event");
  HIP_OK(hipEventRecord(e0, ac->stream), "event");
  int rc = EXA_OK;
  if (use_graph) {
    // one captured step reading the device step counter; receivers go to
    // d_rec + (step - s0) * nrec through the base pointer of this run
    exa::AcousticArgs a = args(ac, ac->d_u);
    a.rec = ac->d_rec - long(s0) * ac->nrec;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_OK(hipStreamBeginCapture(ac->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    const int o = ac->k.order;
    rc = status(exa::launch_acoustic(exa::kAcPredictUpdate, o, a, ac->stream), "acoustic predictor");
    if (!rc) rc = status(exa::launch_acoustic(exa::kAcCorrector, o, a, ac->stream), "acoustic corrector");
    if (!rc && ac->nrec > 0) rc = status(exa::launch_acoustic(exa::kAcReceivers, o, a, ac->stream), "receivers");
    if (!rc) rc = status(exa::launch_acoustic_advance(ac->d_step, ac->stream), "advance");
    const hipError_t e = hipStreamEndCapture(ac->stream, &g);
    if (!rc && e != hipSuccess) rc = fail(e, "end capture");
    if (!rc) {
      const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      if (ei != hipSuccess) rc = fail(ei, "instantiate");
    }
    for (int i = 0; !rc && i < nsteps; i++) {
      const hipError_t el = hipGraphLaunch(ge, ac->stream);
      if (el != hipSuccess) rc = fail(el, "graph launch");
    }
    (void)hipStreamSynchronize(ac->stream);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
  } else {
    if (ac->timing) {
      while (ac->ev.size() < size_t(3) * nsteps) {
        hipEvent_t e = nullptr;
        HIP_OK(hipEventCreate(&e), "event create");
        ac->ev.push_back(e);
      }
      ac->timed_steps = nsteps;
    }
    for (int i = 0; !rc && i < nsteps; i++) {
      rc = enqueue_forward(ac, ac->d_u, s0 + i, ac->d_rec - long(s0) * ac->nrec,
                           ac->timing ? &ac->ev[size_t(3) * i] : nullptr);
    }
  }
  if (!rc) {
    const hipError_t e = hipEventRecord(e1, ac->stream);
    if (e != hipSuccess) rc = fail(e, "event");
  }
  if (!rc) {
    const hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) rc = fail(e, "sync");
  }
  float ms = 0;
  if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  ac->ms[0] = ms;
  ac->ms[1] = ac->ms[2] = 0;
  RC_OK(set_step(ac, s0 + nsteps));
  if (rec && ac->nrec > 0)
    HIP_OK(hipMemcpy(rec, ac->d_rec, sizeof(double) * nsteps * ac->nrec, hipMemcpyDeviceToHost), "receivers");
  return EXA_OK;
}

int exa_acoustic_misfit_gradient(exa_acoustic* ac, int nsteps, int every, const double* obs, double* J,
                                 double* grad_c) {
  if (!ac || nsteps < 1 || every < 1 || !obs || !J || !grad_c)
    return argerr("exa_acoustic_misfit_gradient: invalid argument");
  if (ac->nrec < 1) return argerr("exa_acoustic_misfit_gradient: no receivers");
  if (ac->k.split_ck) {
    exa::set_error("exa_acoustic_misfit_gradient: the adjoint transposes the generic predictor (context dialect "
                   "EXA_DIALECT_SPLIT_CK)");
    return EXA_ERR_UNSUPPORTED;
  }
  const int s0 = ac->step, K = nsteps, R = ac->nrec, o = ac->k.order;
  const int nck = (K + every - 1) / every;
  const int seg = every < K ? every : K;
  const size_t su = size_t(ac->nc) * ac->N * 4, sq = size_t(ac->nc) * 6 * ac->F * 4;
  RC_OK(ensure_source_table(ac, s0 + K));
  RC_OK(grow(&ac->d_rec, &ac->rec_cap, long(K) * R));
  RC_OK(grow(&ac->d_res, &ac->res_cap, long(K) * R));
  if (!ac->d_ubar) {
    HIP_OK(hipMalloc(&ac->d_ubar, sizeof(double) * su), "hipMalloc(adjoint)");
    HIP_OK(hipMalloc(&ac->d_fbar, sizeof(double) * sq), "hipMalloc(adjoint)");
    HIP_OK(hipMalloc(&ac->d_qbbar, sizeof(double) * sq), "hipMalloc(adjoint)");
    HIP_OK(hipMalloc(&ac->d_gc, sizeof(double) * ac->nc), "hipMalloc(adjoint)");
  }
  double* d_ck = nullptr;
  double* d_seg = nullptr;
  HIP_OK(hipMalloc(&d_ck, sizeof(double) * su * nck), "hipMalloc(checkpoints)");
  hipError_t ea = hipMalloc(&d_seg, sizeof(double) * su * seg);
  if (ea != hipSuccess) {
    (void)hipFree(d_ck);
    return fail(ea, "hipMalloc(segment)");
  }
  std::vector<hipEvent_t> ev(2 + 3 * nck, nullptr);
  int rc = EXA_OK;
  for (auto& e : ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) rc = fail(hipErrorUnknown, "event");
  double* rec_base = ac->d_rec - long(s0) * R;
  // forward sweep with checkpoints
  if (!rc && hipEventRecord(ev[0], ac->stream) != hipSuccess) rc = fail(hipErrorUnknown, "event");
  exa::TraceRange tr_fwd("exa.acoustic.adjoint.forward");
  for (int k = 0; !rc && k < K; k++) {
    if (k % every == 0) {
      const hipError_t e = hipMemcpyAsync(d_ck + su * (k / every), ac->d_u, sizeof(double) * su,
                                          hipMemcpyDeviceToDevice, ac->stream);
      if (e != hipSuccess) rc = fail(e, "checkpoint");
    }
    if (!rc) rc = enqueue_forward(ac, ac->d_u, s0 + k, rec_base);
  }
  if (!rc && hipEventRecord(ev[1], ac->stream) != hipSuccess) rc = fail(hipErrorUnknown, "event");
  std::vector<double> rec(size_t(K) * R), res(size_t(K) * R);
  if (!rc) {
    const hipError_t e = hipMemcpyAsync(rec.data(), ac->d_rec, sizeof(double) * rec.size(), hipMemcpyDeviceToHost,
                                        ac->stream);
    if (e != hipSuccess) rc = fail(e, "receivers");
    else if (hipStreamSynchronize(ac->stream) != hipSuccess) rc = fail(hipErrorUnknown, "sync");
  }
  double Jv = 0.0;
  for (size_t i = 0; i < rec.size(); i++) {
    res[i] = rec[i] - obs[i];
    Jv += 0.5 * res[i] * res[i];
  }
  if (!rc) {
    hipError_t e = hipMemcpyAsync(ac->d_res, res.data(), sizeof(double) * res.size(), hipMemcpyHostToDevice,
                                  ac->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ac->d_ubar, 0, sizeof(double) * su, ac->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ac->d_gc, 0, sizeof(double) * ac->nc, ac->stream);
    if (e != hipSuccess) rc = fail(e, "adjoint init");
  }
  // backward sweep, segment by segment
  exa::TraceRange tr_bwd("exa.acoustic.adjoint.backward");
  for (int j = nck - 1; !rc && j >= 0; j--) {
    const int k0 = j * every, k1 = (k0 + every < K) ? k0 + every : K;
    (void)hipEventRecord(ev[2 + 3 * j], ac->stream);
    hipError_t e = hipMemcpyAsync(d_seg, d_ck + su * j, sizeof(double) * su, hipMemcpyDeviceToDevice, ac->stream);
    for (int k = k0; e == hipSuccess && !rc && k < k1 - 1; k++) {
      double* dst = d_seg + su * (k - k0 + 1);
      e = hipMemcpyAsync(dst, d_seg + su * (k - k0), sizeof(double) * su, hipMemcpyDeviceToDevice, ac->stream);
      if (e == hipSuccess) rc = enqueue_forward(ac, dst, s0 + k, nullptr);
    }
    if (e != hipSuccess && !rc) rc = fail(e, "recompute");
    (void)hipEventRecord(ev[3 + 3 * j], ac->stream);
    for (int k = k1 - 1; !rc && k >= k0; k--) {
      double* uk = d_seg + su * (k - k0);
      exa::AcousticArgs a = args(ac, uk);
      a.step_override = s0 + k;
      rc = status(exa::launch_acoustic_receivers_adj(o, a, ac->d_res + long(k) * R, ac->stream), "receiver adjoint");
      if (!rc) rc = status(exa::launch_acoustic(exa::kAcPredictFaces, o, a, ac->stream), "acoustic faces");
      if (!rc) rc = status(exa::launch_acoustic(exa::kAcFbar, o, a, ac->stream), "adjoint fbar");
      if (!rc) rc = status(exa::launch_acoustic(exa::kAcRiemannAdj, o, a, ac->stream), "adjoint riemann");
      if (!rc) rc = status(exa::launch_acoustic(exa::kAcPredictAdj, o, a, ac->stream), "adjoint predictor");
    }
    (void)hipEventRecord(ev[4 + 3 * j], ac->stream);
  }
  if (!rc) {
    const hipError_t e = hipMemcpyAsync(grad_c, ac->d_gc, sizeof(double) * ac->nc, hipMemcpyDeviceToHost, ac->stream);
    if (e != hipSuccess) rc = fail(e, "gradient");
    else if (hipStreamSynchronize(ac->stream) != hipSuccess) rc = fail(hipErrorUnknown, "sync");
  }
  if (!rc) {
    float f = 0;
    (void)hipEventElapsedTime(&f, ev[0], ev[1]);
    ac->ms[0] = f;
    ac->ms[1] = ac->ms[2] = 0;
    for (int j = 0; j < nck; j++) {
      float a1 = 0, a2 = 0;
      (void)hipEventElapsedTime(&a1, ev[2 + 3 * j], ev[3 + 3 * j]);
      (void)hipEventElapsedTime(&a2, ev[3 + 3 * j], ev[4 + 3 * j]);
      ac->ms[1] += a1;
      ac->ms[2] += a2;
    }
  }
  (void)hipStreamSynchronize(ac->stream);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipFree(d_ck);
  (void)hipFree(d_seg);
  if (rc) return rc;
  *J = Jv;
  return set_step(ac, s0 + K);
}

int exa_acoustic_set_timing(exa_acoustic* ac, int enabled) {
  if (!ac) return argerr("exa_acoustic_set_timing: NULL acoustic");
  ac->timing = enabled != 0;
  ac->timed_steps = 0;
  return EXA_OK;
}

int exa_acoustic_set_predictor(exa_acoustic* ac, int variant) {
  if (!ac) return argerr("exa_acoustic_set_predictor: NULL acoustic");
  if (variant != 0 && variant != 1) return argerr("exa_acoustic_set_predictor: variant must be 0 (x-line) or 1 (node)");
  if (variant == 1 && ac->k.split_ck) {
    exa::set_error("exa_acoustic_set_predictor: the split Cauchy-Kowalevski predictor has the x-line form only");
    return EXA_ERR_UNSUPPORTED;
  }
  ac->pred_variant = variant;
  return EXA_OK;
}

int exa_acoustic_phase_times(exa_acoustic* ac, double* ms_predictor, double* ms_corrector, int* nsteps) {
  if (!ac) return argerr("exa_acoustic_phase_times: NULL acoustic");
  double a = 0, b = 0;
  for (int i = 0; i < ac->timed_steps; i++) {
    float x = 0, y = 0;
    HIP_OK(hipEventElapsedTime(&x, ac->ev[size_t(3) * i], ac->ev[size_t(3) * i + 1]), "elapsed");
    HIP_OK(hipEventElapsedTime(&y, ac->ev[size_t(3) * i + 1], ac->ev[size_t(3) * i + 2]), "elapsed");
    a += x;
    b += y;
  }
  if (ms_predictor) *ms_predictor = a;
  if (ms_corrector) *ms_corrector = b;
  if (nsteps) *nsteps = ac->timed_steps;
  return EXA_OK;
}

int exa_acoustic_last_times(exa_acoustic* ac, double* ms_forward, double* ms_recompute, double* ms_adjoint) {
  if (!ac) return argerr("exa_acoustic_last_times: NULL acoustic");
  if (ms_forward) *ms_forward = ac->ms[0];
  if (ms_recompute) *ms_recompute = ac->ms[1];
  if (ms_adjoint) *ms_adjoint = ac->ms[2];
  return EXA_OK;
}

}  // extern "C"

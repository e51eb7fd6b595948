// C ABI of the engine: contexts, per-cell compatibility entry points,
// batched device entry points.  See include/exahype_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/exahype_amd.h"
#include "engine.h"

namespace exa {

bool shape_supported(int dim, int order, int nvar, int npar);

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

}  // namespace exa

using exa::set_error;

struct exa_ctx {
  exa::KernelCfg k{};
  exa::Tables tab{};
  exa::DevTables* d_tab = nullptr;
  hipStream_t stream = nullptr;
  // scratch for the per-cell compatibility entry points
  double* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  // point source of the linear path (per-cell predictor)
  bool has_src = false;
  double src_x[3] = {0, 0, 0}, src_amp = 0, src_f0 = 0, src_t0 = 0;
  // boundaryValues hook of exa_boundaryConditions (the Euler example's default)
  int bc_kind = EXA_BC_DIRICHLET_ENTROPY_WAVE;
  // vector width (doubles) of the optimised dialect's padded arrays
  int vect = 1;
  // Picard-guard list of the batched predictor: {listed, listed in total, list[guard_cap]}
  int* d_guard = nullptr;
  long guard_cap = 0;
  // batched entry points: d_h holds one edge length per cell (0) or dx[D] per cell (dim)
  int cell_dx = 0;
};

namespace {

int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return EXA_ERR_HIP;
}

#define HIP_OK(expr, what)                   \
  do {                                       \
    hipError_t _e = (expr);                  \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

int launch_status(int rc, const char* what) {
  if (rc == -1) {
    set_error(std::string(what) + ": shape not compiled into this library");
    return EXA_ERR_UNSUPPORTED;
  }
  if (rc != 0) return hip_fail(static_cast<hipError_t>(rc), what);
  return EXA_OK;
}

int ensure_scratch(exa_ctx* c, size_t bytes) {
  if (c->scratch_bytes >= bytes) return EXA_OK;
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  c->d_scratch = nullptr;
  c->scratch_bytes = 0;
  HIP_OK(hipMalloc(&c->d_scratch, bytes), "hipMalloc(scratch)");
  c->scratch_bytes = bytes;
  return EXA_OK;
}

// Picard-guard list of exa_batch_predictor_noalloc, in device memory that no
// call allocates: {listed this call, listed in total, list[kNoAllocMaxCells]}
// (4 MiB).  Each GPU the library runs on holds its own copy until the process
// ends; host code uses the per-device address from hipGetSymbolAddress, never
// the array's name.
constexpr long kNoAllocMaxCells = 1L << 20;
__device__ __align__(256) int scratch[2 + kNoAllocMaxCells];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int* g_scratch[kMaxDevices] = {};

// Device address of the scratch array on the device of `stream` (the current
// device for the null stream), looked up once per device with that device current.
int scratch_for_stream(hipStream_t stream, int** out) {
  int cur = 0;
  HIP_OK(hipGetDevice(&cur), "hipGetDevice");
  int dev = cur;
  if (stream) HIP_OK(hipStreamGetDevice(stream, &dev), "hipStreamGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("exa_batch_predictor_noalloc: device index beyond the library's table");
    return EXA_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    if (dev != cur) HIP_OK(hipSetDevice(dev), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (dev != cur) HIP_OK(hipSetDevice(cur), "hipSetDevice");
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = static_cast<int*>(p);
  }
  *out = g_scratch[dev];
  return EXA_OK;
}

// Clears the list length at the start of a call (a kernel, so that a captured
// call records no memset node).
__global__ void guard_clear_kernel(int* guard) {
  if (threadIdx.x == 0) guard[0] = 0;
}

// Guard buffer for ncells cells (grown on demand, counters cleared when allocated).
int ensure_guard(exa_ctx* c, long ncells) {
  if (c->guard_cap >= ncells) return EXA_OK;
  if (c->d_guard) (void)hipFree(c->d_guard);
  c->d_guard = nullptr;
  c->guard_cap = 0;
  HIP_OK(hipMalloc(&c->d_guard, sizeof(int) * (2 + ncells)), "hipMalloc(guard)");
  HIP_OK(hipMemset(c->d_guard, 0, 2 * sizeof(int)), "clear guard");
  c->guard_cap = ncells;
  return EXA_OK;
}

int nspace(const exa_ctx* c) {
  const int B = c->k.order + 1;
  return c->k.dim == 3 ? B * B * B : B * B;
}
int nface(const exa_ctx* c) {
  const int B = c->k.order + 1;
  return c->k.dim == 3 ? B * B : B;
}
int ndata(const exa_ctx* c) { return c->k.nvar + c->k.npar; }

bool is_linear(const exa_ctx* c) { return c->k.pde.id == EXA_PDE_ACOUSTIC; }

// getSizeWithPadding (KernelGenerator/kernelgenerator/controller.py:190-192):
// the next multiple of the vector width
int padded(int n, int vect) { return vect * ((n + vect - 1) / vect); }

// One converter of converter_cpp.template: nblk blocks of nrow rows, each row
// holding n values; the generic rows are n apart, the optimised rows npad.
void convert_rows(const double* in, double* out, long rows, int n, int ld_in, int ld_out) {
  for (long r = 0; r < rows; r++)
    for (int v = 0; v < n; v++) out[r * ld_out + v] = in[r * ld_in + v];
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

// Point-source data of one cell (getPointSources / deltaDistribution,
// generic/c/3d/deltaDistribution.cpph:37-146): returns false when the source
// is not inside [centre - dx/2, centre + dx/2] (closed, as the reference).
bool point_source_terms(const exa_ctx* c, const double* centre, const double* dx, double t, double dt,
                        std::vector<double>& g, std::vector<double>& row) {
  if (!c->has_src) return false;
  for (int d = 0; d < 3; d++)
    if (!(centre[d] - 0.5 * dx[d] <= c->src_x[d] && c->src_x[d] <= centre[d] + 0.5 * dx[d])) return false;
  const int B = c->k.order + 1;
  const double* w = c->tab.w;
  const double* x = c->tab.x;
  double xr[3];
  for (int d = 0; d < 3; d++) xr[d] = (c->src_x[d] - centre[d]) / dx[d] + 0.5;
  g.assign(size_t(B) * B * B, 0.0);
  for (int i = 0; i < B; i++)
    for (int j = 0; j < B; j++)
      for (int k = 0; k < B; k++)
        g[(i * B + j) * B + k] = 1.0 / (w[k] * dx[0] * w[j] * dx[1] * w[i] * dx[2]) *
                                 (lagrange(x, B, k, xr[0]) * lagrange(x, B, j, xr[1]) * lagrange(x, B, i, xr[2]));
  std::vector<double> fn(B + 1), psd(B), tmp(B);
  for (int m = 0; m < B + 1; m++) fn[m] = c->src_amp * ricker(m == 0 ? t : dt * x[m - 1] + t, c->src_f0, c->src_t0);
  if (c->k.split_ck) {
    // split predictor (fusedSPTVI_linear_split_ck_cpp.template:179-199, 313-331):
    // {dt sum_t w_t f_t, s_0 .. s_{B-1}, sum_t w_t f_t}, s_k = derPS_k . f with
    // derPS_0 = FLCoeff and derPS_{k+1}[j] = sum_i derPS_k[i] dudx[i][j]
    row.assign(B + 2, 0.0);
    double sum = 0.0;
    for (int m = 0; m < B; m++) sum += w[m] * fn[m + 1];
    row[0] = dt * sum;
    std::vector<double> der(c->tab.F0, c->tab.F0 + B), nxt(B);
    for (int k = 0; k < B; k++) {
      double v = 0.0;
      for (int j = 0; j < B; j++) v += der[j] * fn[j + 1];
      row[1 + k] = v;
      for (int j = 0; j < B; j++) {
        double a = 0.0;
        for (int i = 0; i < B; i++) a += der[i] * c->tab.dudx[i * B + j];
        nxt[j] = a;
      }
      der = nxt;
    }
    row[B + 1] = sum;
    return true;
  }
  row.assign(B + 1, 0.0);
  double sum = 0.0;
  for (int m = 1; m < B + 1; m++) sum += w[m - 1] * fn[m];
  row[0] = dt * sum;
  if (B > 1) row[1] = fn[0];
  for (int m = 0; m < B; m++) psd[m] = fn[m + 1];
  for (int i = 2; i < B; i++) {
    for (int tt = 0; tt < B; tt++) {
      double s2 = 0.0;
      for (int q = 0; q < B; q++) s2 += psd[q] * c->tab.dudx[tt * B + q] * (1.0 / dt);
      tmp[tt] = s2;
    }
    double v = 0.0;
    for (int tt = 0; tt < B; tt++) v += tmp[tt] * c->tab.F0[tt];
    row[i] = v;
    psd = tmp;
  }
  row[B] = sum;
  return true;
}

}  // namespace

extern "C" {

const char* exa_last_error(void) { return exa::g_error.c_str(); }

int exa_ctx_create(exa_ctx** out, int dim, int order, int nvar, int npar, int pde, int precision,
                   int picard_max, double cfl) {
  if (!out) {
    set_error("exa_ctx_create: out is NULL");
    return EXA_ERR_ARG;
  }
  *out = nullptr;
  if (dim != 2 && dim != 3) {
    set_error("exa_ctx_create: dim must be 2 or 3");
    return EXA_ERR_ARG;
  }
  if (pde != EXA_PDE_EULER && pde != EXA_PDE_ADVECTION && pde != EXA_PDE_ACOUSTIC && pde != EXA_PDE_USER) {
    set_error("exa_ctx_create: unknown pde");
    return EXA_ERR_ARG;
  }
  if (precision < EXA_PREC_FP64 || precision > EXA_PREC_BF16) {
    set_error("exa_ctx_create: unknown precision");
    return EXA_ERR_ARG;
  }
  if (precision == EXA_PREC_BF16) {
    // the BF16 operand form is compiled for the 3D Euler order-5 tile kernel
    // with a fixed Picard count (BF16 iterates cannot meet the 1e-7 rule)
    if (dim != 3 || order != 5 || nvar != 5 || npar != 0 || pde != EXA_PDE_EULER) {
      set_error("exa_ctx_create: BF16 predictor compiled for 3D Euler order 5 (nvar 5, npar 0) only");
      return EXA_ERR_UNSUPPORTED;
    }
    if (picard_max <= 0) {
      set_error("exa_ctx_create: the BF16 predictor needs a fixed Picard count (picard_max > 0)");
      return EXA_ERR_ARG;
    }
  }
  if (!(cfl > 0)) {
    set_error("exa_ctx_create: cfl must be positive");
    return EXA_ERR_ARG;
  }
  if (pde == EXA_PDE_ACOUSTIC) {
    if (dim != 3 || nvar != 4 || npar != 2) {
      set_error("exa_ctx_create: EXA_PDE_ACOUSTIC needs dim 3, nvar 4 (p,u,v,w), npar 2 (rho,c)");
      return EXA_ERR_ARG;
    }
    if (precision != EXA_PREC_FP64 || !exa::acoustic_order_supported(order)) {
      set_error("exa_ctx_create: acoustic path compiled for FP64 and orders 1..5 only");
      return EXA_ERR_UNSUPPORTED;
    }
  } else if (!exa::shape_supported(dim, order, nvar, npar)) {
    set_error("exa_ctx_create: (dim, order, nvar, npar) not compiled into this library");
    return EXA_ERR_UNSUPPORTED;
  } else if (pde == EXA_PDE_EULER && nvar != 5) {
    set_error("exa_ctx_create: EXA_PDE_EULER has 5 variables");
    return EXA_ERR_ARG;
  } else if (pde == EXA_PDE_USER && nvar != exa::kUserNVar) {
    set_error("exa_ctx_create: EXA_PDE_USER has the variable count of the compiled ExaUserPde (kNVar, default 5)");
    return EXA_ERR_ARG;
  }
  exa_ctx* c = new exa_ctx;
  if (!exa::build_tables(order, &c->tab)) {
    delete c;
    set_error("exa_ctx_create: cannot build tables for this order");
    return EXA_ERR_ARG;
  }
  c->k.dim = dim;
  c->k.order = order;
  c->k.nvar = nvar;
  c->k.npar = npar;
  c->k.precision = precision;
  c->k.picard_max = picard_max;
  c->k.cfl = cfl;
  c->k.pnpm = exa::pnpm_for_order(order);
  // relative: the FP32 form's accuracy target (its divergences carry ~1e-8 relative rounding)
  c->k.guard_tol = 1e-6;
  c->k.guard_recount = 0;
  c->k.pde.id = pde;
  c->k.pde.a[0] = 1.0;
  c->k.pde.a[1] = 0.5;
  c->k.pde.a[2] = 0.25;
  exa::DevTables h{};
  const int B = order + 1;
  for (int i = 0; i < B; i++) {
    h.w[i] = c->tab.w[i];
    h.x[i] = c->tab.x[i];
    h.F0[i] = c->tab.F0[i];
    h.F1[i] = c->tab.F1[i];
    h.wf[i] = static_cast<float>(c->tab.w[i]);
    h.iw[i] = 1.0 / c->tab.w[i];
    h.F0f[i] = static_cast<float>(c->tab.F0[i]);
    h.F1f[i] = static_cast<float>(c->tab.F1[i]);
  }
  for (int i = 0; i < B * B; i++) {
    h.Kxi[i] = c->tab.Kxi[i];
    h.iK1[i] = c->tab.iK1[i];
    h.dudx[i] = c->tab.dudx[i];
    h.Kxif[i] = static_cast<float>(c->tab.Kxi[i]);
    h.iK1f[i] = static_cast<float>(c->tab.iK1[i]);
    h.dudxf[i] = static_cast<float>(c->tab.dudx[i]);
    h.Pt[i] = c->tab.iK1[i] * c->tab.w[i % B];
    h.Ptf[i] = static_cast<float>(h.Pt[i]);
  }
  double cm[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < B; j++) {
    double cj = 0.0;
    for (int i = 0; i < B; i++) cj += c->tab.w[i] * c->tab.iK1[i * B + j] * c->tab.w[j];
    h.cwf[j] = static_cast<float>(cj);
    cm[0] += cj;
    cm[1] += cj * c->tab.x[j];
    cm[2] += cj * c->tab.x[j] * c->tab.x[j];
  }
  for (int k = 0; k < 3; k++) h.cmom[k] = static_cast<float>(cm[k]);
  for (int i = 0; i < 3 * B * B; i++) h.fgp[i] = c->tab.fgp[i];
  hipError_t e = hipMalloc(&c->d_tab, sizeof(exa::DevTables));
  if (e == hipSuccess) e = hipMemcpy(c->d_tab, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (c->d_tab) (void)hipFree(c->d_tab);
    delete c;
    return hip_fail(e, "exa_ctx_create");
  }
  c->k.tab = c->d_tab;
  *out = c;
  return EXA_OK;
}

int exa_ctx_destroy(exa_ctx* c) {
  if (!c) return EXA_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_guard) (void)hipFree(c->d_guard);
  if (c->d_scratch) (void)hipFree(c->d_scratch);
  if (c->d_tab) (void)hipFree(c->d_tab);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return EXA_OK;
}

int exa_ctx_set_pde_params(exa_ctx* c, const double* prm, int n) {
  if (!c || n < 0 || n > exa::kUserParams || (n > 0 && !prm)) {
    set_error("exa_ctx_set_pde_params: invalid argument (at most 16 parameters)");
    return EXA_ERR_ARG;
  }
  for (int i = 0; i < exa::kUserParams; i++) c->k.pde.u[i] = i < n ? prm[i] : 0.0;
  return EXA_OK;
}

int exa_ctx_set_advection(exa_ctx* c, const double* a) {
  if (!c || !a) {
    set_error("exa_ctx_set_advection: NULL argument");
    return EXA_ERR_ARG;
  }
  for (int d = 0; d < c->k.dim; d++) c->k.pde.a[d] = a[d];
  return EXA_OK;
}

int exa_ctx_set_cell_dx(exa_ctx* c, int per_direction) {
  if (!c || (per_direction != 0 && per_direction != 1)) {
    set_error("exa_ctx_set_cell_dx: invalid argument (0 or 1)");
    return EXA_ERR_ARG;
  }
  if (is_linear(c)) {
    set_error("exa_ctx_set_cell_dx: the batched entry points serve the nonlinear path");
    return EXA_ERR_UNSUPPORTED;
  }
  c->cell_dx = per_direction ? c->k.dim : 0;
  return EXA_OK;
}

int exa_ctx_set_point_source(exa_ctx* c, const double* x, double amp, double f0, double t0) {
  if (!c) {
    set_error("exa_ctx_set_point_source: NULL context");
    return EXA_ERR_ARG;
  }
  if (!is_linear(c)) {
    set_error("exa_ctx_set_point_source: point sources are compiled for the linear path only");
    return EXA_ERR_UNSUPPORTED;
  }
  c->has_src = x != nullptr && amp != 0.0;
  if (x)
    for (int d = 0; d < 3; d++) c->src_x[d] = x[d];
  c->src_amp = amp;
  c->src_f0 = f0;
  c->src_t0 = t0;
  return EXA_OK;
}

int exa_ctx_set_boundary_values(exa_ctx* c, int kind) {
  if (!c || (kind != EXA_BC_WALL && kind != EXA_BC_DIRICHLET_ENTROPY_WAVE)) {
    set_error("exa_ctx_set_boundary_values: invalid argument (EXA_BC_WALL or EXA_BC_DIRICHLET_ENTROPY_WAVE)");
    return EXA_ERR_ARG;
  }
  if (!exa::boundary_supported(c->k)) {
    set_error("exa_ctx_set_boundary_values: boundary values are defined for the Euler equations only");
    return EXA_ERR_UNSUPPORTED;
  }
  c->bc_kind = kind;
  return EXA_OK;
}

int exa_ctx_set_dialect(exa_ctx* c, int dialect, int vect_size) {
  if (!c || dialect < EXA_DIALECT_GENERIC || dialect > EXA_DIALECT_SPLIT_CK ||
      !(vect_size == 1 || vect_size == 2 || vect_size == 4 || vect_size == 8)) {
    set_error("exa_ctx_set_dialect: invalid argument (dialect 0..3, vect_size 1, 2, 4 or 8)");
    return EXA_ERR_ARG;
  }
  if (is_linear(c) && (dialect == EXA_DIALECT_OPTIMISED || dialect == EXA_DIALECT_OPTIMISED_CERK)) {
    set_error("exa_ctx_set_dialect: the optimised Picard dialect covers the nonlinear kernels");
    return EXA_ERR_UNSUPPORTED;
  }
  if (!is_linear(c) && dialect == EXA_DIALECT_SPLIT_CK) {
    set_error("exa_ctx_set_dialect: the split Cauchy-Kowalevski predictor is linear-only");
    return EXA_ERR_UNSUPPORTED;
  }
  c->k.dialect = (dialect == EXA_DIALECT_OPTIMISED || dialect == EXA_DIALECT_OPTIMISED_CERK) ? 1 : 0;
  c->k.guess = dialect == EXA_DIALECT_OPTIMISED_CERK ? 1 : 0;
  c->k.split_ck = dialect == EXA_DIALECT_SPLIT_CK ? 1 : 0;
  c->vect = vect_size;
  return EXA_OK;
}

int exa_ctx_set_initial_guess(exa_ctx* c, int guess) {
  if (!c || (guess != EXA_GUESS_TRIVIAL && guess != EXA_GUESS_CERK)) {
    set_error("exa_ctx_set_initial_guess: invalid argument");
    return EXA_ERR_ARG;
  }
  if (is_linear(c) && guess != EXA_GUESS_TRIVIAL) {
    set_error("exa_ctx_set_initial_guess: the CERK guess is part of the nonlinear predictor");
    return EXA_ERR_UNSUPPORTED;
  }
  c->k.guess = guess;
  return EXA_OK;
}

int exa_ctx_padded_size(const exa_ctx* c, int which, int* size) {
  if (!c || !size || (which != 0 && which != 1)) {
    set_error("exa_ctx_padded_size: invalid argument");
    return EXA_ERR_ARG;
  }
  *size = padded(which ? ndata(c) : c->k.nvar, c->vect);
  return EXA_OK;
}

int exa_ctx_set_picard_guard(exa_ctx* c, double tol, int recompute_count) {
  if (!c || recompute_count < 0 || recompute_count > 2 * (c->k.order + 1) || !(tol == tol)) {
    set_error("exa_ctx_set_picard_guard: invalid argument");
    return EXA_ERR_ARG;
  }
  c->k.guard_tol = tol;
  c->k.guard_recount = recompute_count;
  return EXA_OK;
}

int exa_ctx_set_predictor_variant(exa_ctx* c, int variant) {
  if (!c || variant < 0 || variant > 4) {
    set_error("exa_ctx_set_predictor_variant: invalid argument");
    return EXA_ERR_ARG;
  }
  c->k.pred_variant = variant;
  return EXA_OK;
}

int exa_tables(int order, double* w, double* x, double* Kxi, double* iK1, double* dudx, double* F0, double* F1,
               double* fgp, double* egp) {
  exa::Tables t;
  if (!exa::build_tables(order, &t)) {
    set_error("exa_tables: unsupported order");
    return EXA_ERR_ARG;
  }
  const int B = order + 1;
  if (w) std::memcpy(w, t.w, sizeof(double) * B);
  if (x) std::memcpy(x, t.x, sizeof(double) * B);
  if (Kxi) std::memcpy(Kxi, t.Kxi, sizeof(double) * B * B);
  if (iK1) std::memcpy(iK1, t.iK1, sizeof(double) * B * B);
  if (dudx) std::memcpy(dudx, t.dudx, sizeof(double) * B * B);
  if (F0) std::memcpy(F0, t.F0, sizeof(double) * B);
  if (F1) std::memcpy(F1, t.F1, sizeof(double) * B);
  if (fgp) std::memcpy(fgp, t.fgp, sizeof(double) * 3 * B * B);
  if (egp) std::memcpy(egp, t.egp, sizeof(double) * B * B);
  return EXA_OK;
}

int exa_ctx_tables(const exa_ctx* c, double* w, double* x, double* Kxi, double* iK1, double* dudx, double* F0,
                   double* F1, double* fgp) {
  if (!c) {
    set_error("exa_ctx_tables: NULL context");
    return EXA_ERR_ARG;
  }
  const int B = c->k.order + 1;
  if (w) std::memcpy(w, c->tab.w, sizeof(double) * B);
  if (x) std::memcpy(x, c->tab.x, sizeof(double) * B);
  if (Kxi) std::memcpy(Kxi, c->tab.Kxi, sizeof(double) * B * B);
  if (iK1) std::memcpy(iK1, c->tab.iK1, sizeof(double) * B * B);
  if (dudx) std::memcpy(dudx, c->tab.dudx, sizeof(double) * B * B);
  if (F0) std::memcpy(F0, c->tab.F0, sizeof(double) * B);
  if (F1) std::memcpy(F1, c->tab.F1, sizeof(double) * B);
  if (fgp) std::memcpy(fgp, c->tab.fgp, sizeof(double) * 3 * B * B);
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Per-cell compatibility entry points
// ---------------------------------------------------------------------------

int exa_fusedSpaceTimePredictorVolumeIntegral(exa_ctx* c, double* lduh, double* lQhbnd, double* lGradQhbnd,
                                              double* lFhbnd, double* luh, const double* centre,
                                              const double* dx, double t, double dt,
                                              int addVolumeIntegralResultToUpdate, int* picard_iterations) {
  (void)lGradQhbnd;
  if (!c || !lduh || !lQhbnd || !lFhbnd || !luh || !dx) {
    set_error("exa_fusedSpaceTimePredictorVolumeIntegral: NULL argument");
    return EXA_ERR_ARG;
  }
  if (is_linear(c)) {
    // linear glue (AbstractGenericADERDGSolverImplementation.template:114-165):
    // point sources of the cell, spaceTimePredictorLinear, volumeIntegralLinear
    if (!centre) {
      set_error("exa_fusedSpaceTimePredictorVolumeIntegral: centre is required by the linear path");
      return EXA_ERR_ARG;
    }
    const int NS = nspace(c), NF = nface(c);
    std::vector<double> g, row;
    const bool src = point_source_terms(c, centre, dx, t, dt, g, row);
    const size_t nL = size_t(NS) * 6, nD = size_t(NS) * 4, nQ = 6 * size_t(NF) * 6, nF = 6 * size_t(NF) * 4;
    const size_t total = nL + nD + nQ + nF + (src ? g.size() + row.size() : 0);
    int rc = ensure_scratch(c, total * sizeof(double));
    if (rc) return rc;
    double* d_luh = c->d_scratch;
    double* d_lduh = d_luh + nL;
    double* d_qb = d_lduh + nD;
    double* d_fb = d_qb + nQ;
    double* d_g = d_fb + nF;
    double* d_row = d_g + (src ? g.size() : 0);
    HIP_OK(hipMemcpyAsync(d_luh, luh, nL * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy luh");
    if (src) {
      HIP_OK(hipMemcpyAsync(d_g, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
      HIP_OK(hipMemcpyAsync(d_row, row.data(), row.size() * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    }
    exa::LinearCellArgs a;
    a.op = c->k.split_ck ? exa::kLinPredictorSplit : exa::kLinPredictor;
    a.tab = c->d_tab;
    a.luh = d_luh;
    for (int d = 0; d < 3; d++) a.dx[d] = dx[d];
    a.dt = dt;
    a.sg = src ? d_g : nullptr;
    a.srow = src ? d_row : nullptr;
    a.lduh = d_lduh;
    a.lQb = d_qb;
    a.lFb = d_fb;
    rc = launch_status(exa::launch_linear_cell(c->k.order, a, c->stream), "linear predictor");
    if (rc) return rc;
    if (!addVolumeIntegralResultToUpdate) {
      exa::LinearCellArgs u;
      u.op = exa::kLinUpdate;
      u.tab = c->d_tab;
      u.luh_out = d_luh;
      u.lduh = d_lduh;
      u.dt = dt;
      rc = launch_status(exa::launch_linear_cell(c->k.order, u, c->stream), "update");
      if (rc) return rc;
      HIP_OK(hipMemcpyAsync(luh, d_luh, nL * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy luh");
    } else {
      HIP_OK(hipMemcpyAsync(lduh, d_lduh, nD * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lduh");
    }
    HIP_OK(hipMemcpyAsync(lQhbnd, d_qb, nQ * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lQhbnd");
    HIP_OK(hipMemcpyAsync(lFhbnd, d_fb, nF * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lFhbnd");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    if (picard_iterations) *picard_iterations = 0;
    return EXA_OK;
  }
  const int NS = nspace(c), NF = nface(c), ND = ndata(c), V = c->k.nvar, D = c->k.dim;
  for (int d = 0; d < D; d++)
    if (!(dx[d] > 0)) {
      set_error("exa_fusedSpaceTimePredictorVolumeIntegral: dx must be positive");
      return EXA_ERR_ARG;
    }
  bool cubic = true;
  for (int d = 1; d < D; d++) cubic = cubic && dx[d] == dx[0];
  const size_t nLuh = size_t(NS) * ND, nLduh = size_t(NS) * V;
  const size_t nQb = size_t(2 * D) * NF * ND, nFb = size_t(2 * D) * NF * V;
  const size_t total = nLuh + nLduh + nQb + nFb + 3 + 1 + 1 + 2;
  int rc = ensure_scratch(c, total * sizeof(double));
  if (rc) return rc;
  double* d_luh = c->d_scratch;
  double* d_lduh = d_luh + nLuh;
  double* d_qb = d_lduh + nLduh;
  double* d_fb = d_qb + nQb;
  double* d_h = d_fb + nFb;  // dx[D] (one value for cubic cells)
  double* d_dt = d_h + 3;
  int* d_pic = reinterpret_cast<int*>(d_dt + 1);
  int* d_guard = reinterpret_cast<int*>(d_dt + 2);  // {count, total, list[1]}
  HIP_OK(hipMemcpyAsync(d_luh, luh, nLuh * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy luh");
  HIP_OK(hipMemcpyAsync(d_h, dx, (cubic ? 1 : D) * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy dx");
  HIP_OK(hipMemcpyAsync(d_dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream), "copy dt");
  HIP_OK(hipMemsetAsync(d_guard, 0, 2 * sizeof(double), c->stream), "clear guard");
  exa::LaunchOpts po;
  po.guard = d_guard;
  po.hdim = cubic ? 0 : D;  // non-cubic cells: FP64 node-column or slab predictor, dx per direction
  rc = launch_status(exa::launch_predictor(c->k, c->stream, 1, d_luh, d_h, d_dt, d_lduh, d_qb, d_fb, d_pic, po),
                     "predictor");
  if (rc) return rc;
  if (!addVolumeIntegralResultToUpdate) {
    // luh += dt/w * volume integral (AbstractGenericADERDGSolverImplementation.template:206-210)
    rc = launch_status(exa::launch_solution_update(c->k, c->stream, d_luh, d_luh, d_lduh, d_dt), "update");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(luh, d_luh, nLuh * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy luh");
  } else {
    HIP_OK(hipMemcpyAsync(lduh, d_lduh, nLduh * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lduh");
  }
  HIP_OK(hipMemcpyAsync(lQhbnd, d_qb, nQb * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lQhbnd");
  HIP_OK(hipMemcpyAsync(lFhbnd, d_fb, nFb * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lFhbnd");
  int pic = 0;
  HIP_OK(hipMemcpyAsync(&pic, d_pic, sizeof(int), hipMemcpyDeviceToHost, c->stream), "copy picard");
  HIP_OK(hipStreamSynchronize(c->stream), "sync");
  if (picard_iterations) *picard_iterations = pic;
  return EXA_OK;
}

int exa_opt_fusedSpaceTimePredictorVolumeIntegral(exa_ctx* c, double* lduh, double* lQhbnd, double* lFhbnd,
                                                  const double* luh, double inverseDx, double dt,
                                                  int* picard_iterations) {
  if (!c || !lduh || !lQhbnd || !lFhbnd || !luh || !(inverseDx > 0)) {
    set_error("exa_opt_fusedSpaceTimePredictorVolumeIntegral: invalid argument");
    return EXA_ERR_ARG;
  }
  if (!c->k.dialect) {
    set_error("exa_opt_fusedSpaceTimePredictorVolumeIntegral: context not in an optimised dialect");
    return EXA_ERR_ARG;
  }
  const int D = c->k.dim, NS = nspace(c), NF = nface(c), V = c->k.nvar, ND = ndata(c);
  const int VP = padded(V, c->vect), NDP = padded(ND, c->vect);
  std::vector<double> u(luh, luh + size_t(NS) * ND), du(size_t(NS) * V), qb(size_t(2 * D) * NF * ND),
      fb(size_t(2 * D) * NF * V);
  const double dx[3] = {1.0 / inverseDx, 1.0 / inverseDx, 1.0 / inverseDx};
  const int rc = exa_fusedSpaceTimePredictorVolumeIntegral(c, du.data(), qb.data(), nullptr, fb.data(), u.data(),
                                                          nullptr, dx, 0.0, dt, 1, picard_iterations);
  if (rc) return rc;
  // the generated kernel zeroes lduh, lQhbnd and lFhbnd (padding included) first
  std::fill(lduh, lduh + size_t(NS) * VP, 0.0);
  std::fill(lQhbnd, lQhbnd + size_t(2 * D) * NF * NDP, 0.0);
  std::fill(lFhbnd, lFhbnd + size_t(2 * D) * NF * VP, 0.0);
  convert_rows(du.data(), lduh, NS, V, V, VP);
  convert_rows(qb.data(), lQhbnd, long(2 * D) * NF, ND, ND, NDP);
  convert_rows(fb.data(), lFhbnd, long(2 * D) * NF, V, V, VP);
  return EXA_OK;
}

// converter_cpp.template: rows of the array (row count, values per row: nVar or nData)
#define EXA_CONVERTER(NAME, ROWS, DATA)                                                          \
  int exa_converter_##NAME##_optimised2generic(const exa_ctx* c, const double* opt, double* gen) { \
    if (!c || !opt || !gen) {                                                                    \
      set_error("exa_converter_" #NAME "_optimised2generic: NULL argument");                     \
      return EXA_ERR_ARG;                                                                        \
    }                                                                                            \
    const int n = (DATA) ? ndata(c) : c->k.nvar;                                                 \
    convert_rows(opt, gen, (ROWS), n, padded(n, c->vect), n);                                    \
    return EXA_OK;                                                                               \
  }                                                                                              \
  int exa_converter_##NAME##_generic2optimised(const exa_ctx* c, const double* gen, double* opt) { \
    if (!c || !opt || !gen) {                                                                    \
      set_error("exa_converter_" #NAME "_generic2optimised: NULL argument");                     \
      return EXA_ERR_ARG;                                                                        \
    }                                                                                            \
    const int n = (DATA) ? ndata(c) : c->k.nvar;                                                 \
    convert_rows(gen, opt, (ROWS), n, n, padded(n, c->vect));                                    \
    return EXA_OK;                                                                               \
  }

EXA_CONVERTER(Qhbnd, long(2 * c->k.dim) * nface(c), 1)
EXA_CONVERTER(Fhbnd, long(2 * c->k.dim) * nface(c), 0)
EXA_CONVERTER(QFace, long(nface(c)), 1)
EXA_CONVERTER(FFace, long(nface(c)), 0)
EXA_CONVERTER(Qhi, long(nspace(c)), 1)
EXA_CONVERTER(Fhi, long(c->k.dim + 1) * nspace(c), 0)
EXA_CONVERTER(PSi, long(c->k.order + 2) * nspace(c), 0)
EXA_CONVERTER(Lduh, long(nspace(c)), 0)
#undef EXA_CONVERTER

int exa_riemannSolver(exa_ctx* c, double* FL, double* FR, const double* QL, const double* QR, double t, double dt,
                      const double* dx, int direction, int isBoundaryFace, int faceIndex) {
  (void)t;
  (void)dt;
  (void)dx;
  (void)isBoundaryFace;
  (void)faceIndex;
  if (!c || !FL || !FR || !QL || !QR) {
    set_error("exa_riemannSolver: NULL argument");
    return EXA_ERR_ARG;
  }
  if (direction < 0 || direction >= c->k.dim) {
    set_error("exa_riemannSolver: direction out of range");
    return EXA_ERR_ARG;
  }
  if (is_linear(c)) {  // riemannSolverLinear: FL, FR are outputs (fluctuations)
    const int NF = nface(c);
    const size_t nq = size_t(NF) * 6, nf = size_t(NF) * 4;
    int rc = ensure_scratch(c, (2 * nq + 2 * nf) * sizeof(double));
    if (rc) return rc;
    double* d_ql = c->d_scratch;
    double* d_qr = d_ql + nq;
    double* d_fl = d_qr + nq;
    double* d_fr = d_fl + nf;
    HIP_OK(hipMemcpyAsync(d_ql, QL, nq * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy QL");
    HIP_OK(hipMemcpyAsync(d_qr, QR, nq * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy QR");
    exa::LinearCellArgs a;
    a.op = exa::kLinRiemann;
    a.tab = c->d_tab;
    a.FL = d_fl;
    a.FR = d_fr;
    a.QL = d_ql;
    a.QR = d_qr;
    a.direction = direction;
    rc = launch_status(exa::launch_linear_cell(c->k.order, a, c->stream), "linear riemann");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(FL, d_fl, nf * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy FL");
    HIP_OK(hipMemcpyAsync(FR, d_fr, nf * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy FR");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    return EXA_OK;
  }
  const int NF = nface(c), ND = ndata(c), V = c->k.nvar, D = c->k.dim;
  // two pseudo cells: cell 0 holds the left state on face 2d+1, cell 1 the right state on face 2d
  const size_t qcell = size_t(2 * D) * NF * ND, fcell = size_t(2 * D) * NF * V;
  const size_t total = 2 * qcell + 2 * fcell + 2;
  int rc = ensure_scratch(c, total * sizeof(double));
  if (rc) return rc;
  double* d_qb = c->d_scratch;
  double* d_fb = d_qb + 2 * qcell;
  int* d_face = reinterpret_cast<int*>(d_fb + 2 * fcell);
  const int d = direction;
  const size_t ql = (size_t(2 * d + 1)) * NF * ND, qr = qcell + size_t(2 * d) * NF * ND;
  const size_t fl = (size_t(2 * d + 1)) * NF * V, fr = fcell + size_t(2 * d) * NF * V;
  const int face[4] = {0, 1, d, 0};
  HIP_OK(hipMemcpyAsync(d_qb + ql, QL, NF * ND * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy QL");
  HIP_OK(hipMemcpyAsync(d_qb + qr, QR, NF * ND * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy QR");
  HIP_OK(hipMemcpyAsync(d_fb + fl, FL, NF * V * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy FL");
  HIP_OK(hipMemcpyAsync(d_fb + fr, FR, NF * V * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy FR");
  HIP_OK(hipMemcpyAsync(d_face, face, sizeof(face), hipMemcpyHostToDevice, c->stream), "copy face");
  rc = launch_status(exa::launch_riemann(c->k, c->stream, 1, d_face, d_qb, d_fb), "riemann");
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(FL, d_fb + fl, NF * V * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy FL");
  HIP_OK(hipMemcpyAsync(FR, d_fb + fr, NF * V * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy FR");
  HIP_OK(hipStreamSynchronize(c->stream), "sync");
  return EXA_OK;
}

int exa_boundaryConditions(exa_ctx* c, double* fluxIn, const double* stateIn, const double* gradStateIn,
                           const double* luh, const double* cellCentre, const double* cellSize, double t, double dt,
                           int direction, int orientation) {
  (void)gradStateIn;  // no viscous flux on this path
  (void)luh;          // unused by the generic kernel as well
  if (!c || !fluxIn || !stateIn || !cellCentre || !cellSize) {
    set_error("exa_boundaryConditions: NULL argument");
    return EXA_ERR_ARG;
  }
  if (direction < 0 || direction >= c->k.dim || (orientation != 0 && orientation != 1)) {
    set_error("exa_boundaryConditions: direction or orientation out of range");
    return EXA_ERR_ARG;
  }
  if (!exa::boundary_supported(c->k)) {
    set_error("exa_boundaryConditions: boundary values are defined for the Euler equations only");
    return EXA_ERR_UNSUPPORTED;
  }
  const int NF = nface(c), ND = ndata(c), V = c->k.nvar, D = c->k.dim;
  // pseudo cell 0 holds the inside face data on its face 2d + o; ghost face 0 the outside
  const size_t qcell = size_t(2 * D) * NF * ND, fcell = size_t(2 * D) * NF * V;
  const size_t face_at = (qcell + fcell + size_t(NF) * (ND + V) + 1) & ~size_t(1);  // int4 needs 16-byte alignment
  int rc = ensure_scratch(c, (face_at + 2) * sizeof(double));
  if (rc) return rc;
  double* d_qb = c->d_scratch;
  double* d_fb = d_qb + qcell;
  double* d_gq = d_fb + fcell;
  double* d_gf = d_gq + size_t(NF) * ND;
  int* d_face = reinterpret_cast<int*>(d_qb + face_at);
  const int d = direction, f = 2 * d + orientation;
  // outside on the left of a left face, on the right of a right face
  const int face[4] = {0, 0, d, orientation == 0 ? 2 : 1};
  HIP_OK(hipMemcpyAsync(d_qb + size_t(f) * NF * ND, stateIn, NF * ND * sizeof(double), hipMemcpyHostToDevice,
                        c->stream),
         "copy stateIn");
  HIP_OK(hipMemcpyAsync(d_fb + size_t(f) * NF * V, fluxIn, NF * V * sizeof(double), hipMemcpyHostToDevice, c->stream),
         "copy fluxIn");
  HIP_OK(hipMemcpyAsync(d_face, face, sizeof(face), hipMemcpyHostToDevice, c->stream), "copy face");
  rc = launch_status(exa::launch_boundary_face(c->k, c->stream, c->bc_kind, cellCentre, cellSize, t, dt, d,
                                               orientation, d_qb + size_t(f) * NF * ND, d_gq, d_gf),
                     "boundary values");
  if (rc) return rc;
  exa::LaunchOpts o;
  o.ghost_q = d_gq;
  o.ghost_f = d_gf;
  rc = launch_status(exa::launch_riemann(c->k, c->stream, 1, d_face, d_qb, d_fb, o), "riemann");
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(fluxIn, d_fb + size_t(f) * NF * V, NF * V * sizeof(double), hipMemcpyDeviceToHost, c->stream),
         "copy fluxIn");
  HIP_OK(hipStreamSynchronize(c->stream), "sync");
  return EXA_OK;
}

int exa_faceIntegral(exa_ctx* c, double* out, const double* lFhbnd, int direction, int orientation,
                     const int* subfaceIndex, int levelDelta, const double* dx, double dt, int addToUpdate) {
  if (!c || !out || !lFhbnd || !dx) {
    set_error("exa_faceIntegral: NULL argument");
    return EXA_ERR_ARG;
  }
  if (direction < 0 || direction >= c->k.dim || orientation < 0 || orientation > 1) {
    set_error("exa_faceIntegral: direction/orientation out of range");
    return EXA_ERR_ARG;
  }
  if (levelDelta < 0 || levelDelta > 8 || (levelDelta > 0 && !subfaceIndex)) {
    set_error("exa_faceIntegral: levelDelta out of range or subfaceIndex missing");
    return EXA_ERR_ARG;
  }
  if (is_linear(c)) {  // faceIntegralLinear (no sign factor); coarse-fine faces restrict first
    if (levelDelta > 0 || !addToUpdate) {
      set_error("exa_faceIntegral: the linear path supports conforming faces added to the update");
      return EXA_ERR_UNSUPPORTED;
    }
    const int NS = nspace(c), NF = nface(c);
    int rc = ensure_scratch(c, (size_t(NS) * 4 + size_t(NF) * 4) * sizeof(double));
    if (rc) return rc;
    double* d_lduh = c->d_scratch;
    double* d_F = d_lduh + size_t(NS) * 4;
    HIP_OK(hipMemcpyAsync(d_lduh, out, size_t(NS) * 4 * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    HIP_OK(hipMemcpyAsync(d_F, lFhbnd, size_t(NF) * 4 * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    exa::LinearCellArgs a;
    a.op = exa::kLinFace;
    a.tab = c->d_tab;
    a.lduh = d_lduh;
    a.Fface = d_F;
    a.direction = direction;
    a.orientation = orientation;
    for (int d = 0; d < 3; d++) a.dx[d] = dx[d];
    rc = launch_status(exa::launch_linear_cell(c->k.order, a, c->stream), "linear face integral");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(out, d_lduh, size_t(NS) * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    return EXA_OK;
  }
  const int NS = nspace(c), NF = nface(c), ND = ndata(c), V = c->k.nvar;
  const size_t total = size_t(NS) * ND + size_t(NS) * V + 3 * size_t(NF) * V + 1;
  int rc = ensure_scratch(c, total * sizeof(double));
  if (rc) return rc;
  double* d_luh = c->d_scratch;
  double* d_lduh = d_luh + size_t(NS) * ND;
  double* d_F = d_lduh + size_t(NS) * V;
  double* d_t1 = d_F + size_t(NF) * V;
  double* d_t2 = d_t1 + size_t(NF) * V;
  double* d_dt = d_t2 + size_t(NF) * V;
  HIP_OK(hipMemcpyAsync(d_F, lFhbnd, NF * V * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy F");
  if (levelDelta > 0) {
    // faceIntegral glue with levelDelta > 0
    // (AbstractGenericADERDGSolverImplementation.template:265-270): restrict the
    // fine flux over levelDelta levels (faceUnknownsRestriction, lowest
    // significance first) and integrate it with the coarse cell size into out.
    int cur[2] = {subfaceIndex[0], c->k.dim == 3 ? subfaceIndex[1] : 0};
    const double* src = d_F;
    double* dst = d_t1;
    for (int l = 1; l <= levelDelta; l++) {
      const int s0 = cur[0] % 3, s1 = cur[1] % 3;
      cur[0] /= 3;
      cur[1] /= 3;
      HIP_OK(hipMemsetAsync(dst, 0, NF * V * sizeof(double), c->stream), "zero");
      rc = launch_status(exa::launch_face_project(c->k, c->stream, V, src, dst, s0, s1, 1), "face restriction");
      if (rc) return rc;
      src = dst;
      dst = (dst == d_t1) ? d_t2 : d_t1;
    }
    double scale = 1.0;
    for (int l = 0; l < levelDelta; l++) scale *= 3.0;
    HIP_OK(hipMemcpyAsync(d_lduh, out, NS * V * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy lduh");
    rc = launch_status(exa::launch_face_integral(c->k, c->stream, d_lduh, src, direction, orientation,
                                                 1.0 / (scale * dx[direction])),
                       "face integral");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(out, d_lduh, NS * V * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lduh");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    return EXA_OK;
  }
  if (addToUpdate) {
    HIP_OK(hipMemcpyAsync(d_lduh, out, NS * V * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy lduh");
  } else {
    HIP_OK(hipMemsetAsync(d_lduh, 0, NS * V * sizeof(double), c->stream), "zero lduh");
    HIP_OK(hipMemcpyAsync(d_luh, out, NS * ND * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy luh");
    HIP_OK(hipMemcpyAsync(d_dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream), "copy dt");
  }
  rc = launch_status(
      exa::launch_face_integral(c->k, c->stream, d_lduh, d_F, direction, orientation, 1.0 / dx[direction]),
      "face integral");
  if (rc) return rc;
  if (addToUpdate) {
    HIP_OK(hipMemcpyAsync(out, d_lduh, NS * V * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy lduh");
  } else {
    rc = launch_status(exa::launch_solution_update(c->k, c->stream, d_luh, d_luh, d_lduh, d_dt), "update");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(out, d_luh, NS * ND * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy luh");
  }
  HIP_OK(hipStreamSynchronize(c->stream), "sync");
  return EXA_OK;
}

int exa_surfaceIntegral(exa_ctx* c, double* lduh, const double* lFhbnd, const double* dx) {
  if (!c || !lduh || !lFhbnd || !dx) {
    set_error("exa_surfaceIntegral: NULL argument");
    return EXA_ERR_ARG;
  }
  const int NF = nface(c), V = c->k.nvar, D = c->k.dim;
  for (int d = 0; d < D; d++)
    for (int o = 0; o < 2; o++) {
      const int rc = exa_faceIntegral(c, lduh, lFhbnd + size_t(2 * d + o) * NF * V, d, o, nullptr, 0, dx, 0.0, 1);
      if (rc) return rc;
    }
  return EXA_OK;
}

int exa_addUpdateToSolution(exa_ctx* c, double* luh, const double* luhOld, const double* lduh, double dt) {
  if (!c || !luh || !luhOld || !lduh) {
    set_error("exa_addUpdateToSolution: NULL argument");
    return EXA_ERR_ARG;
  }
  if (is_linear(c)) {
    const int NS = nspace(c);
    int rc = ensure_scratch(c, size_t(NS) * 10 * sizeof(double));
    if (rc) return rc;
    double* d_luh = c->d_scratch;
    double* d_lduh = d_luh + size_t(NS) * 6;
    HIP_OK(hipMemcpyAsync(d_luh, luhOld, size_t(NS) * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    HIP_OK(hipMemcpyAsync(d_lduh, lduh, size_t(NS) * 4 * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    exa::LinearCellArgs a;
    a.op = exa::kLinUpdate;
    a.tab = c->d_tab;
    a.luh_out = d_luh;
    a.lduh = d_lduh;
    a.dt = dt;
    rc = launch_status(exa::launch_linear_cell(c->k.order, a, c->stream), "update");
    if (rc) return rc;
    std::vector<double> tmp(size_t(NS) * 6);
    HIP_OK(hipMemcpyAsync(tmp.data(), d_luh, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    for (int n = 0; n < NS; n++)
      for (int v = 0; v < 4; v++) luh[n * 6 + v] = tmp[size_t(n) * 6 + v];
    return EXA_OK;
  }
  const int NS = nspace(c), ND = ndata(c), V = c->k.nvar;
  const size_t total = 2 * size_t(NS) * ND + size_t(NS) * V + 1;
  int rc = ensure_scratch(c, total * sizeof(double));
  if (rc) return rc;
  double* d_luh = c->d_scratch;
  double* d_old = d_luh + size_t(NS) * ND;
  double* d_lduh = d_old + size_t(NS) * ND;
  double* d_dt = d_lduh + size_t(NS) * V;
  HIP_OK(hipMemcpyAsync(d_old, luhOld, NS * ND * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
  HIP_OK(hipMemcpyAsync(d_luh, luhOld, NS * ND * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
  HIP_OK(hipMemcpyAsync(d_lduh, lduh, NS * V * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
  HIP_OK(hipMemcpyAsync(d_dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
  rc = launch_status(exa::launch_solution_update(c->k, c->stream, d_luh, d_old, d_lduh, d_dt), "update");
  if (rc) return rc;
  // parameters of luh are left untouched (solutionUpdate.cpph:55-60 updates variables only)
  std::vector<double> tmp(size_t(NS) * ND);
  HIP_OK(hipMemcpyAsync(tmp.data(), d_luh, NS * ND * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
  HIP_OK(hipStreamSynchronize(c->stream), "sync");
  for (int s = 0; s < NS; s++)
    for (int v = 0; v < V; v++) luh[s * ND + v] = tmp[size_t(s) * ND + v];
  return EXA_OK;
}

int exa_stableTimeStepSize(exa_ctx* c, const double* luh, const double* dx, double* dt_out) {
  if (!c || !luh || !dx || !dt_out) {
    set_error("exa_stableTimeStepSize: NULL argument");
    return EXA_ERR_ARG;
  }
  if (is_linear(c)) {
    const int NS = nspace(c);
    int rc = ensure_scratch(c, (size_t(NS) * 6 + 1) * sizeof(double));
    if (rc) return rc;
    double* d_luh = c->d_scratch;
    double* d_out = d_luh + size_t(NS) * 6;
    HIP_OK(hipMemcpyAsync(d_luh, luh, size_t(NS) * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
    exa::LinearCellArgs a;
    a.op = exa::kLinDt;
    a.tab = c->d_tab;
    a.luh = d_luh;
    for (int d = 0; d < 3; d++) a.dx[d] = dx[d];
    a.cflpnpm = c->k.cfl * c->k.pnpm;
    a.out = d_out;
    rc = launch_status(exa::launch_linear_cell(c->k.order, a, c->stream), "stable dt");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(dt_out, d_out, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    return EXA_OK;
  }
  const int NS = nspace(c), ND = ndata(c), D = c->k.dim;
  for (int d = 0; d < D; d++)
    if (!(dx[d] > 0)) {
      set_error("exa_stableTimeStepSize: dx must be positive");
      return EXA_ERR_ARG;
    }
  const size_t total = size_t(NS) * ND + 4;
  int rc = ensure_scratch(c, total * sizeof(double));
  if (rc) return rc;
  double* d_luh = c->d_scratch;
  double* d_h = d_luh + size_t(NS) * ND;  // dx[D]
  double* d_out = d_h + 3;
  const double init[4] = {dx[0], D > 1 ? dx[1] : 0.0, D > 2 ? dx[2] : 0.0, DBL_MAX};
  HIP_OK(hipMemcpyAsync(d_luh, luh, NS * ND * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
  HIP_OK(hipMemcpyAsync(d_h, init, 4 * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy");
  exa::LaunchOpts o;
  o.hdim = D;  // dx per direction
  rc = launch_status(exa::launch_stable_dt(c->k, c->stream, 1, d_luh, d_h, d_out, o), "stable dt");
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(dt_out, d_out, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy");
  HIP_OK(hipStreamSynchronize(c->stream), "sync");
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Batched device entry points
// ---------------------------------------------------------------------------

// ---- AMR projections (per cell), generic/c/{2d,3d}/amrRoutines.cpph ------------
namespace {

// Multi-level projection of one face (K = dim-1) or volume (K = dim) array
// of nv values per node.  Prolongation decodes sub from the most significant
// base-3 digit (amrRoutines.cpph:69-90), restriction from the least (:196-224).
int project_levels(exa_ctx* c, bool face, int nv, const double* h_in, double* h_out, const int* sub, int delta,
                   bool prolong) {
  const int K = face ? c->k.dim - 1 : c->k.dim;
  const int n = (face ? nface(c) : nspace(c)) * nv;
  int rc = ensure_scratch(c, 3 * size_t(n) * sizeof(double));
  if (rc) return rc;
  double* a = c->d_scratch;
  double* b = a + n;
  HIP_OK(hipMemcpyAsync(a, h_in, n * sizeof(double), hipMemcpyHostToDevice, c->stream), "copy in");
  int cur[3] = {sub[0], K > 1 ? sub[1] : 0, K > 2 ? sub[2] : 0};
  for (int l = 1; l <= delta; l++) {
    int si[3] = {0, 0, 0};
    if (prolong) {
      int sig = 1;
      for (int q = 0; q < delta - l; q++) sig *= 3;
      for (int k = 0; k < K; k++) {
        si[k] = cur[k] / sig;
        cur[k] %= sig;
      }
    } else {
      for (int k = 0; k < K; k++) {
        si[k] = cur[k] % 3;
        cur[k] /= 3;
      }
      HIP_OK(hipMemsetAsync(b, 0, n * sizeof(double), c->stream), "zero");
    }
    const int mode = prolong ? 0 : 1;
    rc = face ? exa::launch_face_project(c->k, c->stream, nv, a, b, si[0], si[1], mode)
              : exa::launch_volume_project(c->k, c->stream, nv, a, b, si[0], si[1], si[2], mode);
    rc = launch_status(rc, "projection");
    if (rc) return rc;
    std::swap(a, b);
  }
  if (prolong) {
    HIP_OK(hipMemcpyAsync(h_out, a, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy out");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
  } else {
    std::vector<double> tmp(n);
    HIP_OK(hipMemcpyAsync(tmp.data(), a, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy out");
    HIP_OK(hipStreamSynchronize(c->stream), "sync");
    for (int i = 0; i < n; i++) h_out[i] += tmp[i];  // the reference accumulates into the coarse array
  }
  return EXA_OK;
}

int check_levels(exa_ctx* c, const void* a, const void* b, const void* d, const int* sub, int lc, int lf,
                 const char* what) {
  if (!c || !a || !b || !d || !sub) {
    set_error(std::string(what) + ": NULL argument");
    return EXA_ERR_ARG;
  }
  if (lf <= lc || lf - lc > 8) {
    set_error(std::string(what) + ": fineGridLevel must exceed coarseGridLevel (by at most 8)");
    return EXA_ERR_ARG;
  }
  return EXA_OK;
}

}  // namespace

int exa_faceUnknownsProlongation(exa_ctx* c, double* lQhbndFine, double* lFhbndFine, const double* lQhbndCoarse,
                                 const double* lFhbndCoarse, int coarseGridLevel, int fineGridLevel,
                                 const int* subfaceIndex) {
  int rc = check_levels(c, lQhbndFine, lFhbndFine, lQhbndCoarse, subfaceIndex, coarseGridLevel, fineGridLevel,
                        "exa_faceUnknownsProlongation");
  if (rc) return rc;
  if (!lFhbndCoarse) return check_levels(c, nullptr, nullptr, nullptr, nullptr, 0, 0, "exa_faceUnknownsProlongation");
  const int delta = fineGridLevel - coarseGridLevel;
  rc = project_levels(c, true, ndata(c), lQhbndCoarse, lQhbndFine, subfaceIndex, delta, true);
  if (!rc) rc = project_levels(c, true, c->k.nvar, lFhbndCoarse, lFhbndFine, subfaceIndex, delta, true);
  return rc;
}

int exa_faceUnknownsRestriction(exa_ctx* c, double* lQhbndCoarse, double* lFhbndCoarse, const double* lQhbndFine,
                                const double* lFhbndFine, int coarseGridLevel, int fineGridLevel,
                                const int* subfaceIndex) {
  int rc = check_levels(c, lQhbndCoarse, lFhbndCoarse, lQhbndFine, subfaceIndex, coarseGridLevel, fineGridLevel,
                        "exa_faceUnknownsRestriction");
  if (rc) return rc;
  if (!lFhbndFine) return check_levels(c, nullptr, nullptr, nullptr, nullptr, 0, 0, "exa_faceUnknownsRestriction");
  const int delta = fineGridLevel - coarseGridLevel;
  rc = project_levels(c, true, ndata(c), lQhbndFine, lQhbndCoarse, subfaceIndex, delta, false);
  if (!rc) rc = project_levels(c, true, c->k.nvar, lFhbndFine, lFhbndCoarse, subfaceIndex, delta, false);
  return rc;
}

int exa_volumeUnknownsProlongation(exa_ctx* c, double* luhFine, const double* luhCoarse, int coarseGridLevel,
                                   int fineGridLevel, const int* subcellIndex) {
  const int rc = check_levels(c, luhFine, luhCoarse, luhCoarse, subcellIndex, coarseGridLevel, fineGridLevel,
                              "exa_volumeUnknownsProlongation");
  if (rc) return rc;
  return project_levels(c, false, ndata(c), luhCoarse, luhFine, subcellIndex, fineGridLevel - coarseGridLevel, true);
}

int exa_volumeUnknownsRestriction(exa_ctx* c, double* luhCoarse, const double* luhFine, int coarseGridLevel,
                                  int fineGridLevel, const int* subcellIndex) {
  const int rc = check_levels(c, luhCoarse, luhFine, luhFine, subcellIndex, coarseGridLevel, fineGridLevel,
                              "exa_volumeUnknownsRestriction");
  if (rc) return rc;
  return project_levels(c, false, ndata(c), luhFine, luhCoarse, subcellIndex, fineGridLevel - coarseGridLevel, false);
}

int exa_batch_predictor(exa_ctx* c, void* stream, long ncells, const double* d_luh, const double* d_h,
                        const double* d_dt, double* d_lduh, double* d_lQhbnd, double* d_lFhbnd, int* d_picard) {
  if (!c || ncells < 0 || (ncells > 0 && (!d_luh || !d_h || !d_dt || !d_lduh || !d_lQhbnd || !d_lFhbnd))) {
    set_error("exa_batch_predictor: invalid argument");
    return EXA_ERR_ARG;
  }
  exa::LaunchOpts o;
  o.hdim = c->cell_dx;
  if (ncells > 0 && exa::predictor_uses_guard(c->k)) {
    const int rc = ensure_guard(c, ncells);
    if (rc) return rc;
    o.guard = c->d_guard;
  }
  return launch_status(exa::launch_predictor(c->k, static_cast<hipStream_t>(stream), ncells, d_luh, d_h, d_dt,
                                             d_lduh, d_lQhbnd, d_lFhbnd, d_picard, o),
                       "exa_batch_predictor");
}

int exa_batch_predictor_noalloc(exa_ctx* c, void* stream, long ncells, const double* d_luh, const double* d_h,
                                const double* d_dt, double* d_lduh, double* d_lQhbnd, double* d_lFhbnd,
                                int* d_picard) {
  if (!c || ncells < 0 || (ncells > 0 && (!d_luh || !d_h || !d_dt || !d_lduh || !d_lQhbnd || !d_lFhbnd))) {
    set_error("exa_batch_predictor_noalloc: invalid argument");
    return EXA_ERR_ARG;
  }
  if (ncells > kNoAllocMaxCells) {
    set_error("exa_batch_predictor_noalloc: more than 1048576 cells (the static guard list's size)");
    return EXA_ERR_ARG;
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  exa::LaunchOpts o;
  o.hdim = c->cell_dx;
  if (ncells > 0 && exa::predictor_uses_guard(c->k)) {
    int* g = nullptr;
    const int rc = scratch_for_stream(st, &g);
    if (rc) return rc;
    hipLaunchKernelGGL(guard_clear_kernel, dim3(1), dim3(64), 0, st, g);
    o.guard = g;
  }
  return launch_status(exa::launch_predictor(c->k, st, ncells, d_luh, d_h, d_dt, d_lduh, d_lQhbnd, d_lFhbnd,
                                             d_picard, o),
                       "exa_batch_predictor_noalloc");
}

int exa_batch_riemann(exa_ctx* c, void* stream, long nfaces, const int* d_faces, const double* d_lQhbnd,
                      double* d_lFhbnd) {
  if (!c || nfaces < 0 || (nfaces > 0 && (!d_faces || !d_lQhbnd || !d_lFhbnd))) {
    set_error("exa_batch_riemann: invalid argument");
    return EXA_ERR_ARG;
  }
  return launch_status(exa::launch_riemann(c->k, static_cast<hipStream_t>(stream), nfaces, d_faces, d_lQhbnd,
                                           d_lFhbnd),
                       "exa_batch_riemann");
}

int exa_batch_corrector(exa_ctx* c, void* stream, long ncells, double* d_luh, double* d_lduh,
                        const double* d_lFhbnd, const double* d_h, const double* d_dt, double* d_dt_next) {
  if (!c || ncells < 0 || (ncells > 0 && (!d_luh || !d_lduh || !d_lFhbnd || !d_h || !d_dt || !d_dt_next))) {
    set_error("exa_batch_corrector: invalid argument");
    return EXA_ERR_ARG;
  }
  exa::LaunchOpts o;
  o.hdim = c->cell_dx;
  return launch_status(exa::launch_corrector(c->k, static_cast<hipStream_t>(stream), ncells, d_luh, d_lduh,
                                             d_lFhbnd, d_h, d_dt, d_dt_next, o),
                       "exa_batch_corrector");
}

int exa_batch_stable_dt(exa_ctx* c, void* stream, long ncells, const double* d_luh, const double* d_h,
                        double* d_dt_out) {
  if (!c || ncells < 0 || (ncells > 0 && (!d_luh || !d_h || !d_dt_out))) {
    set_error("exa_batch_stable_dt: invalid argument");
    return EXA_ERR_ARG;
  }
  exa::LaunchOpts o;
  o.hdim = c->cell_dx;
  return launch_status(exa::launch_stable_dt(c->k, static_cast<hipStream_t>(stream), ncells, d_luh, d_h,
                                             d_dt_out, o),
                       "exa_batch_stable_dt");
}

}  // extern "C"

// Accessors for the mesh driver (same library).
namespace exa {
const KernelCfg& ctx_cfg(const exa_ctx* c) { return c->k; }
}  // namespace exa

// Batched uneven work on the device:
//   exa_amr       static two-level mesh (3D): coarse n^3 cells on the unit cube,
//                 selected coarse cells split into 3^3 children; per-leaf cell
//                 size, one global dt (the reference's global time stepping with
//                 its coarse-fine coupling through virtual cells,
//                 ExaHyPE/exahype/solvers/ADERDGSolver.cpp:1340-1399 and 1472-1563)
//   exa_ensemble  many independent uniform periodic meshes of different sizes,
//                 one dt per instance, all cells in one array and every kernel
//                 one launch over all instances
//   exa_group     members stepped by one captured hipGraph per step, each
//                 member on its own branch of the graph
// and the autotuner that picks the launch geometry (persistent grid of the
// predictor and corrector, faces per Riemann workgroup) by timing the
// candidates on the member's own data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exahype_amd.h"
#include "engine.h"
#include "trace.h"

namespace exa {
const KernelCfg& ctx_cfg(const exa_ctx* c);
}

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

#define RC_OK(expr)         \
  do {                      \
    const int _rc = (expr); \
    if (_rc) return _rc;    \
  } while (0)

int status(int rc, const char* what) {
  if (rc == -1) {
    exa::set_error(std::string(what) + ": shape not compiled into this library");
    return EXA_ERR_UNSUPPORTED;
  }
  if (rc != 0) return fail(static_cast<hipError_t>(rc), what);
  return EXA_OK;
}

template <typename T>
int upload(T** dst, const std::vector<T>& v, const char* what) {
  if (v.empty()) return EXA_OK;
  HIP_OK(hipMalloc(dst, sizeof(T) * v.size()), what);
  HIP_OK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), what);
  return EXA_OK;
}

int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

// cells per workgroup of the node-column predictor / corrector (Shape::CPB)
long tiles_of(const exa::KernelCfg& k, long ncells) {
  const int NS = ipow(k.order + 1, k.dim);
  const int cpb = NS >= 128 ? 1 : 256 / NS;
  return (ncells + cpb - 1) / cpb;
}

int compute_units() {
  int dev = 0, cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) return 256;
  return cu;
}

// Launch geometry chosen per member (predictor grid, Riemann waves, corrector grid).
struct Tuning {
  exa::LaunchOpts pred, riem, corr;
  float ms[3] = {0, 0, 0};
  bool tuned = false;
};

// Average milliseconds of `reps` launches of f on stream s (one untimed warm-up).
template <typename F>
int time_launch(hipStream_t s, int reps, F f, float* ms) {
  hipEvent_t a = nullptr, b = nullptr;
  HIP_OK(hipEventCreate(&a), "event");
  HIP_OK(hipEventCreate(&b), "event");
  int rc = f();
  if (!rc && hipEventRecord(a, s) != hipSuccess) rc = fail(hipErrorUnknown, "event");
  for (int i = 0; !rc && i < reps; i++) rc = f();
  if (!rc && hipEventRecord(b, s) != hipSuccess) rc = fail(hipErrorUnknown, "event");
  if (!rc && hipEventSynchronize(b) != hipSuccess) rc = fail(hipErrorUnknown, "event sync");
  float t = 0;
  if (!rc) (void)hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t / reps;
  return rc;
}

// Members share this stepping interface.
struct Member {
  virtual ~Member() = default;
  virtual int enqueue_step(hipStream_t s) = 0;
  virtual hipStream_t own_stream() const = 0;
  virtual void invalidate_graph() = 0;
};

// The common state of a set of cells stepped with the batched kernels.
struct CellSet {
  exa::KernelCfg k{};
  long ncells = 0;   // leaves / instance cells
  long nslots = 0;   // cells with face storage (leaves + virtual cells)
  int NS = 0, NF = 0, ND = 0, V = 0, ninst = 1;
  hipStream_t stream = nullptr;
  double *d_luh = nullptr, *d_lduh = nullptr, *d_qb = nullptr, *d_fb = nullptr, *d_h = nullptr;
  double* d_state = nullptr;  // t[ninst], dt[ninst], dt_next[ninst]
  int* d_cinst = nullptr;
  int* d_picard = nullptr;  // per-cell Picard count of the last step
  int* d_faces = nullptr;
  long nfaces = 0;
  int* d_prolong = nullptr;
  long nprolong = 0;
  int* d_restrict = nullptr;
  long nrestrict = 0;
  Tuning tune;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  void release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    double* b[] = {d_luh, d_lduh, d_qb, d_fb, d_h, d_state};
    for (double* p : b)
      if (p) (void)hipFree(p);
    int* ib[] = {d_cinst, d_picard, d_faces, d_prolong, d_restrict};
    for (int* p : ib)
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  int allocate() {
    const int D = k.dim, B = k.order + 1;
    NS = ipow(B, D);
    NF = ipow(B, D - 1);
    V = k.nvar;
    ND = k.nvar + k.npar;
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
    HIP_OK(hipMalloc(&d_luh, sizeof(double) * ncells * NS * ND), "hipMalloc(luh)");
    HIP_OK(hipMalloc(&d_lduh, sizeof(double) * ncells * NS * V), "hipMalloc(lduh)");
    HIP_OK(hipMalloc(&d_qb, sizeof(double) * nslots * 2 * D * NF * ND), "hipMalloc(lQhbnd)");
    HIP_OK(hipMalloc(&d_fb, sizeof(double) * nslots * 2 * D * NF * V), "hipMalloc(lFhbnd)");
    HIP_OK(hipMalloc(&d_state, sizeof(double) * 3 * ninst), "hipMalloc(state)");
    HIP_OK(hipMalloc(&d_picard, sizeof(int) * ncells), "hipMalloc(picard)");
    HIP_OK(hipMemset(d_picard, 0, sizeof(int) * ncells), "clear");
    std::vector<double> st(3 * ninst, 0.0);
    for (int i = ninst; i < 3 * ninst; i++) st[i] = DBL_MAX;
    HIP_OK(hipMemcpy(d_state, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice), "state");
    HIP_OK(hipMemset(d_qb, 0, sizeof(double) * nslots * 2 * D * NF * ND), "clear");
    HIP_OK(hipMemset(d_fb, 0, sizeof(double) * nslots * 2 * D * NF * V), "clear");
    return EXA_OK;
  }

  exa::LaunchOpts with_cinst(exa::LaunchOpts o) const {
    o.cinst = d_cinst;
    return o;
  }

  int enqueue(hipStream_t s) {
    exa::TraceRange tr("exa.batch.step");
    double* dt = d_state + ninst;
    double* dtn = d_state + 2 * ninst;
    RC_OK(status(exa::launch_predictor(k, s, ncells, d_luh, d_h, dt, d_lduh, d_qb, d_fb, d_picard,
                                       with_cinst(tune.pred)),
                 "predictor"));
    if (nprolong) RC_OK(status(exa::launch_face_prolong(k, s, nprolong, d_prolong, d_qb, d_fb), "face prolongation"));
    RC_OK(status(exa::launch_riemann(k, s, nfaces, d_faces, d_qb, d_fb, tune.riem), "riemann"));
    if (nrestrict) RC_OK(status(exa::launch_face_restrict(k, s, nrestrict, d_restrict, d_fb), "face restriction"));
    RC_OK(status(exa::launch_corrector(k, s, ncells, d_luh, d_lduh, d_fb, d_h, dt, dtn, with_cinst(tune.corr)),
                 "corrector"));
    return status(exa::launch_advance_time(s, d_state, ninst), "advance time");
  }

  int init_dt(double dt) {
    std::vector<double> st(3 * ninst, 0.0);
    HIP_OK(hipMemcpy(st.data(), d_state, sizeof(double) * st.size(), hipMemcpyDeviceToHost), "state");
    for (int i = 0; i < ninst; i++) {
      st[ninst + i] = dt > 0 ? dt : DBL_MAX;
      st[2 * ninst + i] = DBL_MAX;
    }
    HIP_OK(hipMemcpy(d_state, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice), "state");
    if (!(dt > 0)) {
      exa::LaunchOpts o;
      o.cinst = d_cinst;
      RC_OK(status(exa::launch_stable_dt(k, stream, ncells, d_luh, d_h, d_state + ninst, o), "initial dt"));
    }
    HIP_OK(hipStreamSynchronize(stream), "sync");
    return EXA_OK;
  }

  int run(int nsteps, int use_graph) {
    if (use_graph) {
      if (!graph_exec) {
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        const int rc = enqueue(stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) return fail(e, "end capture");
        graph = g;
        HIP_OK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "instantiate");
      }
      for (int i = 0; i < nsteps; i++) HIP_OK(hipGraphLaunch(graph_exec, stream), "graph launch");
    } else {
      for (int i = 0; i < nsteps; i++) RC_OK(enqueue(stream));
    }
    return EXA_OK;
  }

  // Pick the launch geometry by timing candidates on this set's data.
  int autotune(int* cfg, double* ms) {
    const long tiles = tiles_of(k, ncells);
    const int cu = compute_units();
    std::vector<long> grids = {0};
    for (int m : {1, 2, 4, 8, 16})
      if (long(cu) * m < tiles) grids.push_back(long(cu) * m);
    double* dt = d_state + ninst;
    double* dtn = d_state + 2 * ninst;
    // the corrector updates luh and dt_next: keep copies to restore
    double* save_luh = nullptr;
    double* save_state = nullptr;
    HIP_OK(hipMalloc(&save_luh, sizeof(double) * ncells * NS * ND), "hipMalloc(autotune)");
    hipError_t e = hipMalloc(&save_state, sizeof(double) * 3 * ninst);
    if (e == hipSuccess)
      e = hipMemcpyAsync(save_luh, d_luh, sizeof(double) * ncells * NS * ND, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(save_state, d_state, sizeof(double) * 3 * ninst, hipMemcpyDeviceToDevice, stream);
    int rc = e == hipSuccess ? EXA_OK : fail(e, "autotune save");
    Tuning best = tune;
    float bp = 1e30f, br = 1e30f, bc = 1e30f;
    for (long g : grids) {
      if (rc) break;
      exa::LaunchOpts o = with_cinst(exa::LaunchOpts());
      o.grid = g;
      float t = 0;
      rc = time_launch(stream, 3, [&]() {
        return status(exa::launch_predictor(k, stream, ncells, d_luh, d_h, dt, d_lduh, d_qb, d_fb, nullptr, o),
                      "predictor");
      }, &t);
      if (!rc && t < bp) {
        bp = t;
        best.pred.grid = g;
      }
    }
    if (!rc && nprolong) rc = status(exa::launch_face_prolong(k, stream, nprolong, d_prolong, d_qb, d_fb), "prolong");
    for (int w : {1, 2, 4}) {  // the Riemann kernel's launch bound is 4 waves
      if (rc) break;
      exa::LaunchOpts o;
      o.riemann_waves = w;
      float t = 0;
      rc = time_launch(stream, 3, [&]() {
        return status(exa::launch_riemann(k, stream, nfaces, d_faces, d_qb, d_fb, o), "riemann");
      }, &t);
      if (!rc && t < br) {
        br = t;
        best.riem.riemann_waves = w;
      }
    }
    for (long g : grids) {
      if (rc) break;
      exa::LaunchOpts o = with_cinst(exa::LaunchOpts());
      o.grid = g;
      float t = 0;
      rc = time_launch(stream, 3, [&]() {
        return status(exa::launch_corrector(k, stream, ncells, d_luh, d_lduh, d_fb, d_h, dt, dtn, o), "corrector");
      }, &t);
      if (!rc && t < bc) {
        bc = t;
        best.corr.grid = g;
      }
    }
    if (!rc) {
      e = hipMemcpyAsync(d_luh, save_luh, sizeof(double) * ncells * NS * ND, hipMemcpyDeviceToDevice, stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(d_state, save_state, sizeof(double) * 3 * ninst, hipMemcpyDeviceToDevice, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) rc = fail(e, "autotune restore");
    }
    (void)hipFree(save_luh);
    (void)hipFree(save_state);
    if (rc) return rc;
    best.ms[0] = bp;
    best.ms[1] = br;
    best.ms[2] = bc;
    best.tuned = true;
    tune = best;
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    if (cfg) {
      cfg[0] = static_cast<int>(tune.pred.grid);
      cfg[1] = tune.riem.riemann_waves;
      cfg[2] = static_cast<int>(tune.corr.grid);
    }
    if (ms)
      for (int i = 0; i < 3; i++) ms[i] = tune.ms[i];
    return EXA_OK;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// static two-level mesh
// ---------------------------------------------------------------------------
struct exa_amr : Member {
  CellSet cs;
  int n = 0;
  std::vector<int> level, coords;  // per leaf: 0 coarse / 1 fine; integer coordinates on its level
  int enqueue_step(hipStream_t s) override { return cs.enqueue(s); }
  hipStream_t own_stream() const override { return cs.stream; }
  void invalidate_graph() override {
    if (cs.graph_exec) (void)hipGraphExecDestroy(cs.graph_exec);
    if (cs.graph) (void)hipGraphDestroy(cs.graph);
    cs.graph_exec = nullptr;
    cs.graph = nullptr;
  }
};

struct exa_ensemble : Member {
  CellSet cs;
  std::vector<int> n;        // cells per side per instance
  std::vector<long> offset;  // first cell of each instance
  int enqueue_step(hipStream_t s) override { return cs.enqueue(s); }
  hipStream_t own_stream() const override { return cs.stream; }
  void invalidate_graph() override {
    if (cs.graph_exec) (void)hipGraphExecDestroy(cs.graph_exec);
    if (cs.graph) (void)hipGraphDestroy(cs.graph);
    cs.graph_exec = nullptr;
    cs.graph = nullptr;
  }
};

struct exa_group {
  std::vector<Member*> members;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> ev;  // fork + one join per member
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};

extern "C" {

// ---- two-level mesh ---------------------------------------------------------
int exa_amr_create(exa_amr** out, exa_ctx* ctx, int n, const unsigned char* refined) {
  if (!out || !ctx || n < 1 || !refined) return argerr("exa_amr_create: invalid argument");
  *out = nullptr;
  const exa::KernelCfg& k = exa::ctx_cfg(ctx);
  if (k.dim != 3 || k.pde.id == EXA_PDE_ACOUSTIC) {
    exa::set_error("exa_amr_create: the two-level mesh is compiled for the 3D nonlinear path");
    return EXA_ERR_UNSUPPORTED;
  }
  exa_amr* a = new exa_amr;
  a->n = n;
  CellSet& cs = a->cs;
  cs.k = k;
  const long nc = long(n) * n * n, N3 = 3L * n;
  std::vector<int> coarse_leaf(nc, -1), fine_leaf(N3 * N3 * N3, -1);
  for (long c = 0; c < nc; c++) {
    const int ix = int(c % n), iy = int((c / n) % n), iz = int(c / (long(n) * n));
    if (!refined[c]) {
      coarse_leaf[c] = static_cast<int>(a->level.size());
      a->level.push_back(0);
      a->coords.insert(a->coords.end(), {ix, iy, iz});
    } else {
      for (int sz = 0; sz < 3; sz++)
        for (int sy = 0; sy < 3; sy++)
          for (int sx = 0; sx < 3; sx++) {
            const int X = 3 * ix + sx, Y = 3 * iy + sy, Z = 3 * iz + sz;
            fine_leaf[(long(Z) * N3 + Y) * N3 + X] = static_cast<int>(a->level.size());
            a->level.push_back(1);
            a->coords.insert(a->coords.end(), {X, Y, Z});
          }
    }
  }
  const long L = static_cast<long>(a->level.size());
  std::vector<int> faces, prolong, restr;
  long nvirt = 0;
  for (long i = 0; i < L; i++) {
    const int* c0 = &a->coords[3 * i];
    for (int d = 0; d < 3; d++) {
      int c[3] = {c0[0], c0[1], c0[2]};
      int j;
      if (a->level[i] == 0) {
        c[d] = (c[d] + 1) % n;
        j = coarse_leaf[(long(c[2]) * n + c[1]) * n + c[0]];
      } else {
        c[d] = int((c[d] + 1) % N3);
        j = fine_leaf[(long(c[2]) * N3 + c[1]) * N3 + c[0]];
      }
      if (j >= 0) faces.insert(faces.end(), {int(i), j, d, 0});
    }
  }
  for (long i = 0; i < L; i++) {
    if (a->level[i] != 0) continue;
    const int* c0 = &a->coords[3 * i];
    for (int f = 0; f < 6; f++) {
      const int d = f / 2, o = f % 2;
      int c[3] = {c0[0], c0[1], c0[2]};
      c[d] = (c[d] + (o ? 1 : n - 1)) % n;
      if (!refined[(long(c[2]) * n + c[1]) * n + c[0]]) continue;
      const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
      const long v0 = L + nvirt;
      restr.insert(restr.end(), {int(i), f, int(v0), 0});
      for (int s1 = 0; s1 < 3; s1++)
        for (int s0 = 0; s0 < 3; s0++) {
          const long v = L + nvirt++;
          prolong.insert(prolong.end(), {int(i), f, s0 + 3 * s1, int(v)});
          int X[3];
          X[d] = o ? 3 * c[d] : 3 * c[d] + 2;
          X[t0] = 3 * c0[t0] + s0;
          X[t1] = 3 * c0[t1] + s1;
          const int j = fine_leaf[(long(X[2]) * N3 + X[1]) * N3 + X[0]];
          if (o)
            faces.insert(faces.end(), {int(v), j, d, 0});
          else
            faces.insert(faces.end(), {j, int(v), d, 0});
        }
    }
  }
  cs.ncells = L;
  cs.nslots = L + nvirt;
  cs.nfaces = static_cast<long>(faces.size() / 4);
  cs.nprolong = static_cast<long>(prolong.size() / 4);
  cs.nrestrict = static_cast<long>(restr.size() / 4);
  std::vector<double> h(L);
  for (long i = 0; i < L; i++) h[i] = a->level[i] ? 1.0 / (3.0 * n) : 1.0 / n;
  int rc = cs.allocate();
  if (!rc) rc = upload(&cs.d_h, h, "upload h");
  if (!rc) rc = upload(&cs.d_faces, faces, "upload faces");
  if (!rc) rc = upload(&cs.d_prolong, prolong, "upload prolongation list");
  if (!rc) rc = upload(&cs.d_restrict, restr, "upload restriction list");
  if (rc) {
    cs.release();
    delete a;
    return rc;
  }
  *out = a;
  return EXA_OK;
}

int exa_amr_destroy(exa_amr* a) {
  if (!a) return EXA_OK;
  if (a->cs.stream) (void)hipStreamSynchronize(a->cs.stream);
  a->cs.release();
  delete a;
  return EXA_OK;
}

long exa_amr_nleaves(const exa_amr* a) { return a ? a->cs.ncells : 0; }

int exa_amr_leaves(exa_amr* a, int* level, int* coords) {
  if (!a) return argerr("exa_amr_leaves: NULL mesh");
  if (level) std::copy(a->level.begin(), a->level.end(), level);
  if (coords) std::copy(a->coords.begin(), a->coords.end(), coords);
  return EXA_OK;
}

int exa_amr_set_solution(exa_amr* a, const double* luh) {
  if (!a || !luh) return argerr("exa_amr_set_solution: NULL argument");
  HIP_OK(hipMemcpy(a->cs.d_luh, luh, sizeof(double) * a->cs.ncells * a->cs.NS * a->cs.ND, hipMemcpyHostToDevice),
         "set_solution");
  return EXA_OK;
}

int exa_amr_get_solution(exa_amr* a, double* luh) {
  if (!a || !luh) return argerr("exa_amr_get_solution: NULL argument");
  HIP_OK(hipStreamSynchronize(a->cs.stream), "sync");
  HIP_OK(hipMemcpy(luh, a->cs.d_luh, sizeof(double) * a->cs.ncells * a->cs.NS * a->cs.ND, hipMemcpyDeviceToHost),
         "get_solution");
  return EXA_OK;
}

int exa_amr_init_dt(exa_amr* a, double dt) {
  if (!a) return argerr("exa_amr_init_dt: NULL mesh");
  return a->cs.init_dt(dt);
}

int exa_amr_run(exa_amr* a, int nsteps, int use_graph) {
  if (!a || nsteps < 0) return argerr("exa_amr_run: invalid argument");
  return a->cs.run(nsteps, use_graph);
}

int exa_amr_sync(exa_amr* a) {
  if (!a) return argerr("exa_amr_sync: NULL mesh");
  HIP_OK(hipStreamSynchronize(a->cs.stream), "sync");
  return EXA_OK;
}

int exa_amr_time(exa_amr* a, double* t, double* dt) {
  if (!a) return argerr("exa_amr_time: NULL mesh");
  double st[3];
  HIP_OK(hipStreamSynchronize(a->cs.stream), "sync");
  HIP_OK(hipMemcpy(st, a->cs.d_state, sizeof(st), hipMemcpyDeviceToHost), "time");
  if (t) *t = st[0];
  if (dt) *dt = st[1];
  return EXA_OK;
}

int exa_amr_picard(exa_amr* a, int* counts) {
  if (!a || !counts) return argerr("exa_amr_picard: NULL argument");
  HIP_OK(hipStreamSynchronize(a->cs.stream), "sync");
  HIP_OK(hipMemcpy(counts, a->cs.d_picard, sizeof(int) * a->cs.ncells, hipMemcpyDeviceToHost), "picard");
  return EXA_OK;
}

int exa_amr_autotune(exa_amr* a, int* cfg, double* ms) {
  if (!a) return argerr("exa_amr_autotune: NULL mesh");
  return a->cs.autotune(cfg, ms);
}

// ---- ensemble of independent uniform meshes --------------------------------
int exa_ensemble_create(exa_ensemble** out, exa_ctx* ctx, int ninst, const int* n) {
  if (!out || !ctx || ninst < 1 || !n) return argerr("exa_ensemble_create: invalid argument");
  *out = nullptr;
  const exa::KernelCfg& k = exa::ctx_cfg(ctx);
  if (k.pde.id == EXA_PDE_ACOUSTIC) {
    exa::set_error("exa_ensemble_create: the ensemble is compiled for the nonlinear path");
    return EXA_ERR_UNSUPPORTED;
  }
  const int D = k.dim;
  exa_ensemble* e = new exa_ensemble;
  CellSet& cs = e->cs;
  cs.k = k;
  cs.ninst = ninst;
  long total = 0;
  for (int i = 0; i < ninst; i++) {
    if (n[i] < 1) {
      delete e;
      return argerr("exa_ensemble_create: every instance needs at least one cell per side");
    }
    e->n.push_back(n[i]);
    e->offset.push_back(total);
    total += (D == 3) ? long(n[i]) * n[i] * n[i] : long(n[i]) * n[i];
  }
  e->offset.push_back(total);
  std::vector<int> cinst(total), faces;
  std::vector<double> h(total);
  for (int i = 0; i < ninst; i++) {
    const long off = e->offset[i], cnt = e->offset[i + 1] - off;
    const int m = n[i];
    for (long c = 0; c < cnt; c++) {
      cinst[off + c] = i;
      h[off + c] = 1.0 / m;
    }
    for (int d = 0; d < D; d++)
      for (long c = 0; c < cnt; c++) {
        long ix = c % m, iy = (c / m) % m, iz = (D == 3) ? c / (long(m) * m) : 0;
        if (d == 0) ix = (ix + 1) % m;
        if (d == 1) iy = (iy + 1) % m;
        if (d == 2) iz = (iz + 1) % m;
        const long r = (D == 3) ? (iz * m + iy) * m + ix : iy * m + ix;
        faces.insert(faces.end(), {int(off + c), int(off + r), d, 0});
      }
  }
  cs.ncells = total;
  cs.nslots = total;
  cs.nfaces = static_cast<long>(faces.size() / 4);
  int rc = cs.allocate();
  if (!rc) rc = upload(&cs.d_h, h, "upload h");
  if (!rc) rc = upload(&cs.d_cinst, cinst, "upload instance index");
  if (!rc) rc = upload(&cs.d_faces, faces, "upload faces");
  if (rc) {
    cs.release();
    delete e;
    return rc;
  }
  *out = e;
  return EXA_OK;
}

int exa_ensemble_destroy(exa_ensemble* e) {
  if (!e) return EXA_OK;
  if (e->cs.stream) (void)hipStreamSynchronize(e->cs.stream);
  e->cs.release();
  delete e;
  return EXA_OK;
}

long exa_ensemble_ncells(const exa_ensemble* e, int inst) {
  if (!e) return 0;
  if (inst < 0) return e->cs.ncells;
  if (inst >= int(e->n.size())) return 0;
  return e->offset[inst + 1] - e->offset[inst];
}

int exa_ensemble_set_solution(exa_ensemble* e, int inst, const double* luh) {
  if (!e || !luh || inst < 0 || inst >= int(e->n.size())) return argerr("exa_ensemble_set_solution: invalid argument");
  const size_t per = size_t(e->cs.NS) * e->cs.ND;
  HIP_OK(hipMemcpy(e->cs.d_luh + per * e->offset[inst], luh, sizeof(double) * per * (e->offset[inst + 1] - e->offset[inst]),
                   hipMemcpyHostToDevice),
         "set_solution");
  return EXA_OK;
}

int exa_ensemble_get_solution(exa_ensemble* e, int inst, double* luh) {
  if (!e || !luh || inst < 0 || inst >= int(e->n.size())) return argerr("exa_ensemble_get_solution: invalid argument");
  const size_t per = size_t(e->cs.NS) * e->cs.ND;
  HIP_OK(hipStreamSynchronize(e->cs.stream), "sync");
  HIP_OK(hipMemcpy(luh, e->cs.d_luh + per * e->offset[inst], sizeof(double) * per * (e->offset[inst + 1] - e->offset[inst]),
                   hipMemcpyDeviceToHost),
         "get_solution");
  return EXA_OK;
}

int exa_ensemble_init_dt(exa_ensemble* e, double dt) {
  if (!e) return argerr("exa_ensemble_init_dt: NULL ensemble");
  return e->cs.init_dt(dt);
}

int exa_ensemble_run(exa_ensemble* e, int nsteps, int use_graph) {
  if (!e || nsteps < 0) return argerr("exa_ensemble_run: invalid argument");
  return e->cs.run(nsteps, use_graph);
}

int exa_ensemble_sync(exa_ensemble* e) {
  if (!e) return argerr("exa_ensemble_sync: NULL ensemble");
  HIP_OK(hipStreamSynchronize(e->cs.stream), "sync");
  return EXA_OK;
}

int exa_ensemble_time(exa_ensemble* e, int inst, double* t, double* dt) {
  if (!e || inst < 0 || inst >= int(e->n.size())) return argerr("exa_ensemble_time: invalid argument");
  const int m = e->cs.ninst;
  std::vector<double> st(3 * m);
  HIP_OK(hipStreamSynchronize(e->cs.stream), "sync");
  HIP_OK(hipMemcpy(st.data(), e->cs.d_state, sizeof(double) * st.size(), hipMemcpyDeviceToHost), "time");
  if (t) *t = st[inst];
  if (dt) *dt = st[m + inst];
  return EXA_OK;
}

int exa_ensemble_picard(exa_ensemble* e, int* counts) {
  if (!e || !counts) return argerr("exa_ensemble_picard: NULL argument");
  HIP_OK(hipStreamSynchronize(e->cs.stream), "sync");
  HIP_OK(hipMemcpy(counts, e->cs.d_picard, sizeof(int) * e->cs.ncells, hipMemcpyDeviceToHost), "picard");
  return EXA_OK;
}

int exa_ensemble_autotune(exa_ensemble* e, int* cfg, double* ms) {
  if (!e) return argerr("exa_ensemble_autotune: NULL ensemble");
  return e->cs.autotune(cfg, ms);
}

// ---- group: one captured graph per step ------------------------------------
int exa_group_create(exa_group** out, int namr, exa_amr* const* amrs, int nens, exa_ensemble* const* ens) {
  if (!out || namr < 0 || nens < 0 || namr + nens < 1 || (namr && !amrs) || (nens && !ens))
    return argerr("exa_group_create: invalid argument");
  *out = nullptr;
  exa_group* g = new exa_group;
  for (int i = 0; i < namr; i++) g->members.push_back(amrs[i]);
  for (int i = 0; i < nens; i++) g->members.push_back(ens[i]);
  for (Member* m : g->members)
    if (!m) {
      delete g;
      return argerr("exa_group_create: NULL member");
    }
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  g->ev.assign(1 + g->members.size(), nullptr);
  for (auto& x : g->ev)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
  if (e != hipSuccess) {
    for (auto& x : g->ev)
      if (x) (void)hipEventDestroy(x);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return fail(e, "exa_group_create");
  }
  *out = g;
  return EXA_OK;
}

int exa_group_destroy(exa_group* g) {
  if (!g) return EXA_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->graph_exec) (void)hipGraphExecDestroy(g->graph_exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (auto& x : g->ev)
    if (x) (void)hipEventDestroy(x);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return EXA_OK;
}

// Enqueue one step of every member: fork from the group stream to each
// member's stream, join back.  Captured, this is one graph whose members are
// parallel branches.
static int group_enqueue(exa_group* g) {
  HIP_OK(hipEventRecord(g->ev[0], g->stream), "fork");
  for (size_t i = 0; i < g->members.size(); i++) {
    Member* m = g->members[i];
    HIP_OK(hipStreamWaitEvent(m->own_stream(), g->ev[0], 0), "fork wait");
    RC_OK(m->enqueue_step(m->own_stream()));
    HIP_OK(hipEventRecord(g->ev[1 + i], m->own_stream()), "join");
    HIP_OK(hipStreamWaitEvent(g->stream, g->ev[1 + i], 0), "join wait");
  }
  return EXA_OK;
}

int exa_group_run(exa_group* g, int nsteps, int use_graph) {
  if (!g || nsteps < 0) return argerr("exa_group_run: invalid argument");
  // members may have queued work of their own
  for (Member* m : g->members) HIP_OK(hipStreamSynchronize(m->own_stream()), "sync member");
  if (use_graph) {
    if (!g->graph_exec) {
      HIP_OK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      const int rc = group_enqueue(g);
      hipGraph_t gr = nullptr;
      const hipError_t e = hipStreamEndCapture(g->stream, &gr);
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (e != hipSuccess) return fail(e, "end capture");
      g->graph = gr;
      HIP_OK(hipGraphInstantiate(&g->graph_exec, g->graph, nullptr, nullptr, 0), "instantiate");
    }
    for (int i = 0; i < nsteps; i++) HIP_OK(hipGraphLaunch(g->graph_exec, g->stream), "graph launch");
  } else {
    for (int i = 0; i < nsteps; i++) RC_OK(group_enqueue(g));
  }
  HIP_OK(hipStreamSynchronize(g->stream), "sync");
  return EXA_OK;
}

// Drop the captured graph (after a member was re-tuned).
int exa_group_reset(exa_group* g) {
  if (!g) return argerr("exa_group_reset: NULL group");
  if (g->graph_exec) (void)hipGraphExecDestroy(g->graph_exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->graph_exec = nullptr;
  g->graph = nullptr;
  return EXA_OK;
}

}  // extern "C"

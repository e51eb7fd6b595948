// Host mesh driver: the part of the reference's Peano traversal + solver
// orchestration that the hot path needs, re-implemented for device-resident
// patch arrays.  One global time step (non-fused global scheme of
// ExaHyPE/exahype/solvers/ADERDGSolver.cpp):
//   1. predictor + volume integral for all cells   (predictionAndVolumeIntegralBody, :887-947)
//   2. Riemann solve on all faces                   (solveRiemannProblemAtInterface, :1764-1823)
//   3. face integrals into the update, update       (surfaceIntegral :1190-1239, addUpdateToSolution :1241-1256)
//   4. dt_{n+1} = min over cells stableTimeStepSize (startNewTimeStep :1031-1049, reduce :458-462)
// The first dt comes from the initial solution (updateTimeStepSize, :1051-1060).
// Periodic neighbours are this driver's own rule: the reference ADER-DG path
// has no periodic boundary.  Non-periodic domain faces take the reference's
// boundary path (applyBoundaryConditions, :1836-1900): after the predictor a
// boundary kernel writes the outside face data into ghost faces, which the
// Riemann kernels read in place of a neighbour.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/exahype_amd.h"
#include "engine.h"
#include "trace.h"

namespace exa {
const KernelCfg& ctx_cfg(const exa_ctx* c);
}

struct exa_mesh {
  exa_ctx* ctx = nullptr;
  exa::KernelCfg k{};
  int n = 0;
  long ncells = 0, nfaces = 0;
  int NS = 0, NF = 0, ND = 0;
  double h = 0;
  hipStream_t stream = nullptr;
  double *d_luh = nullptr, *d_lduh = nullptr, *d_qb = nullptr, *d_fb = nullptr, *d_h = nullptr;
  double* d_state = nullptr;  // {t, dt, dt_next}
  int* d_faces = nullptr;
  int* d_picard = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool timing = false;
  std::vector<hipEvent_t> ev;  // 4 events per timed step: before/after each phase
  int timed_steps = 0;
  // fused step: the predictor adds the volume update to luh and one kernel
  // solves each cell's face Riemann problems and applies the face update
  bool fused = false;
  // boundary kinds per domain face; ghost faces [2D][n^(D-1)][NF][ND | V]
  exa::BcKinds bc{};
  unsigned bcmask = 0;
  bool hooks = false;  // some face takes its ghost data from a boundary-value hook
  double *d_gq = nullptr, *d_gf = nullptr;
  double* d_halo = nullptr;  // staging of the host-side halo entry points
  // per-workgroup stable-dt minima spread over kDtSlots cache lines (DBL_MAX between steps)
  double* d_slots = nullptr;
  // Picard guard of the one-iteration FP32 CERK predictor {listed, listed in total, list[ncells]}
  int* d_guard = nullptr;
  // slab-pipelined fused step (exa_mesh_set_overlap): `slabs` slabs of z planes;
  // the predictors run in slab order on `stream`, the corrector of a slab on
  // stream2 as soon as the predictors of its two neighbour slabs are done.
  // sev: slab predictor done [slabs], fork, join (no timing).
  int slabs = 0;
  hipStream_t stream2 = nullptr;
  std::vector<hipEvent_t> sev;
};

namespace {

int fail(hipError_t e, const char* what) {
  exa::set_error(std::string(what) + ": " + hipGetErrorString(e));
  return EXA_ERR_HIP;
}

#define HIP_OK(expr, what)                          \
  do {                                              \
    hipError_t _e = (expr);                         \
    if (_e != hipSuccess) return fail(_e, what);    \
  } while (0)

int status(int rc, const char* what) {
  if (rc == -1) {
    exa::set_error(std::string(what) + ": shape not compiled into this library");
    return EXA_ERR_UNSUPPORTED;
  }
  if (rc != 0) return fail(static_cast<hipError_t>(rc), what);
  return EXA_OK;
}

int enqueue_boundary(exa_mesh* m) {
  if (!m->hooks) return EXA_OK;
  return status(exa::launch_boundary_mesh(m->k, m->stream, m->n, m->h, m->bc, m->d_qb, m->d_gq, m->d_gf, m->d_state),
                "boundary conditions");
}

// FP32 extrapolated faces (half the face traffic, exact for the FP32 forms):
// fused step, all faces periodic, and an FP32 predictor that can write them.
bool f32_faces(const exa_mesh* m) {
  return m->fused && !m->bcmask && exa::predictor_f32_faces_supported(m->k);
}

exa::LaunchOpts ghost_opts(const exa_mesh* m) {
  exa::LaunchOpts o;
  if (m->bcmask) {
    o.ghost_q = m->d_gq;
    o.ghost_f = m->d_gf;
    o.bcmask = m->bcmask;
  }
  o.faces_f32 = f32_faces(m);
  o.dt_slots = m->d_slots;
  return o;
}

// One global time step on m->stream in three parts, so that a caller can
// exchange halo faces between predict and correct and reduce dt_next across
// processes between correct and advance.  Events (timing): ev[0] before the
// predictor, ev[1] after it, ev[2] after the Riemann solves (four-kernel form)
// or before the fused corrector, ev[3] after the step.
int enqueue_predict(exa_mesh* m, hipEvent_t* ev) {
  exa::TraceRange tr("exa.mesh.predictor");
  double* st = m->d_state;
  if (ev) HIP_OK(hipEventRecord(ev[0], m->stream), "event");
  exa::LaunchOpts o;
  if (m->fused) o.luh_update = m->d_luh;
  o.faces_f32 = f32_faces(m);
  o.guard = m->d_guard;
  int rc = status(exa::launch_predictor(m->k, m->stream, m->ncells, m->d_luh, m->d_h, st + 1, m->d_lduh, m->d_qb,
                                        m->d_fb, m->d_picard, o),
                  "predictor");
  if (rc) return rc;
  if (ev) HIP_OK(hipEventRecord(ev[1], m->stream), "event");
  return enqueue_boundary(m);
}

int enqueue_correct(exa_mesh* m, hipEvent_t* ev) {
  exa::TraceRange tr("exa.mesh.corrector");
  double* st = m->d_state;
  int rc;
  if (m->fused) {
    if (ev) HIP_OK(hipEventRecord(ev[2], m->stream), "event");
    rc = status(exa::launch_corrector_riemann(m->k, m->stream, m->ncells, m->n, m->d_luh, m->d_qb, m->d_fb, m->d_h,
                                              st + 1, st + 2, ghost_opts(m)),
                "corrector");
  } else {
    rc = status(exa::launch_riemann(m->k, m->stream, m->nfaces, m->d_faces, m->d_qb, m->d_fb, ghost_opts(m)),
                "riemann");
    if (rc) return rc;
    if (ev) HIP_OK(hipEventRecord(ev[2], m->stream), "event");
    exa::LaunchOpts o;
    o.dt_slots = m->d_slots;
    rc = status(exa::launch_corrector(m->k, m->stream, m->ncells, m->d_luh, m->d_lduh, m->d_fb, m->d_h, st + 1,
                                      st + 2, o),
                "corrector");
  }
  if (rc) return rc;
  return status(exa::launch_dt_reduce(m->stream, m->d_slots, st + 2), "dt reduction");
}

int enqueue_advance(exa_mesh* m, hipEvent_t* ev) {
  const int rc = status(exa::launch_advance_time(m->stream, m->d_state), "advance time");
  if (rc) return rc;
  if (ev) HIP_OK(hipEventRecord(ev[3], m->stream), "event");
  return EXA_OK;
}

// The slab form applies: 3D, fused, all domain faces periodic, the predictor
// covers parts of the mesh, and at least one z plane per slab.
bool slabs_apply(const exa_mesh* m, int slabs) {
  return slabs >= 3 && m->k.dim == 3 && m->fused && m->bcmask == 0 && slabs <= m->n &&
         exa::predictor_supports_range(m->k);
}

// One fused step in slabs of z planes, slab k = planes [k n / S, (k + 1) n / S).
// The corrector of slab k reads the faces of slabs k - 1, k, k + 1 (periodic),
// so it waits for the predictor of slab k + 1 (the predictors run in order on
// one stream, their guard recomputations included); correctors 1 .. S - 2
// overlap the predictors of the later slabs, S - 1 and 0 run last.  A predictor
// writes only its own cells' luh and faces, a corrector only its own cells'
// luh, so the overlapped kernels touch disjoint data.  Same kernels, same
// operations per cell: the results equal the one-launch step's bit for bit.
int enqueue_step_slabs(exa_mesh* m) {
  exa::TraceRange tr("exa.mesh.step(slabs)");
  const int S = m->slabs;
  const long plane = long(m->n) * m->n;
  auto first = [&](int k) { return plane * ((long(k) * m->n) / S); };
  double* st = m->d_state;
  hipEvent_t fork = m->sev[S], join = m->sev[S + 1];
  HIP_OK(hipEventRecord(fork, m->stream), "event");
  HIP_OK(hipStreamWaitEvent(m->stream2, fork, 0), "stream wait");
  auto predict = [&](int k) -> int {
    exa::LaunchOpts o;
    o.luh_update = m->d_luh;
    o.faces_f32 = f32_faces(m);
    o.guard = m->d_guard;
    o.cell0 = first(k);
    o.cell_end = first(k + 1);
    const int rc = status(exa::launch_predictor(m->k, m->stream, m->ncells, m->d_luh, m->d_h, st + 1, m->d_lduh,
                                                m->d_qb, m->d_fb, m->d_picard, o),
                          "predictor");
    if (rc) return rc;
    HIP_OK(hipEventRecord(m->sev[k], m->stream), "event");
    return EXA_OK;
  };
  auto correct = [&](int k, int after) -> int {
    HIP_OK(hipStreamWaitEvent(m->stream2, m->sev[after], 0), "stream wait");
    exa::LaunchOpts o = ghost_opts(m);
    o.cell0 = first(k);
    o.cell_end = first(k + 1);
    return status(exa::launch_corrector_riemann(m->k, m->stream2, m->ncells, m->n, m->d_luh, m->d_qb, m->d_fb,
                                                m->d_h, st + 1, st + 2, o),
                  "corrector");
  };
  int rc = EXA_OK;
  for (int k = 0; k < S && !rc; k++) {
    rc = predict(k);
    if (!rc && k >= 2) rc = correct(k - 1, k);
  }
  if (!rc) rc = correct(S - 1, S - 1);
  if (!rc) rc = correct(0, S - 1);
  if (!rc) rc = status(exa::launch_dt_reduce(m->stream2, m->d_slots, st + 2), "dt reduction");
  // join (also after an error, so that a capture does not end with stream2 forked)
  HIP_OK(hipEventRecord(join, m->stream2), "event");
  HIP_OK(hipStreamWaitEvent(m->stream, join, 0), "stream wait");
  if (rc) return rc;
  return enqueue_advance(m, nullptr);
}

int enqueue_step(exa_mesh* m, hipEvent_t* ev) {
  // timed runs measure each phase in isolation: the one-launch step
  if (!ev && slabs_apply(m, m->slabs)) return enqueue_step_slabs(m);
  int rc = enqueue_predict(m, ev);
  if (!rc) rc = enqueue_correct(m, ev);
  if (!rc) rc = enqueue_advance(m, ev);
  return rc;
}

void free_mesh(exa_mesh* m) {
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (hipEvent_t e : m->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : m->sev) (void)hipEventDestroy(e);
  if (m->stream2) (void)hipStreamDestroy(m->stream2);
  double* bufs[] = {m->d_luh, m->d_lduh, m->d_qb, m->d_fb, m->d_h, m->d_state, m->d_gq, m->d_gf, m->d_halo, m->d_slots};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (m->d_faces) (void)hipFree(m->d_faces);
  if (m->d_picard) (void)hipFree(m->d_picard);
  if (m->d_guard) (void)hipFree(m->d_guard);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

// Face list: by direction, then by left cell (x fastest).  Periodic
// directions wrap; in the others the last cell's right face pairs with its
// ghost face (w = 1) and the first cell's left face with its ghost (w = 2).
std::vector<int> build_faces(const exa_mesh* m) {
  const int D = m->k.dim, n = m->n;
  const long per = (D == 3) ? long(n) * n : long(n);
  std::vector<int> faces;
  faces.reserve(size_t(4) * D * (n + 1) * per);
  auto push = [&faces](long a, long b, int d, int w) {
    faces.push_back(static_cast<int>(a));
    faces.push_back(static_cast<int>(b));
    faces.push_back(d);
    faces.push_back(w);
  };
  for (int d = 0; d < D; d++) {
    const bool periodic = m->bc.k[2 * d] == exa::kBcPeriodic;
    for (long c = 0; c < m->ncells; c++) {
      long ix = c % n, iy = (c / n) % n, iz = (D == 3) ? c / (long(n) * n) : 0;
      const long id = (d == 0) ? ix : (d == 1 ? iy : iz);
      const long tr = (D == 3) ? ((d == 0) ? iz * n + iy : (d == 1 ? iz * n + ix : iy * n + ix)) : ((d == 0) ? iy : ix);
      if (!periodic && id == 0) push(long(2 * d) * per + tr, c, d, 2);
      if (!periodic && id == n - 1) {
        push(c, long(2 * d + 1) * per + tr, d, 1);
        continue;
      }
      if (d == 0) ix = (ix + 1) % n;
      if (d == 1) iy = (iy + 1) % n;
      if (d == 2) iz = (iz + 1) % n;
      push(c, (D == 3) ? (iz * n + iy) * n + ix : iy * n + ix, d, 0);
    }
  }
  return faces;
}

void drop_graph(exa_mesh* m) {
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
}

}  // namespace

extern "C" {

int exa_mesh_create_uniform(exa_mesh** out, exa_ctx* ctx, int n, double h) {
  if (!out || !ctx || n < 1 || !(h > 0)) {
    exa::set_error("exa_mesh_create_uniform: invalid argument");
    return EXA_ERR_ARG;
  }
  *out = nullptr;
  exa_mesh* m = new exa_mesh;
  m->ctx = ctx;
  m->k = exa::ctx_cfg(ctx);
  const int D = m->k.dim, B = m->k.order + 1;
  m->n = n;
  m->h = h;
  m->ncells = (D == 3) ? long(n) * n * n : long(n) * n;
  m->nfaces = D * m->ncells;
  m->NS = (D == 3) ? B * B * B : B * B;
  m->NF = (D == 3) ? B * B : B;
  m->ND = m->k.nvar + m->k.npar;
  const int V = m->k.nvar;
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&m->d_luh, sizeof(double) * m->ncells * m->NS * m->ND);
  if (e == hipSuccess) e = hipMalloc(&m->d_lduh, sizeof(double) * m->ncells * m->NS * V);
  if (e == hipSuccess) e = hipMalloc(&m->d_qb, sizeof(double) * m->ncells * 2 * D * m->NF * m->ND);
  if (e == hipSuccess) e = hipMalloc(&m->d_fb, sizeof(double) * m->ncells * 2 * D * m->NF * V);
  if (e == hipSuccess) e = hipMalloc(&m->d_h, sizeof(double) * m->ncells);
  if (e == hipSuccess) e = hipMalloc(&m->d_state, sizeof(double) * 4);
  // room for the face list of any boundary choice: (n + 1) n^(D-1) entries per direction
  if (e == hipSuccess) e = hipMalloc(&m->d_faces, sizeof(int) * 4 * D * (n + 1) * (m->ncells / n));
  if (e == hipSuccess) e = hipMalloc(&m->d_picard, sizeof(int) * m->ncells);
  if (e == hipSuccess) e = hipMalloc(&m->d_slots, sizeof(double) * exa::kDtSlots * exa::kDtSlotStride);
  if (e == hipSuccess) e = hipMalloc(&m->d_guard, sizeof(int) * (2 + m->ncells));
  if (e != hipSuccess) {
    free_mesh(m);
    return fail(e, "exa_mesh_create_uniform");
  }
  const std::vector<int> faces = build_faces(m);  // all periodic
  std::vector<double> hs(m->ncells, h);
  const double st[4] = {0.0, 0.0, DBL_MAX, 0.0};
  e = hipMemcpy(m->d_faces, faces.data(), sizeof(int) * faces.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->d_h, hs.data(), sizeof(double) * hs.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->d_state, st, sizeof(st), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->d_picard, 0, sizeof(int) * m->ncells);
  if (e == hipSuccess) e = hipMemset(m->d_guard, 0, sizeof(int) * 2);
  if (e == hipSuccess) e = static_cast<hipError_t>(exa::launch_dt_slots_init(m->stream, m->d_slots));
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    free_mesh(m);
    return fail(e, "exa_mesh_create_uniform(upload)");
  }
  m->fused = exa::corrector_riemann_supported(m->k);
  *out = m;
  return EXA_OK;
}

int exa_mesh_destroy(exa_mesh* m) {
  if (!m) return EXA_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_mesh(m);
  return EXA_OK;
}

long exa_mesh_ncells(const exa_mesh* m) { return m ? m->ncells : 0; }

int exa_mesh_set_solution(exa_mesh* m, const double* luh) {
  if (!m || !luh) {
    exa::set_error("exa_mesh_set_solution: NULL argument");
    return EXA_ERR_ARG;
  }
  HIP_OK(hipMemcpyAsync(m->d_luh, luh, sizeof(double) * m->ncells * m->NS * m->ND, hipMemcpyHostToDevice,
                        m->stream),
         "set_solution");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_get_solution(exa_mesh* m, double* luh) {
  if (!m || !luh) {
    exa::set_error("exa_mesh_get_solution: NULL argument");
    return EXA_ERR_ARG;
  }
  HIP_OK(hipMemcpyAsync(luh, m->d_luh, sizeof(double) * m->ncells * m->NS * m->ND, hipMemcpyDeviceToHost,
                        m->stream),
         "get_solution");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

double* exa_mesh_device_solution(exa_mesh* m) { return m ? m->d_luh : nullptr; }

int exa_mesh_adjust_solution(exa_mesh* m, int ic, const double* params, int nparams) {
  if (!m || !params || nparams < 1) {
    exa::set_error("exa_mesh_adjust_solution: invalid argument");
    return EXA_ERR_ARG;
  }
  if ((ic != EXA_IC_DENSITY_WAVE && ic != EXA_IC_ENTROPY_WAVE) || m->k.pde.id != exa::kPdeEuler || m->k.nvar != 5) {
    exa::set_error("exa_mesh_adjust_solution: unknown initial condition for this PDE");
    return EXA_ERR_UNSUPPORTED;
  }
  const double* d_xi = reinterpret_cast<const double*>(reinterpret_cast<const char*>(m->k.tab) +
                                                        offsetof(exa::DevTables, x));
  if (ic == EXA_IC_ENTROPY_WAVE) {
    const int rc = status(exa::launch_entropy_wave(m->stream, m->ncells, m->n, m->k.dim, m->k.order + 1, m->ND, m->h,
                                                   d_xi, params[0], m->d_luh),
                          "entropy wave");
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(m->stream), "sync");
    return EXA_OK;
  }
  if (nparams < 7) {
    exa::set_error("exa_mesh_adjust_solution: the density wave takes at least 7 params");
    return EXA_ERR_ARG;
  }
  const int M = static_cast<int>(params[5]);
  if (M < 0 || nparams < 7 + 8 * M) {
    exa::set_error("exa_mesh_adjust_solution: params too short for the number of modes");
    return EXA_ERR_ARG;
  }
  double* d_prm = nullptr;
  HIP_OK(hipMalloc(&d_prm, sizeof(double) * nparams), "hipMalloc(params)");
  hipError_t e = hipMemcpyAsync(d_prm, params, sizeof(double) * nparams, hipMemcpyHostToDevice, m->stream);
  int rc = EXA_OK;
  if (e == hipSuccess) {
    rc = status(exa::launch_density_wave(m->stream, m->ncells, m->n, m->k.dim, m->k.order + 1, m->ND, m->h, d_xi,
                                         d_prm, m->d_luh),
                "density wave");
    e = hipStreamSynchronize(m->stream);
  }
  (void)hipFree(d_prm);
  if (rc) return rc;
  if (e != hipSuccess) return fail(e, "exa_mesh_adjust_solution");
  return EXA_OK;
}

int exa_mesh_init_dt(exa_mesh* m, double dt) {
  if (!m) {
    exa::set_error("exa_mesh_init_dt: NULL mesh");
    return EXA_ERR_ARG;
  }
  const double st[4] = {0.0, dt > 0 ? dt : DBL_MAX, DBL_MAX, 0.0};
  HIP_OK(hipMemcpyAsync(m->d_state, st, sizeof(st), hipMemcpyHostToDevice, m->stream), "init_dt");
  if (!(dt > 0)) {
    exa::LaunchOpts o;
    o.dt_slots = m->d_slots;
    int rc = status(exa::launch_stable_dt(m->k, m->stream, m->ncells, m->d_luh, m->d_h, m->d_state + 1, o),
                    "initial dt");
    if (!rc) rc = status(exa::launch_dt_reduce(m->stream, m->d_slots, m->d_state + 1), "dt reduction");
    if (rc) return rc;
  }
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_run(exa_mesh* m, int nsteps, int use_graph) {
  if (!m || nsteps < 0) {
    exa::set_error("exa_mesh_run: invalid argument");
    return EXA_ERR_ARG;
  }
  exa::TraceRange tr(use_graph ? "exa.mesh.run(graph)" : "exa.mesh.run");
  if (m->timing) {
    // events are recorded on the launch stream; nothing synchronises here
    while (m->ev.size() < size_t(4) * nsteps) {
      hipEvent_t e = nullptr;
      HIP_OK(hipEventCreate(&e), "event create");
      m->ev.push_back(e);
    }
    m->timed_steps = nsteps;
    for (int i = 0; i < nsteps; i++) {
      const int rc = enqueue_step(m, &m->ev[size_t(4) * i]);
      if (rc) return rc;
    }
    return EXA_OK;
  }
  if (use_graph) {
    if (!m->graph_exec) {
      HIP_OK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      const int rc = enqueue_step(m, nullptr);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(m->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(e, "end capture");
      m->graph = g;
      HIP_OK(hipGraphInstantiate(&m->graph_exec, m->graph, nullptr, nullptr, 0), "instantiate");
    }
    for (int i = 0; i < nsteps; i++) HIP_OK(hipGraphLaunch(m->graph_exec, m->stream), "graph launch");
    return EXA_OK;
  }
  for (int i = 0; i < nsteps; i++) {
    const int rc = enqueue_step(m, nullptr);
    if (rc) return rc;
  }
  return EXA_OK;
}

int exa_mesh_sync(exa_mesh* m) {
  if (!m) return EXA_ERR_ARG;
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_time(exa_mesh* m, double* t, double* dt) {
  if (!m) return EXA_ERR_ARG;
  double st[4];
  HIP_OK(hipMemcpyAsync(st, m->d_state, sizeof(st), hipMemcpyDeviceToHost, m->stream), "time");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  if (t) *t = st[0];
  if (dt) *dt = st[1];
  return EXA_OK;
}

int exa_mesh_picard_guard(exa_mesh* m, long* recomputed) {
  if (!m || !recomputed) {
    exa::set_error("exa_mesh_picard_guard: NULL argument");
    return EXA_ERR_ARG;
  }
  int g[2] = {0, 0};
  HIP_OK(hipMemcpyAsync(g, m->d_guard, sizeof(g), hipMemcpyDeviceToHost, m->stream), "picard guard");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  *recomputed = g[1];
  return EXA_OK;
}

int exa_mesh_picard(exa_mesh* m, int* counts) {
  if (!m || !counts) return EXA_ERR_ARG;
  HIP_OK(hipMemcpyAsync(counts, m->d_picard, sizeof(int) * m->ncells, hipMemcpyDeviceToHost, m->stream), "picard");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

void* exa_mesh_stream(exa_mesh* m) { return m ? static_cast<void*>(m->stream) : nullptr; }

int exa_mesh_set_fused(exa_mesh* m, int enabled) {
  if (!m) return EXA_ERR_ARG;
  if (enabled && !exa::corrector_riemann_supported(m->k)) {
    exa::set_error("exa_mesh_set_fused: the fused corrector is not compiled for this shape");
    return EXA_ERR_UNSUPPORTED;
  }
  m->fused = enabled != 0;
  drop_graph(m);
  return EXA_OK;
}

int exa_mesh_set_overlap(exa_mesh* m, int slabs) {
  if (!m || slabs < 0) {
    exa::set_error("exa_mesh_set_overlap: invalid argument");
    return EXA_ERR_ARG;
  }
  if (slabs > 0 && !slabs_apply(m, slabs)) {
    exa::set_error(
        "exa_mesh_set_overlap: needs 3 <= slabs <= n, a 3D fused step with all faces periodic and the "
        "one-iteration FP32 CERK predictor");
    return EXA_ERR_UNSUPPORTED;
  }
  if (slabs > 0) {
    if (!m->stream2) HIP_OK(hipStreamCreateWithFlags(&m->stream2, hipStreamNonBlocking), "stream create");
    while (m->sev.size() < size_t(slabs) + 2) {
      hipEvent_t e = nullptr;
      HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
      m->sev.push_back(e);
    }
  }
  m->slabs = slabs;
  drop_graph(m);
  return EXA_OK;
}

int exa_mesh_set_boundary(exa_mesh* m, const int* kinds) {
  if (!m || !kinds) {
    exa::set_error("exa_mesh_set_boundary: NULL argument");
    return EXA_ERR_ARG;
  }
  const int D = m->k.dim;
  exa::BcKinds bc{};
  unsigned mask = 0;
  for (int f = 0; f < 2 * D; f++) {
    if (kinds[f] < EXA_BC_PERIODIC || kinds[f] > EXA_BC_EXCHANGE) {
      exa::set_error("exa_mesh_set_boundary: unknown boundary kind");
      return EXA_ERR_ARG;
    }
    bc.k[f] = kinds[f];
    if (kinds[f] != EXA_BC_PERIODIC) mask |= 1u << f;
  }
  for (int d = 0; d < D; d++)
    if ((bc.k[2 * d] == EXA_BC_PERIODIC) != (bc.k[2 * d + 1] == EXA_BC_PERIODIC)) {
      exa::set_error("exa_mesh_set_boundary: a direction is periodic on both faces or on neither");
      return EXA_ERR_ARG;
    }
  bool hooks = false;
  for (int f = 0; f < 2 * D; f++) hooks = hooks || bc.k[f] == EXA_BC_WALL || bc.k[f] == EXA_BC_DIRICHLET_ENTROPY_WAVE;
  if (hooks && !exa::boundary_supported(m->k)) {
    exa::set_error("exa_mesh_set_boundary: the boundary-value hooks are defined for the Euler equations only");
    return EXA_ERR_UNSUPPORTED;
  }
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  if (mask && !m->d_gq) {
    const size_t nghost = size_t(2 * D) * (m->ncells / m->n) * m->NF;
    HIP_OK(hipMalloc(&m->d_gq, sizeof(double) * nghost * m->ND), "hipMalloc(ghost faces)");
    HIP_OK(hipMalloc(&m->d_gf, sizeof(double) * nghost * m->k.nvar), "hipMalloc(ghost fluxes)");
  }
  m->bc = bc;
  m->bcmask = mask;
  m->hooks = hooks;
  const std::vector<int> faces = build_faces(m);
  m->nfaces = static_cast<long>(faces.size() / 4);
  HIP_OK(hipMemcpy(m->d_faces, faces.data(), sizeof(int) * faces.size(), hipMemcpyHostToDevice), "upload faces");
  drop_graph(m);
  return EXA_OK;
}

int exa_mesh_global_observables(exa_mesh* m, double* out, int nout) {
  if (!m || !out || nout < 2 + 3 * m->ND) {
    exa::set_error("exa_mesh_global_observables: invalid argument (out needs 2 + 3 nData values)");
    return EXA_ERR_ARG;
  }
  const int nobs = 2 + 3 * m->ND;
  double *d_part = nullptr, *d_out = nullptr;
  hipError_t e = hipMalloc(&d_part, sizeof(double) * exa::observables_blocks(m->ncells) * 3 * m->ND);
  if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * nobs);
  int rc = EXA_OK;
  if (e == hipSuccess) {
    rc = status(exa::launch_observables(m->k, m->stream, m->ncells, m->h, m->d_luh, d_part, d_out),
                "global observables");
    if (!rc) e = hipMemcpyAsync(out, d_out, sizeof(double) * nobs, hipMemcpyDeviceToHost, m->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(m->stream);
  }
  if (d_part) (void)hipFree(d_part);
  if (d_out) (void)hipFree(d_out);
  if (rc) return rc;
  if (e != hipSuccess) return fail(e, "exa_mesh_global_observables");
  return EXA_OK;
}

int exa_mesh_probe(exa_mesh* m, int npoints, const double* points, double* values) {
  if (!m || npoints < 0 || (npoints > 0 && (!points || !values))) {
    exa::set_error("exa_mesh_probe: invalid argument");
    return EXA_ERR_ARG;
  }
  if (npoints == 0) return EXA_OK;
  const int D = m->k.dim;
  const double L = m->n * m->h;
  for (long i = 0; i < long(npoints) * D; i++)
    if (!(points[i] >= 0.0 && points[i] < L)) {
      exa::set_error("exa_mesh_probe: a point lies outside the mesh");
      return EXA_ERR_ARG;
    }
  double *d_pts = nullptr, *d_out = nullptr;
  const size_t npb = sizeof(double) * npoints * D, nob = sizeof(double) * npoints * m->ND;
  hipError_t e = hipMalloc(&d_pts, npb);
  if (e == hipSuccess) e = hipMalloc(&d_out, nob);
  if (e == hipSuccess) e = hipMemcpyAsync(d_pts, points, npb, hipMemcpyHostToDevice, m->stream);
  int rc = EXA_OK;
  if (e == hipSuccess) {
    rc = status(exa::launch_probe(m->k, m->stream, m->n, m->h, m->d_luh, npoints, d_pts, d_out), "probe");
    if (!rc) e = hipMemcpyAsync(values, d_out, nob, hipMemcpyDeviceToHost, m->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(m->stream);
  }
  if (d_pts) (void)hipFree(d_pts);
  if (d_out) (void)hipFree(d_out);
  if (rc) return rc;
  if (e != hipSuccess) return fail(e, "exa_mesh_probe");
  return EXA_OK;
}

int exa_mesh_step_predict(exa_mesh* m) {
  if (!m) return EXA_ERR_ARG;
  return enqueue_predict(m, nullptr);
}

int exa_mesh_step_correct(exa_mesh* m) {
  if (!m) return EXA_ERR_ARG;
  return enqueue_correct(m, nullptr);
}

int exa_mesh_step_advance(exa_mesh* m) {
  if (!m) return EXA_ERR_ARG;
  return enqueue_advance(m, nullptr);
}

int exa_mesh_halo_size(const exa_mesh* m, long* n_q, long* n_f) {
  if (!m) return EXA_ERR_ARG;
  const long per = m->ncells / m->n;
  if (n_q) *n_q = per * m->NF * m->ND;
  if (n_f) *n_f = per * m->NF * m->k.nvar;
  return EXA_OK;
}

int exa_mesh_halo_pack(exa_mesh* m, int face, double* d_msg_q, double* d_msg_f) {
  if (!m || !d_msg_q || !d_msg_f || face < 0 || face >= 2 * m->k.dim) {
    exa::set_error("exa_mesh_halo_pack: invalid argument");
    return EXA_ERR_ARG;
  }
  if (f32_faces(m)) {
    // the all-periodic FP32 step keeps its faces as FP32 [face][var][node]
    exa::set_error("exa_mesh_halo_pack: the mesh's faces are FP32 (all-periodic FP32 step); "
                   "halo messages need a domain face of kind EXA_BC_EXCHANGE");
    return EXA_ERR_UNSUPPORTED;
  }
  return status(exa::launch_halo_pack(m->k, m->stream, m->n, face, m->d_qb, m->d_fb, d_msg_q, d_msg_f), "halo pack");
}

int exa_mesh_halo_unpack(exa_mesh* m, int face, const double* d_msg_q, const double* d_msg_f) {
  if (!m || !d_msg_q || !d_msg_f || face < 0 || face >= 2 * m->k.dim) {
    exa::set_error("exa_mesh_halo_unpack: invalid argument");
    return EXA_ERR_ARG;
  }
  if (m->bc.k[face] != EXA_BC_EXCHANGE) {
    exa::set_error("exa_mesh_halo_unpack: the face is not of kind EXA_BC_EXCHANGE");
    return EXA_ERR_ARG;
  }
  long nq = 0, nf = 0;
  exa_mesh_halo_size(m, &nq, &nf);
  HIP_OK(hipMemcpyAsync(m->d_gq + size_t(face) * nq, d_msg_q, sizeof(double) * nq, hipMemcpyDeviceToDevice, m->stream),
         "halo unpack");
  HIP_OK(hipMemcpyAsync(m->d_gf + size_t(face) * nf, d_msg_f, sizeof(double) * nf, hipMemcpyDeviceToDevice, m->stream),
         "halo unpack");
  return EXA_OK;
}

int exa_mesh_halo_pack_host(exa_mesh* m, int face, double* msg_q, double* msg_f) {
  if (!m || !msg_q || !msg_f) return EXA_ERR_ARG;
  long nq = 0, nf = 0;
  exa_mesh_halo_size(m, &nq, &nf);
  if (!m->d_halo) HIP_OK(hipMalloc(&m->d_halo, sizeof(double) * (nq + nf)), "hipMalloc(halo staging)");
  const int rc = exa_mesh_halo_pack(m, face, m->d_halo, m->d_halo + nq);
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(msg_q, m->d_halo, sizeof(double) * nq, hipMemcpyDeviceToHost, m->stream), "halo copy");
  HIP_OK(hipMemcpyAsync(msg_f, m->d_halo + nq, sizeof(double) * nf, hipMemcpyDeviceToHost, m->stream), "halo copy");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_halo_unpack_host(exa_mesh* m, int face, const double* msg_q, const double* msg_f) {
  if (!m || !msg_q || !msg_f) return EXA_ERR_ARG;
  long nq = 0, nf = 0;
  exa_mesh_halo_size(m, &nq, &nf);
  if (!m->d_halo) HIP_OK(hipMalloc(&m->d_halo, sizeof(double) * (nq + nf)), "hipMalloc(halo staging)");
  // the previous use of the staging buffer has completed: every halo call synchronises
  HIP_OK(hipMemcpyAsync(m->d_halo, msg_q, sizeof(double) * nq, hipMemcpyHostToDevice, m->stream), "halo copy");
  HIP_OK(hipMemcpyAsync(m->d_halo + nq, msg_f, sizeof(double) * nf, hipMemcpyHostToDevice, m->stream), "halo copy");
  const int rc = exa_mesh_halo_unpack(m, face, m->d_halo, m->d_halo + nq);
  if (rc) return rc;
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

double* exa_mesh_device_state(exa_mesh* m) { return m ? m->d_state : nullptr; }

int exa_mesh_dt_next(exa_mesh* m, double* dt_next) {
  if (!m || !dt_next) return EXA_ERR_ARG;
  HIP_OK(hipMemcpyAsync(dt_next, m->d_state + 2, sizeof(double), hipMemcpyDeviceToHost, m->stream), "dt_next");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_set_dt_next(exa_mesh* m, double dt_next) {
  if (!m || !(dt_next > 0)) return EXA_ERR_ARG;
  HIP_OK(hipMemcpyAsync(m->d_state + 2, &dt_next, sizeof(double), hipMemcpyHostToDevice, m->stream), "dt_next");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_set_time(exa_mesh* m, double t) {
  if (!m) return EXA_ERR_ARG;
  HIP_OK(hipMemcpyAsync(m->d_state, &t, sizeof(double), hipMemcpyHostToDevice, m->stream), "set_time");
  HIP_OK(hipStreamSynchronize(m->stream), "sync");
  return EXA_OK;
}

int exa_mesh_set_timing(exa_mesh* m, int enabled) {
  if (!m) return EXA_ERR_ARG;
  m->timing = enabled != 0;
  return EXA_OK;
}

int exa_mesh_phase_times(exa_mesh* m, double* a, double* b, double* c, int* n) {
  if (!m) return EXA_ERR_ARG;
  double ms[3] = {0, 0, 0};
  if (m->timed_steps > 0) HIP_OK(hipEventSynchronize(m->ev[size_t(4) * m->timed_steps - 1]), "event sync");
  for (int i = 0; i < m->timed_steps; i++) {
    for (int p = 0; p < 3; p++) {
      float x = 0;
      HIP_OK(hipEventElapsedTime(&x, m->ev[size_t(4) * i + p], m->ev[size_t(4) * i + p + 1]), "elapsed");
      ms[p] += x;
    }
  }
  if (a) *a = ms[0];
  if (b) *b = ms[1];
  if (c) *c = ms[2];
  if (n) *n = m->timed_steps;
  return EXA_OK;
}

}  // extern "C"

"""GPU: the batched device entry points (exa_batch_predictor, exa_batch_riemann,
exa_batch_corrector, exa_batch_stable_dt) driven by a host traversal that owns
the cell ordering, as a host code replacing the reference's per-cell loop
(ADERDGSolver.cpp: fusedSpaceTimePredictorVolumeIntegral per cell, riemannSolver
per face, surfaceIntegral + update + stableTimeStepSize per cell) would call them.

Device arrays come from the HIP runtime through ctypes.  One global step on a
periodic mesh, face list {left, right, direction} built here, must equal the
mesh driver's four-kernel step bit for bit (the same kernels, faces are
independent) and its fused step to the FP64 parity bar; the stable dt of the
batch and the next dt of the corrector equal the mesh's.  Empty batches (0
cells or faces, null pointers) are no-ops; negative counts and missing arrays
are argument errors.
"""
import ctypes

import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P

pytestmark = pytest.mark.gpu

_hip = ctypes.CDLL("libamdhip64.so")
_H2D, _D2H = 1, 2


class Dev:
    """A device array of float64 / int32 values (hipMalloc, synchronous copies)."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = ctypes.c_void_p()
        assert _hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 8))) == 0
        assert _hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes),
                              _H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert _hip.hipDeviceSynchronize() == 0
        assert _hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), _D2H) == 0
        return out

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = None


def rel_l2(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


def periodic_faces(n, D):
    """{left, right, direction, 0} for every face of the periodic n^D mesh
    (cells x fastest)."""
    faces = []
    for c in range(n ** D):
        idx = [(c // n ** d) % n for d in range(D)]
        for d in range(D):
            j = list(idx)
            j[d] = (j[d] + 1) % n
            faces.append([c, sum(j[e] * n ** e for e in range(D)), d, 0])
    return np.array(faces, dtype=np.int32)


def state(n, order, D, seed):
    pts = P.node_points(n, order, D)
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, 6)
    tp = 2 * np.pi
    x, y = pts[..., 0], pts[..., 1]
    z = pts[..., 2] if D == 3 else 0.0 * x
    rho = 1 + 0.2 * np.sin(tp * (x + c[0])) * np.cos(tp * (y + c[1])) + 0.05 * np.sin(tp * (z + c[2]))
    u, v = 0.3 * np.cos(tp * (y + c[3])), 0.2 * np.sin(tp * (x + c[4]))
    w = 0.1 * np.cos(tp * (x + z + c[5])) if D == 3 else 0.0 * x
    p = 1 + 0.1 * np.sin(tp * (x + y))
    return np.ascontiguousarray(P.euler_state(rho, u, v, w, p).ravel())


@pytest.mark.parametrize("D,order,n", [(3, 3, 4), (2, 4, 5), (3, 5, 3)])
def test_batched_step_equals_the_mesh_step(D, order, n):
    print(f"[batch step] D={D} order={order} n={n}", flush=True)
    ctx = E.Context(D, order, precision=E.PREC_FP64)
    B, V = order + 1, 5
    nc = n ** D
    nface_vals = 2 * D * B ** (D - 1) * V
    luh = state(n, order, D, seed=D + order)
    h = 1.0 / n
    faces = periodic_faces(n, D)
    bufs = {
        "luh": Dev(luh), "h": Dev(np.full(nc, h)), "dtmin": Dev(np.array([np.finfo(np.float64).max])),
        "lduh": Dev(np.zeros(nc * B ** D * V)), "qb": Dev(np.zeros(nc * nface_vals)),
        "fb": Dev(np.zeros(nc * nface_vals)), "pic": Dev(np.zeros(nc, dtype=np.int32)), "faces": Dev(faces),
        "dt_next": Dev(np.array([np.finfo(np.float64).max])),
    }
    b = {k: v.ptr for k, v in bufs.items()}
    try:
        # global dt = min over cells of stableTimeStepSize
        ctx.batch_stable_dt(None, nc, b["luh"], b["h"], b["dtmin"])
        dt = bufs["dtmin"].get()[0]
        ctx.batch_predictor(None, nc, b["luh"], b["h"], b["dtmin"], b["lduh"], b["qb"], b["fb"], b["pic"])
        ctx.batch_riemann(None, len(faces), b["faces"], b["qb"], b["fb"])
        ctx.batch_corrector(None, nc, b["luh"], b["lduh"], b["fb"], b["h"], b["dtmin"], b["dt_next"])
        got = bufs["luh"].get()
        dt_next = bufs["dt_next"].get()[0]
        assert np.all(bufs["pic"].get() >= 0)
    finally:
        for v in bufs.values():
            v.free()

    for fused in (False, True):
        m = E.Mesh(ctx, n, h)
        try:
            m.set_fused(fused)
        except E.ExaError:
            assert fused  # fused corrector not compiled for this shape
            m.close()
            continue
        m.set_solution(luh)
        m.init_dt(0.0)
        assert m.time()[1] == dt
        m.run(1)
        m.sync()
        ref = m.get_solution()
        t, dtm = m.time()
        m.close()
        assert t == dt
        assert dtm == dt_next
        if fused:
            assert rel_l2(got, ref) <= 1e-12
        else:
            np.testing.assert_array_equal(got, ref)
    assert rel_l2(got, luh) > 1e-6  # the step changed the state
    ctx.close()


def test_empty_batches_and_argument_errors():
    print("[batch args]", flush=True)
    ctx = E.Context(3, 3, precision=E.PREC_FP64)
    # empty batches: no launch, no pointer needed
    ctx.batch_predictor(None, 0, None, None, None, None, None, None)
    ctx.batch_riemann(None, 0, None, None, None)
    ctx.batch_corrector(None, 0, None, None, None, None, None, None)
    ctx.batch_stable_dt(None, 0, None, None, None)
    with pytest.raises(E.ExaError):
        ctx.batch_predictor(None, -1, None, None, None, None, None, None)
    with pytest.raises(E.ExaError):
        ctx.batch_riemann(None, 4, None, None, None)
    with pytest.raises(E.ExaError):
        ctx.batch_corrector(None, 2, None, None, None, None, None, None)
    with pytest.raises(E.ExaError):
        ctx.batch_stable_dt(None, -3, None, None, None)
    ctx.close()


def test_batched_step_with_per_direction_dx():
    """exa_ctx_set_cell_dx(1): d_h holds dx[D] per cell.  One step of a periodic
    box of non-cubic cells through the batched entry points against the
    oracle's per-cell kernels composed in the reference's order (stable dt,
    fusedSpaceTimePredictorVolumeIntegral, riemannSolver per face,
    surfaceIntegral, solutionUpdate), FP64 bar 1e-12."""
    import oracle as O

    D, order, n = 3, 3, 3
    B, V = order + 1, 5
    nc = n ** D
    dx = np.array([1.0 / n, 0.7 / n, 1.3 / n])
    nfv = 2 * D * B ** (D - 1) * V
    luh = state(n, order, D, seed=11)
    faces = periodic_faces(n, D)
    ctx = E.Context(D, order)
    ctx.set_cell_dx(True)
    bufs = {
        "luh": Dev(luh), "h": Dev(np.tile(dx, nc)), "dtmin": Dev(np.array([np.finfo(np.float64).max])),
        "lduh": Dev(np.zeros(nc * B ** D * V)), "qb": Dev(np.zeros(nc * nfv)), "fb": Dev(np.zeros(nc * nfv)),
        "pic": Dev(np.zeros(nc, dtype=np.int32)), "faces": Dev(faces),
        "dt_next": Dev(np.array([np.finfo(np.float64).max])),
    }
    b = {k: v.ptr for k, v in bufs.items()}
    try:
        ctx.batch_stable_dt(None, nc, b["luh"], b["h"], b["dtmin"])
        dt = bufs["dtmin"].get()[0]
        ctx.batch_predictor(None, nc, b["luh"], b["h"], b["dtmin"], b["lduh"], b["qb"], b["fb"], b["pic"])
        ctx.batch_riemann(None, len(faces), b["faces"], b["qb"], b["fb"])
        ctx.batch_corrector(None, nc, b["luh"], b["lduh"], b["fb"], b["h"], b["dtmin"], b["dt_next"])
        got, pic = bufs["luh"].get(), bufs["pic"].get()
    finally:
        for v in bufs.values():
            v.free()
    ctx.close()

    o = O.Oracle(D, order)
    cells = luh.reshape(nc, -1)
    dt_o = min(o.stable_dt(cells[c], dx) for c in range(nc))
    assert abs(dt / dt_o - 1) <= 1e-14
    res = [o.fused(cells[c], dx, dt_o) for c in range(nc)]
    nf1 = B ** (D - 1) * V
    qb = np.stack([r["lQhbnd"] for r in res]).reshape(nc, 2 * D, -1)
    fb = np.stack([r["lFhbnd"] for r in res]).reshape(nc, 2 * D, nf1).copy()
    for c, r, d, _ in faces:
        FL, FR = o.riemann(fb[c, 2 * d + 1], fb[r, 2 * d], qb[c, 2 * d + 1], qb[r, 2 * d], d)
        fb[c, 2 * d + 1], fb[r, 2 * d] = FL, FR
    ref = np.concatenate([o.solution_update(cells[c], o.surface_integral(res[c]["lduh"], fb[c].ravel(), dx), dt_o)
                          for c in range(nc)])
    assert rel_l2(got, ref) <= 1e-12, rel_l2(got, ref)
    assert rel_l2(got - luh, ref - luh) <= 1e-10
    np.testing.assert_array_equal(pic, [r["iters"] for r in res])

"""GPU parity: the HIP library (through the C ABI) against the CPU oracle.

FP64 bar: <= 1e-12 relative L2 per array (the parity bar of the FP64 form)
and identical per-cell Picard counts.  Inputs are seeded; sizes are
small enough for the oracle to finish in seconds.
"""
import json
import os

import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_testdata.json")))["data"]
TOL = 1e-12


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def smooth_state(dim, order, npar=0, seed=0, amp=0.05):
    """A non-uniform smooth Euler state on one cell (positive density/pressure)."""
    rng = np.random.default_rng(seed)
    B = order + 1
    xi = np.polynomial.legendre.leggauss(B)[0] * 0.5 + 0.5
    grids = np.meshgrid(*([xi] * dim), indexing="ij")  # z,y,x order for dim 3
    ns = B ** dim
    Q = np.zeros((ns, 5 + npar))
    c = rng.uniform(-1, 1, size=(5, dim + 1))
    for v in range(5):
        f = c[v, 0] + sum(c[v, d + 1] * np.sin(2 * np.pi * grids[d] + v) for d in range(dim))
        Q[:, v] = f.ravel()
    rho = 1.0 + amp * Q[:, 0]
    u = 0.3 + amp * Q[:, 1]
    vv = -0.2 + amp * Q[:, 2]
    w = (0.1 + amp * Q[:, 3]) if dim == 3 else 0.0 * rho
    p = 1.0 + amp * Q[:, 4]
    Q[:, 0] = rho
    Q[:, 1] = rho * u
    Q[:, 2] = rho * vv
    Q[:, 3] = rho * w
    Q[:, 4] = p / 0.4 + 0.5 * rho * (u * u + vv * vv + w * w)
    if npar:
        Q[:, 5:] = rng.uniform(0.5, 1.5, size=(ns, npar))
    return Q.ravel()


@pytest.fixture(scope="module")
def ctxs():
    cache = {}

    def get(dim, order, npar=0, picard_max=0, precision=E.PREC_FP64):
        key = (dim, order, npar, picard_max, precision)
        if key not in cache:
            cache[key] = E.Context(dim, order, 5, npar, precision=precision, picard_max=picard_max)
        return cache[key]

    yield get
    for c in cache.values():
        c.close()


# --------------------------------------------------------------------------
# per-cell compatibility entry points vs oracle and golden vectors
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,dt,dx", [(3, 1.083937460199773e-3, 0.05), (2, 1.686854344081342e-3, 0.05)])
def test_predictor_reference_golden(ctxs, dim, dt, dx):
    g = GOLD[f"dim{dim}"]["testSpaceTimePredictorNonlinear"]
    B, ns = 4, 4 ** dim
    luh = np.zeros((ns, 6))
    luh[:, :5] = np.array(g["luh"]).reshape(ns, 5)
    r = ctxs(dim, 3, npar=1).fusedSpaceTimePredictorVolumeIntegral(luh.ravel(), [dx] * dim, 0.0, dt)
    np.testing.assert_allclose(r["lQhbnd"].reshape(-1, 6)[:, :5].ravel(), g["lQhbnd"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(r["lFhbnd"], g["lFhbnd"], atol=1e-10, rtol=0)
    assert r["iters"] == (0 if dim == 3 else 2 * 4 + 2)


@pytest.mark.parametrize("dim,order,npar", [(3, 5, 0), (3, 3, 0), (3, 4, 0), (2, 3, 0), (2, 4, 0), (3, 3, 1), (2, 3, 1)])
@pytest.mark.parametrize("seed", [0, 1])
def test_predictor_vs_oracle(ctxs, dim, order, npar, seed):
    ctx = ctxs(dim, order, npar)
    orc = O.Oracle(dim, order, 5, npar)
    luh = smooth_state(dim, order, npar, seed)
    h = 1.0 / 27
    dt = 0.9 * orc.stable_dt(luh, [h] * dim)
    r = ctx.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * dim, 0.0, dt)
    o = orc.fused(luh, [h] * dim, dt)
    assert r["iters"] == o["iters"]
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        assert rel_l2(r[k], o[k]) < TOL, (k, rel_l2(r[k], o[k]))


@pytest.mark.parametrize("dim", [2, 3])
def test_predictor_add_to_solution(ctxs, dim):
    ctx = ctxs(dim, 3)
    orc = O.Oracle(dim, 3)
    luh = smooth_state(dim, 3, 0, 3)
    h, dt = 0.1, 1e-3
    r = ctx.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * dim, 0.0, dt, addVolumeIntegralResultToUpdate=False)
    o = orc.fused(luh, [h] * dim, dt)
    expect = orc.solution_update(luh, o["lduh"], dt)
    assert rel_l2(r["luh"], expect) < TOL


@pytest.mark.parametrize("dim,order,picard", [(3, 5, 6), (2, 3, 4), (3, 4, 2)])
def test_predictor_fixed_iterations(ctxs, dim, order, picard):
    ctx = ctxs(dim, order, picard_max=picard)
    orc = O.Oracle(dim, order, max_picard=picard)
    luh = smooth_state(dim, order, 0, 5)
    h = 1.0 / 9
    dt = 0.9 * orc.stable_dt(luh, [h] * dim)
    r = ctx.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * dim, 0.0, dt)
    o = orc.fused(luh, [h] * dim, dt)
    assert r["iters"] == o["iters"] == picard
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        assert rel_l2(r[k], o[k]) < TOL, k


def test_riemann_reference_golden_3d(ctxs):
    g = GOLD["dim3"]
    QL = np.zeros((16, 6))
    QR = np.zeros((16, 6))
    QL[:, :5] = np.array(g["testRiemannSolver"]["QL"]).reshape(16, 5)
    QR[:, :5] = np.array(g["testRiemannSolver"]["QR"]).reshape(16, 5)
    FL, FR = ctxs(3, 3, npar=1).riemannSolver(np.arange(80) + 1.0, np.arange(80) - 1.0, QL.ravel(), QR.ravel(), 1)
    np.testing.assert_allclose(FL, g["testRiemannSolverNonlinear"]["FL"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(FR, g["testRiemannSolverNonlinear"]["FR"], atol=1e-10, rtol=0)


@pytest.mark.parametrize("case,direction", [(1, 0), (2, 1)])
def test_riemann_reference_golden_2d(ctxs, case, direction):
    R = GOLD["dim2"]["testRiemannSolverNonlinear"]
    QL = np.zeros((4, 6))
    QR = np.zeros((4, 6))
    QL[:, :5] = np.array(R[f"QL_{case}_in"]).reshape(4, 5)
    QR[:, :5] = np.array(R[f"QR_{case}_in"]).reshape(4, 5)
    FL, FR = ctxs(2, 3, npar=1).riemannSolver(R[f"FL_{case}_in"], R[f"FR_{case}_in"], QL.ravel(), QR.ravel(), direction)
    np.testing.assert_allclose(FL, R[f"F_{case}_out"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(FR, R[f"F_{case}_out"], atol=1e-10, rtol=0)


@pytest.mark.parametrize("dim,order", [(3, 5), (2, 3)])
def test_riemann_vs_oracle(ctxs, dim, order):
    ctx = ctxs(dim, order)
    orc = O.Oracle(dim, order)
    nf = (order + 1) ** (dim - 1)
    rng = np.random.default_rng(7)
    for d in range(dim):
        QL = smooth_state(dim - 1 if dim > 2 else 2, order, 0, 11 + d)[: nf * 5]
        QR = smooth_state(dim - 1 if dim > 2 else 2, order, 0, 21 + d)[: nf * 5]
        FL0 = rng.standard_normal(nf * 5)
        FR0 = rng.standard_normal(nf * 5)
        FL, FR = ctx.riemannSolver(FL0, FR0, QL, QR, d)
        oL, oR = orc.riemann(FL0, FR0, QL, QR, d)
        assert rel_l2(FL, oL) < TOL and rel_l2(FR, oR) < TOL
        np.testing.assert_array_equal(FL, FR)


@pytest.mark.parametrize("dim,order", [(3, 5), (3, 3), (2, 3)])
def test_face_surface_update_dt_vs_oracle(ctxs, dim, order):
    ctx = ctxs(dim, order)
    orc = O.Oracle(dim, order)
    rng = np.random.default_rng(3)
    ns, nf = (order + 1) ** dim, (order + 1) ** (dim - 1)
    lduh = rng.standard_normal(ns * 5)
    lF = rng.standard_normal(2 * dim * nf * 5)
    dx = [0.07] * dim
    for d in range(dim):
        for o in range(2):
            face = lF[(2 * d + o) * nf * 5:(2 * d + o + 1) * nf * 5]
            got = ctx.faceIntegral(lduh, face, d, o, dx)
            exp = orc.face_integral(lduh, face, d, o, dx)
            assert rel_l2(got, exp) < TOL
    assert rel_l2(ctx.surfaceIntegral(lduh, lF, dx), orc.surface_integral(lduh, lF, dx)) < TOL
    luh = smooth_state(dim, order, 0, 9)
    assert rel_l2(ctx.addUpdateToSolution(luh, lduh, 1e-3), orc.solution_update(luh, lduh, 1e-3)) < TOL
    assert abs(ctx.stableTimeStepSize(luh, dx) / orc.stable_dt(luh, dx) - 1) < 1e-14
    # faceIntegral with addToUpdate == false adds dt/w * integral to luh
    face = lF[:nf * 5]
    got = ctx.faceIntegral(luh, face, 0, 0, dx, dt=2e-3, addToUpdate=False)
    exp = orc.solution_update(luh, orc.face_integral(np.zeros(ns * 5), face, 0, 0, dx), 2e-3)
    assert rel_l2(got, exp) < TOL


def test_error_paths_do_not_abort(ctxs):
    ctx = ctxs(3, 3)
    with pytest.raises(E.ExaError) as ei:
        ctx.riemannSolver(np.zeros(80), np.zeros(80), np.ones(80), np.ones(80), 3)
    assert ei.value.code == E.EXA_ERR_ARG
    with pytest.raises(E.ExaError):
        ctx.faceIntegral(np.zeros(320), np.zeros(80), 5, 0, [0.1] * 3)


# --------------------------------------------------------------------------
# whole time steps on a periodic mesh vs the oracle driver
# --------------------------------------------------------------------------
def run_both(dim, order, n, steps, picard_max=0, precision=E.PREC_FP64, init=None):
    ctx = E.Context(dim, order, 5, 0, precision=precision, picard_max=picard_max)
    orc = O.Oracle(dim, order, max_picard=picard_max if picard_max > 0 else None)
    luh = init if init is not None else (P.density_wave_3d(n, order) if dim == 3 else P.density_wave_2d(n, order))
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.set_solution(luh)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got = mesh.get_solution()
    t, dt_next = mesh.time()
    pic = mesh.picard()
    ref = luh.copy()
    dt_o, pic_o = orc.run_uniform(ref, n, 1.0 / n, steps, 0.0, want_picard=True)
    mesh.close()
    ctx.close()
    return got, ref, luh, (t, dt_next, pic), (dt_o, pic_o)


def test_config1_2d_euler_order3_27x27_10_steps():
    """The reference CPU-runnable configuration, run to the end on both sides."""
    got, ref, luh0, (t, dtn, pic), (dt_o, pic_o) = run_both(2, 3, 27, 10)
    assert rel_l2(got, ref) < TOL, rel_l2(got, ref)
    assert rel_l2(got - luh0, ref - luh0) < 1e-9
    assert abs(dtn / dt_o - 1) < 1e-12
    np.testing.assert_array_equal(pic, pic_o)


@pytest.mark.parametrize("order,n,steps", [(5, 4, 3), (4, 5, 3), (3, 6, 4)])
def test_mesh_3d_vs_oracle(order, n, steps):
    got, ref, luh0, (t, dtn, pic), (dt_o, pic_o) = run_both(3, order, n, steps)
    assert rel_l2(got, ref) < TOL, rel_l2(got, ref)
    assert rel_l2(got - luh0, ref - luh0) < 1e-9
    assert abs(dtn / dt_o - 1) < 1e-12
    np.testing.assert_array_equal(pic, pic_o)


def test_graph_replay_matches_eager():
    ctx = E.Context(3, 3, 5, 0)
    n = 6
    luh = P.density_wave_3d(n, 3)
    outs = []
    for use_graph in (False, True):
        mesh = E.Mesh(ctx, n, 1.0 / n)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        res = []
        for _ in range(3):  # several replays, host work and syncs in between
            mesh.run(2, use_graph=use_graph)
            mesh.sync()
            res.append(mesh.get_solution())
        outs.append((res, mesh.time()))
        mesh.close()
    for a, b in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(a, b)
    assert outs[0][1] == outs[1][1]
    ctx.close()


def test_fp32_predictor_accuracy_3d_order5():
    """FP32 predictor with the adaptive Picard rule (line kernel, two-tile form)
    stays within 1e-6 of the adaptive FP64 oracle."""
    n, steps = 4, 3
    got, ref, luh0, _, _ = run_both(3, 5, n, steps, picard_max=0, precision=E.PREC_FP32)
    g = got.reshape(-1, 5)
    r = ref.reshape(-1, 5)
    for v in range(5):
        assert rel_l2(g[:, v], r[:, v]) < 1e-6, (v, rel_l2(g[:, v], r[:, v]))


def test_conservation_and_determinism_full_size():
    """At the full 81^3 N=5 size: total mass is conserved and a rerun is bitwise identical."""
    n, order = 81, 5
    ctx = E.Context(3, order, 5, 0)
    luh = P.density_wave_3d(n, order, modes=1)
    w = E.tables(order)["w"]
    W = np.einsum("i,j,k->ijk", w, w, w).ravel()
    h = 1.0 / n

    def mass(u):
        return float(np.einsum("cn,n->", u.reshape(-1, 216, 5)[:, :, 0], W)) * h ** 3

    m0 = mass(luh)
    results = []
    for _ in range(2):
        mesh = E.Mesh(ctx, n, h)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        mesh.run(2)
        mesh.sync()
        results.append(mesh.get_solution())
        mesh.close()
    assert abs(mass(results[0]) / m0 - 1) < 1e-12
    np.testing.assert_array_equal(results[0], results[1])
    ctx.close()


def test_fp32_cerk_form_conservation_and_determinism_full_size():
    """The FP32 form with the CERK guess and one collapsed Picard iteration (the
    one-iteration line kernel, FP32 faces in the fused step) at the full 81^3
    N=5 size: each face flux is the same bits for both cells, so mass changes
    only by the FP32 rounding of the volume terms (bound 1e-9 over two steps);
    reruns are bitwise identical."""
    print("[full size fp32 cerk] 81^3 N=5", flush=True)
    n, order = 81, 5
    ctx = E.Context(3, order, 5, 0, precision=E.PREC_FP32, picard_max=1)
    ctx.set_initial_guess(E.GUESS_CERK)
    luh = P.density_wave_3d(n, order, modes=1)
    w = E.tables(order)["w"]
    W = np.einsum("i,j,k->ijk", w, w, w).ravel()
    h = 1.0 / n

    def mass(u):
        return float(np.einsum("cn,n->", u.reshape(-1, 216, 5)[:, :, 0], W)) * h ** 3

    m0 = mass(luh)
    results = []
    for _ in range(2):
        mesh = E.Mesh(ctx, n, h)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        mesh.run(2)
        mesh.sync()
        assert np.all(mesh.picard() == 1)
        results.append(mesh.get_solution())
        mesh.close()
    assert abs(mass(results[0]) / m0 - 1) < 1e-9
    np.testing.assert_array_equal(results[0], results[1])
    assert rel_l2(results[0], luh) > 1e-6  # the steps changed the state
    ctx.close()


@pytest.mark.parametrize("dim,n,order", [(3, 3, 5), (3, 4, 3), (2, 9, 3)])
def test_device_initial_condition_matches_host(dim, n, order):
    """exa_mesh_adjust_solution (device) reproduces the host generator."""
    ctx = E.Context(dim, order)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.adjust_solution(P.density_wave_params(dim))
    got = mesh.get_solution()
    ref = P.density_wave_3d(n, order) if dim == 3 else P.density_wave_2d(n, order)
    mesh.close()
    ctx.close()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize("precision,variant,tol", [(E.PREC_FP64, 2, 1e-13), (E.PREC_FP32, 2, 2e-6),
                                                   (E.PREC_FP32, 3, 2e-6), (E.PREC_FP32, 4, 2e-6),
                                                   (E.PREC_FP32, 0, 2e-6)])
def test_predictor_variants_agree(precision, variant, tol):
    """The line-decomposed 3D order-5 predictor (VALU and matrix-core forms)
    and the FP32 space-time tile predictor agree with the generic node-column
    kernel."""
    n, order, steps = 4, 5, 2
    luh = P.density_wave_3d(n, order)
    outs = []
    for variant in (variant, 1):
        ctx = E.Context(3, order, 5, 0, precision=precision, picard_max=0 if precision == E.PREC_FP64 else 3)
        ctx.set_predictor_variant(variant)
        mesh = E.Mesh(ctx, n, 1.0 / n)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        mesh.run(steps)
        mesh.sync()
        outs.append((mesh.get_solution(), mesh.picard()))
        mesh.close()
        ctx.close()
    assert rel_l2(outs[0][0] - luh, outs[1][0] - luh) < tol * 1e3
    assert rel_l2(outs[0][0], outs[1][0]) < tol
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("picard", [1, 2, 4])
def test_fp32_fixed_picard_counts_agree(picard):
    """The FP32 matrix-core line predictor with a fixed Picard count (time
    averages formed by the time lanes: fluxes of iteration picard - 2's output,
    or of the initial guess for one iteration) against the FP32 node-column
    kernel with the same count."""
    n, order, steps = 3, 5, 2
    luh = P.density_wave_3d(n, order)
    outs = []
    for variant in (3, 1):
        ctx = E.Context(3, order, 5, 0, precision=E.PREC_FP32, picard_max=picard)
        ctx.set_predictor_variant(variant)
        mesh = E.Mesh(ctx, n, 1.0 / n)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        mesh.run(steps)
        mesh.sync()
        outs.append((mesh.get_solution(), mesh.picard()))
        mesh.close()
        ctx.close()
    assert rel_l2(outs[0][0] - luh, outs[1][0] - luh) < 2e-3
    assert rel_l2(outs[0][0], outs[1][0]) < 2e-6
    np.testing.assert_array_equal(outs[0][1], outs[1][1])

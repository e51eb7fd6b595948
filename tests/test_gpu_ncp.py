"""GPU parity of the user-PDE hooks with a non-conservative product and an
algebraic source (HIP library through the C ABI vs the oracle).

The compiled user PDE is the test PDE of include/exahype_amd_pde.h,
dQ/dt + div(a Q) + b . grad Q = -kappa Q per variable, whose oracle
restatement is pinned to the exact solution Q0(x - (a + b) t) exp(-kappa t)
(tests/test_oracle_ncp.py).  Checked here, FP64 bar 1e-12 relative L2:
* the per-cell fused predictor + volume integral (source and NCP terms in the
  Picard loop, spaceTimePredictorNonlinear.cpph:104-176, 323-367; W lShi in
  the volume integral, volumeIntegralNonlinear.cpph:204-222), 2D and 3D;
* the per-face Rusanov solve with NCP fluctuations
  (riemannSolverNonlinear.cpph:95-120);
* uniform periodic meshes stepped by the four-kernel form, with equal dt, and
  convergence of the GPU mesh to the exact solution;
* what is not compiled is refused: the fused two-kernel step, FP32, the CERK
  guess, more than 16 parameters.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12
A = (0.6, -0.4, 0.3)
BNCP = (0.5, 0.35, -0.45)
KAPPA = 0.8
PRM = list(A) + list(BNCP) + [KAPPA]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def q0(x):
    s = x.sum(-1)
    return np.stack([1.0 + 0.3 * np.sin(2 * np.pi * s + v) for v in range(5)], -1)


def user_ctx(dim, order, **kw):
    ctx = E.Context(dim, order, 5, 0, pde=E.PDE_USER, **kw)
    ctx.set_pde_params(PRM)
    return ctx


def oracle(dim, order):
    return O.Oracle(dim, order, pde=O.PDE_ADV_NCP_SRC, adv=A, ncp_b=BNCP, kappa=KAPPA)


@pytest.mark.parametrize("dim,order", [(3, 2), (3, 3), (3, 5), (2, 3), (2, 4)])
def test_cell_predictor_with_source_and_ncp_vs_oracle(dim, order):
    print(f"[ncp cell] dim={dim} order={order}", flush=True)
    rng = np.random.default_rng(10 * dim + order)
    B = order + 1
    luh = 1.0 + 0.3 * rng.normal(size=B ** dim * 5)
    h = [0.1] * dim
    dt = 0.2 * oracle(dim, order).stable_dt(luh, h)
    ctx = user_ctx(dim, order)
    r = ctx.fusedSpaceTimePredictorVolumeIntegral(luh, h, 0.0, dt)
    ctx.close()
    o = oracle(dim, order).fused(luh, h, dt)
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        assert rel_l2(r[k], o[k]) <= TOL, (k, rel_l2(r[k], o[k]))
    assert r["iters"] == o["iters"]
    # the source and NCP terms matter: without them the volume integral differs at O(1)
    plain = O.Oracle(dim, order, pde=O.PDE_ADV_NCP_SRC, adv=A).fused(luh, h, dt)
    assert rel_l2(plain["lduh"], o["lduh"]) > 1e-3


@pytest.mark.parametrize("dim", [2, 3])
def test_riemann_with_ncp_vs_oracle(dim):
    order = 3
    B = order + 1
    nf = B ** (dim - 1)
    rng = np.random.default_rng(dim)
    QL, QR = 1.0 + 0.2 * rng.normal(size=(2, nf * 5))
    FL0, FR0 = rng.normal(size=(2, nf * 5))
    ctx = user_ctx(dim, order)
    for d in range(dim):
        gl, gr = ctx.riemannSolver(FL0, FR0, QL, QR, d)
        ol, orr = oracle(dim, order).riemann(FL0, FR0, QL, QR, d)
        assert rel_l2(gl, ol) <= TOL and rel_l2(gr, orr) <= TOL
        assert rel_l2(gl, gr) > 1e-3  # fluctuations: the two sides differ
    ctx.close()


@pytest.mark.parametrize("dim,order,n,steps", [(3, 3, 3, 3), (3, 4, 4, 2), (2, 3, 6, 5)])
def test_mesh_with_source_and_ncp_vs_oracle(dim, order, n, steps):
    print(f"[ncp mesh] dim={dim} order={order} n={n}", flush=True)
    luh = np.ascontiguousarray(q0(P.node_points(n, order, dim)).ravel())
    ctx = user_ctx(dim, order)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    with pytest.raises(E.ExaError):
        mesh.set_fused(True)  # the fused corrector solves conservative problems only
    mesh.set_solution(luh)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got = mesh.get_solution()
    t, dtn = mesh.time()
    pic = mesh.picard()
    mesh.close()
    ctx.close()
    ref = luh.copy()
    dt_o, pic_o = oracle(dim, order).run_uniform(ref, n, 1.0 / n, steps, 0.0, want_picard=True)
    assert rel_l2(got, ref) <= TOL, rel_l2(got, ref)
    assert rel_l2(got - luh, ref - luh) <= 1e-10
    assert abs(dtn / dt_o - 1) <= 1e-12
    np.testing.assert_array_equal(pic, pic_o)


def test_gpu_mesh_converges_to_the_exact_solution():
    dim, order, T = 2, 3, 0.1
    errs = []
    for n in (6, 12):
        pts = P.node_points(n, order, dim)
        luh = np.ascontiguousarray(q0(pts).ravel())
        ctx = user_ctx(dim, order)
        mesh = E.Mesh(ctx, n, 1.0 / n)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        steps = int(np.ceil(T / mesh.time()[1]))
        for _ in range(steps):
            mesh.init_dt(T / steps)
            mesh.run(1)
        mesh.sync()
        got = mesh.get_solution().reshape(n ** dim, -1, 5)
        mesh.close()
        ctx.close()
        shift = np.array(A[:dim]) + np.array(BNCP[:dim])
        exact = q0(pts - shift * T) * np.exp(-KAPPA * T)
        w = np.array(O.tables(order)["w"])
        W = np.einsum("i,j->ij", w, w).ravel()
        errs.append(np.sqrt(np.sum((got - exact) ** 2 * W[None, :, None]) / n ** dim))
    rate = np.log(errs[0] / errs[1]) / np.log(2.0)
    assert rate >= order + 0.5, (errs, rate)


def test_unsupported_forms_are_refused():
    ctx = user_ctx(3, 3, precision=E.PREC_FP32, picard_max=2)
    with pytest.raises(E.ExaError):
        ctx.fusedSpaceTimePredictorVolumeIntegral(np.ones(4 ** 3 * 5), [0.1] * 3, 0.0, 1e-3)
    ctx.close()
    ctx = user_ctx(3, 3)
    ctx.set_initial_guess(E.GUESS_CERK)
    with pytest.raises(E.ExaError):
        ctx.fusedSpaceTimePredictorVolumeIntegral(np.ones(4 ** 3 * 5), [0.1] * 3, 0.0, 1e-3)
    with pytest.raises(E.ExaError):
        ctx.set_pde_params(np.ones(17))
    ctx.close()

"""Edge cases of the periodic mesh drivers on the GPU (through the C ABI):
a single cell (every face is the cell's own opposite face), two cells (both
neighbours of a cell are the same cell) and odd sizes, in both step forms and
precisions, against the FP64 oracle's periodic driver.  FP64: <= 1e-12
relative L2 with identical Picard counts; FP32 forms: <= 1e-6.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def state(D, n, order, seed):
    pts = P.node_points(n, order, D)
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 3, size=D)
    phase = (pts * k).sum(-1) * 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(phase)
    u = np.zeros(pts.shape[:-1] + (5,))
    u[..., 0] = rho
    u[..., 1] = rho * 0.6
    u[..., 2] = rho * (-0.3 if D >= 2 else 0.0)
    if D == 3:
        u[..., 3] = rho * 0.2
    vel2 = (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / rho ** 2
    u[..., 4] = 1.0 / 0.4 + 0.5 * rho * vel2
    return np.ascontiguousarray(u.ravel())


@pytest.mark.parametrize("D,order,n,fused", [(3, 5, 1, True), (3, 5, 1, False), (3, 3, 2, True), (3, 4, 5, False),
                                             (2, 3, 1, True), (2, 4, 2, False), (2, 3, 7, True)])
def test_small_and_odd_periodic_meshes_fp64(D, order, n, fused):
    print(f"[edge] D={D} order={order} n={n} fused={fused}", flush=True)
    luh = state(D, n, order, seed=n + order)
    steps = 3
    ctx = E.Context(D, order)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.set_fused(fused)
    mesh.set_solution(luh)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got, pic, (_, dtn) = mesh.get_solution(), mesh.picard(), mesh.time()
    mesh.close()
    ctx.close()
    ref = luh.copy()
    dt_o, pic_o = O.Oracle(D, order).run_uniform(ref, n, 1.0 / n, steps, 0.0, want_picard=True)
    assert rel_l2(got, ref) <= 1e-12, rel_l2(got, ref)
    assert rel_l2(got - luh, ref - luh) <= 1e-10
    np.testing.assert_array_equal(pic, pic_o)
    assert abs(dtn / dt_o - 1) <= 1e-12


@pytest.mark.parametrize("n", [1, 2, 5])
def test_small_periodic_meshes_fp32_headline_form(n):
    D, order, steps = 3, 5, 3
    luh = state(D, n, order, seed=3 * n)
    ctx = E.Context(D, order, precision=E.PREC_FP32, picard_max=1)
    ctx.set_initial_guess(E.GUESS_CERK)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.set_solution(luh)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got = mesh.get_solution()
    mesh.close()
    ctx.close()
    ref = luh.copy()
    O.Oracle(D, order).run_uniform(ref, n, 1.0 / n, steps, 0.0)
    g, r = got.reshape(-1, 5), ref.reshape(-1, 5)
    for v in range(5):
        assert rel_l2(g[:, v], r[:, v]) <= 1e-6, (v, rel_l2(g[:, v], r[:, v]))

"""GPU: partitioned stepping with halo messages (EXA_BC_EXCHANGE faces).

Two meshes on one GPU stand for two ranks of an n x n x 2n (2D: n x 2n)
domain split in the last direction: each step both predict, pack the
extrapolated-predictor and fluctuation messages of their faces in the split
direction, receive the other's (device-to-device copies in place of sends),
correct, take the global minimum of dt_next and advance (messages staged
through host memory, as MPI messages would be).  Checks: with the
same data on both halves the pair reproduces the periodic single mesh bit for
bit (fused and four-kernel step forms); with different data the union equals
the oracle's periodic box to the FP64 parity bar (walls in x included).
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def state(pts, D, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, 8)
    x, y = pts[..., 0], pts[..., 1]
    z = pts[..., 2] if D == 3 else 0.0 * x
    tp = 2 * np.pi
    rho = 1 + 0.1 * np.sin(tp * (x + c[0])) * np.cos(tp * (y + c[1])) + 0.05 * np.cos(tp * (z + c[2]))
    u = 0.3 * np.sin(tp * (y + c[3]))
    v = 0.2 * np.cos(tp * (x + c[4]))
    w = 0.1 * np.sin(tp * (z + x + c[5])) if D == 3 else 0.0 * x
    p = 1 + 0.1 * np.cos(tp * (x + y + c[6]))
    return np.ascontiguousarray(P.euler_state(rho, u, v, w, p))


def coupled_run(ctx, n, luh_a, luh_b, steps, bc_other, fused):
    D = ctx.dim
    lo, hi = 2 * (D - 1), 2 * (D - 1) + 1  # faces of the split direction
    kinds = list(bc_other) + [E.BC_EXCHANGE, E.BC_EXCHANGE]
    meshes = []
    for luh in (luh_a, luh_b):
        m = E.Mesh(ctx, n, 1.0 / n)
        m.set_fused(fused)
        m.set_boundary(kinds)
        m.set_solution(luh)
        m.init_dt(0.0)
        meshes.append(m)
    a, b = meshes
    dt0 = min(a.time()[1], b.time()[1])
    for m in meshes:  # the first dt is global too
        m.init_dt(dt0)
    for _ in range(steps):
        for m in meshes:
            m.step_predict()
        msg = {(i, f): m.halo_pack_host(f) for i, m in enumerate(meshes) for f in (lo, hi)}
        # a's top face meets b's bottom face and (periodic) b's top meets a's bottom
        a.halo_unpack_host(hi, *msg[(1, lo)])
        a.halo_unpack_host(lo, *msg[(1, hi)])
        b.halo_unpack_host(lo, *msg[(0, hi)])
        b.halo_unpack_host(hi, *msg[(0, lo)])
        for m in meshes:
            m.step_correct()
        dtn = min(a.dt_next(), b.dt_next())
        for m in meshes:
            m.set_dt_next(dtn)
            m.step_advance()
            m.sync()
    out = [m.get_solution() for m in meshes]
    t = [m.time() for m in meshes]
    for m in meshes:
        m.close()
    return out, t


@pytest.mark.parametrize("D,order,n,precision", [(3, 3, 4, E.PREC_FP64), (2, 4, 5, E.PREC_FP64),
                                                 (3, 5, 3, E.PREC_FP64), (3, 5, 3, E.PREC_FP32)])
def test_same_halves_reproduce_the_periodic_mesh(D, order, n, precision):
    """In the FP32 form the periodic mesh stores its extrapolated faces in FP32
    while exchange faces keep FP64 storage: equality pins that the FP32
    storage is exact."""
    print(f"[halo same] D={D} order={order} n={n} precision={precision}", flush=True)
    ctx = E.Context(D, order, precision=precision, picard_max=0 if precision == E.PREC_FP64 else 3)
    luh = state(P.node_points(n, order, D), D, 1)
    for fused in (True, False):
        # the single periodic mesh in the same step form (the two forms agree only to rounding)
        single = E.Mesh(ctx, n, 1.0 / n)
        try:
            single.set_fused(fused)
        except E.ExaError:
            assert fused  # fused corrector not compiled for this shape
            single.close()
            continue
        single.set_solution(luh)
        single.init_dt(0.0)
        single.run(3)
        single.sync()
        ref = single.get_solution()
        t_ref = single.time()
        single.close()
        (a, b), t = coupled_run(ctx, n, luh, luh, 3, [E.BC_PERIODIC] * (2 * (D - 1)), fused)
        assert np.array_equal(a, ref) and np.array_equal(b, ref), fused
        assert t[0] == t[1] == t_ref
    ctx.close()


@pytest.mark.parametrize("D,order,n,walls", [(3, 3, 4, False), (3, 2, 4, True), (2, 4, 5, True)])
def test_halves_match_the_oracle_box(D, order, n, walls):
    print(f"[halo box] D={D} order={order} n={n} walls={walls}", flush=True)
    B = order + 1
    o = O.Oracle(D, order)
    pts = P.node_points(n, order, D)
    ua, ub = state(pts, D, 2), state(pts, D, 3)
    other = [E.BC_WALL, E.BC_WALL] + [E.BC_PERIODIC] * (2 * (D - 2)) if walls else [E.BC_PERIODIC] * (2 * (D - 1))
    box = np.ascontiguousarray(np.concatenate([ua.reshape((n,) * D + (B ** D, 5)), ub.reshape((n,) * D + (B ** D, 5))],
                                              axis=0).reshape(-1, B ** D, 5))
    t_ref, _ = o.run_box(box, [n] * (D - 1) + [2 * n], 1.0 / n, 3, list(other) + [O.BC_PERIODIC] * 2)
    ref = box.reshape((2 * n,) + (n,) * (D - 1) + (B ** D, 5))
    ctx = E.Context(D, order)
    (a, b), t = coupled_run(ctx, n, ua, ub, 3, other, False)
    ctx.close()
    assert rel_l2(a, ref[:n]) <= 1e-12 and rel_l2(b, ref[n:]) <= 1e-12
    assert abs(t[0][0] - t_ref) <= 1e-14


def test_halo_pack_rejects_fp32_faces():
    """The all-periodic FP32 step stores its faces as FP32 [face][var][node]:
    packing FP64 messages from them is refused, not read as FP64."""
    ctx = E.Context(3, 5, precision=E.PREC_FP32, picard_max=3)
    m = E.Mesh(ctx, 3, 1.0 / 3)
    m.set_solution(P.density_wave_3d(3, 5))
    m.init_dt(0.0)
    m.step_predict()
    with pytest.raises(E.ExaError):
        m.halo_pack_host(0)
    m.close()
    ctx.close()

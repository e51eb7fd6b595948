"""Slab-pipelined fused step (GPU, through the C ABI).

exa_mesh_set_overlap(mesh, S) splits a periodic 3D mesh into S slabs of z
planes: the one-iteration FP32 CERK predictors run slab by slab on the mesh
stream, and the fused corrector of a slab runs on a second stream once the
predictors of its neighbour slabs (periodic) are done.  The kernels and their
per-cell operations are those of the one-launch step, so the bar is bitwise
equality with it: solution, dt of the next step, Picard counts and the Picard
guard's recomputed-cell counter, over several steps, for slab counts that
divide n and that do not, S = n (one plane per slab), with the guard firing on
a rough state (its list is recomputed slab by slab), and replayed from a
captured graph.  The arguments the mode refuses are checked too.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P

pytestmark = pytest.mark.gpu
ORDER, B, V = 5, 6, 5


def headline_ctx(guard=True):
    ctx = E.Context(3, ORDER, V, 0, precision=E.PREC_FP32, picard_max=1)
    ctx.set_initial_guess(E.GUESS_CERK)
    if not guard:
        ctx.set_picard_guard(0.0, 0)
    return ctx


def rough_state(n, seed=5):
    """Independent random nodal values (rho, p in [0.9, 1.1], |u_d| <= 0.2)."""
    rng = np.random.default_rng(seed)
    m = n ** 3 * B ** 3
    rho = rng.uniform(0.9, 1.1, m)
    u = rng.uniform(-0.2, 0.2, (m, 3))
    p = rng.uniform(0.9, 1.1, m)
    Q = np.stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2], p / 0.4 + 0.5 * rho * (u ** 2).sum(1)], 1)
    return np.ascontiguousarray(Q.ravel())


def run(ctx, n, steps, slabs, luh=None, graph=False):
    mesh = E.Mesh(ctx, n, 1.0 / n)
    if luh is None:
        mesh.adjust_solution(P.density_wave_params(3, 0, 8))
    else:
        mesh.set_solution(luh)
    mesh.init_dt(0.0)
    if slabs:
        mesh.set_overlap(slabs)
    snaps = []
    if graph:
        # one replay per call, the program's own synchronisation and download between replays
        for _ in range(steps):
            mesh.run(1, use_graph=True)
            mesh.sync()
            snaps.append(mesh.get_solution())
    else:
        mesh.run(steps)
        mesh.sync()
        snaps.append(mesh.get_solution())
    out = dict(sol=snaps, dt=mesh.time(), pic=mesh.picard(), rec=mesh.picard_guard_recomputed())
    mesh.close()
    return out


def assert_same(a, b):
    assert len(a["sol"]) == len(b["sol"])
    for x, y in zip(a["sol"], b["sol"]):
        assert np.array_equal(x, y)
    assert a["dt"] == b["dt"]
    assert np.array_equal(a["pic"], b["pic"])
    assert a["rec"] == b["rec"]


@pytest.mark.parametrize("n,slabs", [(9, 3), (9, 4), (10, 10), (7, 5)])
def test_slab_step_equals_one_launch_step(n, slabs):
    print(f"[overlap] n={n} slabs={slabs}", flush=True)
    ctx = headline_ctx()
    ref = run(ctx, n, 3, 0)
    got = run(ctx, n, 3, slabs)
    ctx.close()
    assert_same(got, ref)


def test_slab_step_with_the_guard_firing():
    n, slabs = 6, 3
    print(f"[overlap] rough state n={n} slabs={slabs}", flush=True)
    luh = rough_state(n)
    ctx = headline_ctx()
    ref = run(ctx, n, 2, 0, luh)
    got = run(ctx, n, 2, slabs, luh)
    ctx.close()
    assert ref["rec"] > 0, "the rough state must make the guard recompute cells"
    assert_same(got, ref)


def test_slab_step_from_a_captured_graph():
    n, slabs = 9, 3
    print(f"[overlap] graph n={n} slabs={slabs}", flush=True)
    ctx = headline_ctx()
    ref = run(ctx, n, 3, 0, graph=True)
    got = run(ctx, n, 3, slabs, graph=True)
    eager = run(ctx, n, 3, 0)
    ctx.close()
    assert_same(got, ref)
    assert np.array_equal(got["sol"][-1], eager["sol"][-1])


def test_overlap_arguments():
    ctx = headline_ctx()
    mesh = E.Mesh(ctx, 4, 0.25)
    for bad in (1, 2, 5):  # fewer than 3 slabs, more slabs than planes
        with pytest.raises(E.ExaError):
            mesh.set_overlap(bad)
    with pytest.raises(E.ExaError):
        mesh.set_overlap(-1)
    mesh.set_overlap(4)
    mesh.set_overlap(0)
    mesh.set_fused(False)
    with pytest.raises(E.ExaError):  # the four-kernel step has no slab form
        mesh.set_overlap(3)
    mesh.close()
    ctx.close()
    ctx64 = E.Context(3, ORDER)  # FP64 node-column predictor: whole-mesh launches only
    mesh = E.Mesh(ctx64, 4, 0.25)
    with pytest.raises(E.ExaError):
        mesh.set_overlap(3)
    mesh.close()
    ctx64.close()

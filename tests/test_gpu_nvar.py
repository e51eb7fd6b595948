"""Variable counts other than five (GPU vs the oracle, through the C ABI).

The reference's kernels take NumberOfVariables from the solver
(AbstractGenericADERDGSolverHeader.template:34-68); the library compiles
linear advection (F_d = a_d Q, any number of variables) for 1 variable in 2D
and 3D at order 3 and 3 variables in 3D at order 2, and a user PDE for its
declared kNVar.  Checked at the FP64 bar 1e-12 relative L2 with identical
Picard counts: the per-cell fused predictor + volume integral, the per-face
Rusanov solve, uniform periodic meshes in both step forms; the FP32 node-column
form within 1e-6; 3D order 5 with five advected variables (the node-column
kernel's general flux slab, which Euler no longer uses), both initial guesses;
and the host refuses Euler or the user PDE with another variable count.  The oracle's advection path is pinned to the exact space-time
solution (tests/test_oracle_dialect.py, CERK exactness for quadratic data).
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu
A = (0.7, -0.4, 0.3)


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def ctx(D, order, V, **kw):
    c = E.Context(D, order, nvar=V, pde=E.PDE_ADVECTION, **kw)
    c.set_advection(A[:D])
    return c


def oracle(D, order, V):
    return O.Oracle(D, order, nvar=V, pde=O.PDE_ADVECTION, adv=A)


def smooth(D, n, order, V):
    pts = P.node_points(n, order, D)
    s = pts.sum(-1)
    return np.ascontiguousarray(np.stack([1.0 + 0.3 * np.sin(2 * np.pi * s + 0.7 * v) for v in range(V)], -1).ravel())


SHAPES = [(2, 3, 1), (3, 3, 1), (3, 2, 3)]


@pytest.mark.parametrize("D,order,V", SHAPES)
def test_cell_predictor_and_riemann_vs_oracle(D, order, V):
    print(f"[nvar cell] D={D} order={order} V={V}", flush=True)
    B = order + 1
    rng = np.random.default_rng(D * 10 + V)
    luh = 1.0 + 0.3 * rng.normal(size=B ** D * V)
    h = [0.1, 0.12, 0.09][:D]
    o = oracle(D, order, V)
    dt = 0.3 * o.stable_dt(luh, h)
    c = ctx(D, order, V)
    r = c.fusedSpaceTimePredictorVolumeIntegral(luh, h, 0.0, dt)
    ref = o.fused(luh, h, dt)
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        assert rel_l2(r[k], ref[k]) <= 1e-12, (k, rel_l2(r[k], ref[k]))
    assert r["iters"] == ref["iters"]
    nf = B ** (D - 1)
    QL, QR = 1.0 + 0.2 * rng.normal(size=(2, nf * V))
    FL0, FR0 = rng.normal(size=(2, nf * V))
    for d in range(D):
        gl, gr = c.riemannSolver(FL0, FR0, QL, QR, d)
        ol, orr = o.riemann(FL0, FR0, QL, QR, d)
        assert rel_l2(gl, ol) <= 1e-12 and rel_l2(gr, orr) <= 1e-12
    c.close()


@pytest.mark.parametrize("D,order,V", SHAPES)
@pytest.mark.parametrize("fused", [True, False])
def test_mesh_vs_oracle(D, order, V, fused):
    n, steps = (6, 4) if D == 2 else (3, 3)
    luh = smooth(D, n, order, V)
    c = ctx(D, order, V)
    mesh = E.Mesh(c, n, 1.0 / n)
    mesh.set_fused(fused)
    mesh.set_solution(luh)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got, pic, (_, dtn) = mesh.get_solution(), mesh.picard(), mesh.time()
    mesh.close()
    c.close()
    ref = luh.copy()
    dt_o, pic_o = oracle(D, order, V).run_uniform(ref, n, 1.0 / n, steps, 0.0, want_picard=True)
    assert rel_l2(got, ref) <= 1e-12, rel_l2(got, ref)
    assert rel_l2(got - luh, ref - luh) <= 1e-10
    np.testing.assert_array_equal(pic, pic_o)
    assert abs(dtn / dt_o - 1) <= 1e-12


def test_fp32_node_column_form():
    D, order, V, n, steps = 3, 3, 1, 3, 3
    luh = smooth(D, n, order, V)
    c = ctx(D, order, V, precision=E.PREC_FP32, picard_max=4)
    mesh = E.Mesh(c, n, 1.0 / n)
    mesh.set_solution(luh)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got = mesh.get_solution()
    mesh.close()
    c.close()
    ref = luh.copy()
    oracle(D, order, V).run_uniform(ref, n, 1.0 / n, steps, 0.0)
    assert rel_l2(got, ref) <= 1e-6


@pytest.mark.parametrize("guess", ["trivial", "cerk"])
def test_order5_advection_five_variables_vs_oracle(guess):
    """3D order 5 with five variables and a PDE other than Euler: the FP64
    node-column kernel with its general flux slab (Euler runs the 12-slot
    form), both initial guesses, per-cell and on a periodic mesh."""
    D, order, V, n, steps = 3, 5, 5, 3, 2
    print(f"[order5 advection] guess={guess}", flush=True)
    dial = "cerk" if guess == "cerk" else None
    c = ctx(D, order, V)
    if guess == "cerk":
        c.set_initial_guess(E.GUESS_CERK)
    rng = np.random.default_rng(5)
    B = order + 1
    luh = 1.0 + 0.3 * rng.normal(size=B ** D * V)
    h = [0.1, 0.12, 0.09]
    o = O.Oracle(D, order, nvar=V, pde=O.PDE_ADVECTION, adv=A, dialect=dial)
    dt = 0.3 * o.stable_dt(luh, h)
    r = c.fusedSpaceTimePredictorVolumeIntegral(luh, h, 0.0, dt)
    ref = o.fused(luh, h, dt)
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        assert rel_l2(r[k], ref[k]) <= 1e-12, (k, rel_l2(r[k], ref[k]))
    assert r["iters"] == ref["iters"]
    u0 = smooth(D, n, order, V)
    mesh = E.Mesh(c, n, 1.0 / n)
    mesh.set_solution(u0)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got, pic, (_, dtn) = mesh.get_solution(), mesh.picard(), mesh.time()
    mesh.close()
    c.close()
    refm = u0.copy()
    dt_o, pic_o = O.Oracle(D, order, nvar=V, pde=O.PDE_ADVECTION, adv=A, dialect=dial).run_uniform(
        refm, n, 1.0 / n, steps, 0.0, want_picard=True)
    assert rel_l2(got, refm) <= 1e-12, rel_l2(got, refm)
    np.testing.assert_array_equal(pic, pic_o)
    assert abs(dtn / dt_o - 1) <= 1e-12


def test_variable_counts_are_checked():
    with pytest.raises(E.ExaError):
        E.Context(3, 3, nvar=1, pde=E.PDE_EULER)
    with pytest.raises(E.ExaError):
        E.Context(3, 3, nvar=1, pde=E.PDE_USER)
    with pytest.raises(E.ExaError):  # not compiled
        E.Context(3, 5, nvar=1, pde=E.PDE_ADVECTION)

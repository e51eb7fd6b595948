"""GPU parity of the optimised-kernel dialect (HIP library through the C ABI
vs the oracle).

* exa_opt_fusedSpaceTimePredictorVolumeIntegral (padded layouts, the
  optimised Picard rule, trivial or CERK initial guess) against the oracle's
  optimised dialect, 2D and 3D, several orders and vector widths;
* a uniform periodic mesh stepped with the CERK dialect against the oracle's
  mesh driver with the same dialect;
* the layout converters: round trips, untouched padding, and agreement with
  the optimised entry point's outputs.
FP64 bar: <= 1e-12 relative L2.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def euler_cell(rng, B, D, npar=0):
    n = B ** D
    rho = rng.uniform(0.9, 1.1, n)
    u = rng.uniform(-0.2, 0.2, (n, 3))
    if D == 2:
        u[:, 2] = 0
    p = rng.uniform(0.9, 1.1, n)
    E_ = p / 0.4 + 0.5 * rho * (u ** 2).sum(1)
    cols = [rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2], E_] + [rng.uniform(0.5, 1.5, n)] * npar
    return np.ascontiguousarray(np.stack(cols, 1).ravel())


def unpad(a, rows, n, npad):
    a = np.asarray(a).reshape(rows, npad)
    return a[:, :n].ravel(), a[:, n:]


DIALECTS = {E.DIALECT_OPTIMISED: "optimised", E.DIALECT_OPTIMISED_CERK: "optimised-cerk"}


@pytest.mark.parametrize("D,order,vect", [(3, 5, 8), (3, 3, 4), (3, 2, 1), (2, 4, 8), (2, 6, 2)])
@pytest.mark.parametrize("dialect", [E.DIALECT_OPTIMISED, E.DIALECT_OPTIMISED_CERK])
def test_opt_predictor_vs_oracle(D, order, vect, dialect):
    print(f"[dialect cell] D={D} order={order} vect={vect} dialect={dialect}", flush=True)
    rng = np.random.default_rng(100 * D + 10 * order + dialect)
    B, V, h = order + 1, 5, 0.05
    ctx = E.Context(D, order)
    ctx.set_dialect(dialect, vect)
    VP, NDP = ctx.padded_size(0), ctx.padded_size(1)
    assert VP == vect * ((V + vect - 1) // vect) and NDP == VP
    orc = O.Oracle(D, order, dialect=DIALECTS[dialect])
    for _ in range(3):
        luh = euler_cell(rng, B, D)
        dt = 0.7 * orc.stable_dt(luh, [h] * D)
        got = ctx.opt_fusedSpaceTimePredictorVolumeIntegral(luh, 1.0 / h, dt)
        ref = orc.fused(luh, [h] * D, dt)
        assert got["iters"] == ref["iters"]
        NS, NF = B ** D, B ** (D - 1)
        lduh, pad_d = unpad(got["lduh"], NS, V, VP)
        qb, pad_q = unpad(got["lQhbnd"], 2 * D * NF, V, NDP)
        fb, pad_f = unpad(got["lFhbnd"], 2 * D * NF, V, VP)
        assert rel_l2(lduh, ref["lduh"]) <= TOL
        assert rel_l2(qb, ref["lQhbnd"]) <= TOL
        assert rel_l2(fb, ref["lFhbnd"]) <= TOL
        for pad in (pad_d, pad_q, pad_f):
            assert (pad == 0).all()
    ctx.close()


@pytest.mark.parametrize("D,order,n,steps", [(3, 5, 3, 2), (3, 4, 4, 3), (2, 3, 9, 5)])
@pytest.mark.parametrize("dialect", [E.DIALECT_OPTIMISED, E.DIALECT_OPTIMISED_CERK])
def test_mesh_in_optimised_dialect_vs_oracle(D, order, n, steps, dialect):
    print(f"[dialect mesh] D={D} order={order} n={n} dialect={dialect}", flush=True)
    luh0 = P.density_wave_3d(n, order) if D == 3 else P.density_wave_2d(n, order)
    ref = luh0.copy()
    orc = O.Oracle(D, order, dialect=DIALECTS[dialect])
    dt_ref, pic_ref = orc.run_uniform(ref, n, 1.0 / n, steps, 0.0, want_picard=True)
    ctx = E.Context(D, order)
    ctx.set_dialect(dialect, 4)
    for fused in (True, False):
        mesh = E.Mesh(ctx, n, 1.0 / n)
        try:
            mesh.set_fused(fused)
        except E.ExaError:
            assert fused
            mesh.close()
            continue
        mesh.set_solution(luh0)
        mesh.init_dt(0.0)
        mesh.run(steps)
        mesh.sync()
        got = mesh.get_solution()
        pic = mesh.picard()
        t, dtn = mesh.time()
        mesh.close()
        assert rel_l2(got, ref) <= TOL, (fused, rel_l2(got, ref))
        assert rel_l2(got - luh0, ref - luh0) <= 1e-9
        np.testing.assert_array_equal(np.asarray(pic), np.asarray(pic_ref))
        assert abs(dtn / dt_ref - 1) <= 1e-12
    ctx.close()


SIZES = {  # rows of each converter's array and whether a row holds nData (else nVar) values
    "Qhbnd": lambda D, B: (2 * D * B ** (D - 1), True),
    "Fhbnd": lambda D, B: (2 * D * B ** (D - 1), False),
    "QFace": lambda D, B: (B ** (D - 1), True),
    "FFace": lambda D, B: (B ** (D - 1), False),
    "Qhi": lambda D, B: (B ** D, True),
    "Fhi": lambda D, B: ((D + 1) * B ** D, False),
    "PSi": lambda D, B: ((B + 1) * B ** D, False),
    "Lduh": lambda D, B: (B ** D, False),
}


@pytest.mark.parametrize("D,order,npar,vect", [(3, 3, 1, 8), (2, 3, 1, 4), (3, 5, 0, 2), (2, 4, 0, 1)])
def test_converters(D, order, npar, vect):
    rng = np.random.default_rng(D + order + vect)
    B = order + 1
    ctx = E.Context(D, order, 5, npar)
    ctx.set_dialect(E.DIALECT_OPTIMISED, vect)
    VP, NDP = ctx.padded_size(0), ctx.padded_size(1)
    for name in E.CONVERTERS:
        rows, data = SIZES[name](D, B)
        n = 5 + npar if data else 5
        npad = NDP if data else VP
        gen = rng.normal(size=rows * n)
        opt = ctx.convert(name, "generic2optimised", gen, rows * npad)
        body, pad = unpad(opt, rows, n, npad)
        np.testing.assert_array_equal(body, gen)
        assert np.isnan(pad).all()  # padding untouched
        opt_in = np.where(np.isnan(opt), np.inf, opt)  # padding must not be read
        back = ctx.convert(name, "optimised2generic", opt_in, rows * n)
        np.testing.assert_array_equal(back, gen)
    ctx.close()


def test_converters_agree_with_the_optimised_entry_point():
    D, order, vect = 3, 4, 8
    B = order + 1
    rng = np.random.default_rng(5)
    ctx = E.Context(D, order)
    ctx.set_dialect(E.DIALECT_OPTIMISED_CERK, vect)
    luh = euler_cell(rng, B, D)
    h = 0.05
    got = ctx.opt_fusedSpaceTimePredictorVolumeIntegral(luh, 1.0 / h, 1e-3)
    gen = ctx.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * D, 0.0, 1e-3)
    assert got["iters"] == gen["iters"]
    NS, NF = B ** D, B ** (D - 1)
    np.testing.assert_array_equal(ctx.convert("Lduh", "optimised2generic", got["lduh"], NS * 5), gen["lduh"])
    np.testing.assert_array_equal(ctx.convert("Qhbnd", "optimised2generic", got["lQhbnd"], 2 * D * NF * 5),
                                  gen["lQhbnd"])
    np.testing.assert_array_equal(ctx.convert("Fhbnd", "optimised2generic", got["lFhbnd"], 2 * D * NF * 5),
                                  gen["lFhbnd"])
    ctx.close()


def test_dialect_error_paths():
    ctx = E.Context(3, 3)
    with pytest.raises(E.ExaError):
        ctx.set_dialect(3, 1)
    with pytest.raises(E.ExaError):
        ctx.set_dialect(E.DIALECT_OPTIMISED, 3)
    with pytest.raises(E.ExaError):  # generic dialect: no optimised entry point
        ctx.opt_fusedSpaceTimePredictorVolumeIntegral(np.ones(64 * 5), 10.0, 1e-3)
    ctx.close()
    lin = E.Context(3, 2, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    with pytest.raises(E.ExaError):
        lin.set_dialect(E.DIALECT_OPTIMISED, 1)
    lin.close()


@pytest.mark.parametrize("D,order,n,steps", [(3, 5, 3, 3), (2, 4, 9, 4)])
def test_mesh_with_cerk_guess_and_generic_rule_vs_oracle(D, order, n, steps):
    """FP64 node-column kernel: CERK initial guess with the generic adaptive rule."""
    print(f"[cerk generic] D={D} order={order} n={n}", flush=True)
    luh0 = P.density_wave_3d(n, order) if D == 3 else P.density_wave_2d(n, order)
    ref = luh0.copy()
    orc = O.Oracle(D, order, dialect="cerk")
    dt_ref, pic_ref = orc.run_uniform(ref, n, 1.0 / n, steps, 0.0, want_picard=True)
    ctx = E.Context(D, order)
    ctx.set_initial_guess(E.GUESS_CERK)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.set_solution(luh0)
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got = mesh.get_solution()
    pic = mesh.picard()
    t, dtn = mesh.time()
    mesh.close()
    ctx.close()
    assert rel_l2(got, ref) <= TOL
    np.testing.assert_array_equal(np.asarray(pic), np.asarray(pic_ref))
    assert abs(dtn / dt_ref - 1) <= 1e-12


@pytest.mark.parametrize("picard", [1, 2, 3])
def test_fp32_line_kernel_with_cerk_guess(picard):
    """The FP32 matrix-core line kernel with the CERK guess and a fixed count
    agrees with the FP32 node-column kernel with the same guess and count, and
    stays within 1e-6 of the FP64 generic adaptive oracle."""
    n, order, steps = 3, 5, 3
    luh = P.density_wave_3d(n, order)
    outs = []
    for variant in (3, 1):
        ctx = E.Context(3, order, 5, 0, precision=E.PREC_FP32, picard_max=picard)
        ctx.set_predictor_variant(variant)
        ctx.set_initial_guess(E.GUESS_CERK)
        ctx.set_picard_guard(0.0)  # both kernels at the same fixed count (the guard has its own tests)
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
    ref = luh.copy()
    O.Oracle(3, order).run_uniform(ref, n, 1.0 / n, steps, 0.0)
    g, r = outs[0][0].reshape(-1, 5), ref.reshape(-1, 5)
    for v in range(5):
        assert rel_l2(g[:, v], r[:, v]) < 1e-6, (v, rel_l2(g[:, v], r[:, v]))


@pytest.mark.parametrize("D,order,n,steps,bc", [(3, 4, 3, 2, [1, 1, 0, 0, 2, 2]), (2, 3, 6, 4, [2, 2, 1, 1])])
def test_mesh_with_boundaries_and_cerk_guess_vs_oracle(D, order, n, steps, bc):
    """CERK guess (generic rule, FP64 node-column kernel) on a mesh with wall and
    Dirichlet faces: the boundary path and the guess together."""
    print(f"[cerk boundary] D={D} order={order} n={n} bc={bc}", flush=True)
    o = O.Oracle(D, order, dialect="cerk")
    pts = P.node_points(n, order, D)
    x = np.full(pts.shape[:-1] + (3,), 0.5)
    x[..., :D] = pts
    t0 = 0.8
    luh0 = np.ascontiguousarray(np.array([[o.entropy_wave(p, t0) for p in cell] for cell in x]))
    ref = luh0.copy()
    t_ref, dt_ref = o.run_bc(ref, n, 1.0 / n, steps, bc, t=t0)
    ctx = E.Context(D, order)
    ctx.set_initial_guess(E.GUESS_CERK)
    for fused in (True, False):
        mesh = E.Mesh(ctx, n, 1.0 / n)
        try:
            mesh.set_fused(fused)
        except E.ExaError:
            assert fused
            mesh.close()
            continue
        mesh.set_boundary(bc)
        mesh.set_solution(luh0)
        mesh.init_dt(0.0)
        mesh.set_time(t0)
        mesh.run(steps)
        mesh.sync()
        got = mesh.get_solution()
        t, dtn = mesh.time()
        mesh.close()
        assert rel_l2(got, ref) <= TOL, (fused, rel_l2(got, ref))
        assert abs(t - t_ref) <= 1e-14 and abs(dtn / dt_ref - 1) <= 1e-12
    ctx.close()


def test_fp32_cerk_line_kernel_with_walls_agrees_with_node_kernel():
    """FP32, CERK guess, one fixed iteration (the one-iteration line kernel with
    FP64 faces, since walls rule out the FP32-face step) against the FP32
    node-column kernel with the same guess and count."""
    n, order, steps = 3, 5, 2
    bc = [1, 1, 0, 0, 0, 0]
    o = O.Oracle(3, order)
    pts = P.node_points(n, order, 3)
    luh = np.ascontiguousarray(np.array([[o.entropy_wave(p, 0.8) for p in cell] for cell in pts])).ravel()
    outs = []
    for variant in (3, 1):
        ctx = E.Context(3, order, 5, 0, precision=E.PREC_FP32, picard_max=1)
        ctx.set_predictor_variant(variant)
        ctx.set_initial_guess(E.GUESS_CERK)
        ctx.set_picard_guard(0.0)  # both kernels at one fixed iteration
        mesh = E.Mesh(ctx, n, 1.0 / n)
        mesh.set_boundary(bc)
        mesh.set_solution(luh)
        mesh.init_dt(0.0)
        mesh.set_time(0.8)
        mesh.run(steps)
        mesh.sync()
        outs.append(mesh.get_solution())
        mesh.close()
        ctx.close()
    assert rel_l2(outs[0] - luh, outs[1] - luh) < 2e-3
    assert rel_l2(outs[0], outs[1]) < 2e-6

"""The linear (Cauchy-Kowalevski) oracle against the reference's golden vectors
and against closed forms.

Golden vectors:
* ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp, used by the
  active tests GenericEulerKernelTest3D.cpp:165-197 (volumeIntegralLinear),
  :233-275 (surfaceIntegralLinear) and GenericEulerKernelTest2D.cpp:193-265,
  :623-679 (same, 2D).
* ExaHyPE/exahype/tests/kernels/testdata/elasticity_testdata.cpp, used by
  ElasticityKernelTest2D.cpp:229-480 (spaceTimePredictorLinear with the NCP
  only, riemannSolverLinear with the flux only, volume and surface
  integrals; order 4, 9 variables, 3 parameters).  The reference runs these
  with validate(false); the vectors nevertheless agree with the generic
  kernels to the printed precision, so they pin the same code paths.

Closed forms (3D acoustics, MyLinearWaveSolver.cpp hooks):
* a polynomial plane wave p = P(n.x - c t), u = -n P/(rho c) of degree <= N is
  reproduced exactly by the Taylor series, so lQhi equals the exact time
  average, lduh equals W (Q(dt) - Q(0))/dt and lQbnd the exact face values;
* for jumps in (p, u_n) the fluctuations are the exact upwind ones
  A^- dQ and A^+ dQ of the normal Jacobian (eigenvalues +-c);
* point-source moments: PSi[n][k] W_k h^3 = f(t_n) phi_k(x_s), and the source
  terms of the Taylor levels are the time derivatives of the interpolating
  polynomial of f at the time nodes.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
EUL = json.load(open(os.path.join(HERE, "golden", "reference_testdata.json")))["data"]
ELA = json.load(open(os.path.join(HERE, "golden", "elasticity_testdata.json")))["data"]
EPS = 1e-10


# ---- generic Euler test vectors (Order 3, 5 variables, 1 parameter) ---------
def test_volume_integral_linear_3d_golden():
    # GenericEulerKernelTest3D.cpp:165-197
    o = O.Oracle(3, 3, nvar=5, npar=1)
    lFhi = np.zeros((4, 64, 5))
    lFhi[0, :, 1] = 1.0
    lFhi[1, :, 2] = 2.0
    lFhi[2, :, 3] = 3.0
    lduh = np.zeros(320)
    O.lib().orc_lin_volume(o.ps, O._p(lduh), O._p(lFhi.ravel()))
    np.testing.assert_allclose(lduh, EUL["dim3"]["testVolumeIntegralLinear"]["lduh"], atol=EPS, rtol=0)


def test_surface_integral_linear_3d_golden():
    # GenericEulerKernelTest3D.cpp:233-275
    o = O.Oracle(3, 3, nvar=5, npar=1)
    F = np.zeros((6, 16, 5))
    F[0:2, :, 1] = 1.0
    F[2:4, :, 2] = 1.0
    F[4:6, :, 3] = 1.0
    lduh = np.arange(320) / 10.0
    O.lib().orc_lin_surface(o.ps, O._p(lduh), O._p(F.ravel()), O._p(np.array([0.1] * 3)))
    np.testing.assert_allclose(lduh, EUL["dim3"]["testSurfaceIntegralLinear"]["lduh"], atol=EPS, rtol=0)


def test_surface_integral_linear_2d_golden():
    # GenericEulerKernelTest2D.cpp:193-265 (two cases)
    o = O.Oracle(2, 3, nvar=5, npar=1)
    F = np.zeros((4, 4, 5))
    F[0:2, :, 1] = 1.0
    F[2:4, :, 2] = 1.0
    lin = np.array(EUL["dim2"]["testSurfaceIntegral"]["lduh_in"])
    for case, Fin in ((1, F.ravel()), (2, np.array(EUL["dim2"]["testSurfaceIntegral"]["lFhbnd_in"]))):
        lduh = lin.copy()
        O.lib().orc_lin_surface(o.ps, O._p(lduh), O._p(np.ascontiguousarray(Fin)), O._p(np.array([0.1, 0.1])))
        np.testing.assert_allclose(lduh, EUL["dim2"]["testSurfaceIntegralLinear"][f"lduh_out_{case}"],
                                   atol=EPS, rtol=0)


def test_volume_integral_linear_2d_golden():
    # GenericEulerKernelTest2D.cpp:623-679 (two cases)
    o = O.Oracle(2, 3, nvar=5, npar=1)
    lF1 = np.zeros(3 * 80)
    lF1[:160] = EUL["dim2"]["testVolumeIntegral"]["lFhi"][:160]
    lF2 = np.zeros((3, 16, 5))
    lF2[0, :, 1] = 1.0
    lF2[1, :, 2] = 1.0
    for case, lF in ((1, lF1), (2, lF2.ravel())):
        lduh = np.zeros(80)
        O.lib().orc_lin_volume(o.ps, O._p(lduh), O._p(np.ascontiguousarray(lF)))
        np.testing.assert_allclose(lduh, EUL["dim2"]["testVolumeIntegralLinear"][f"lduh_{case}"], atol=EPS, rtol=0)


# ---- 2D elasticity test vectors (Order 4, 9 variables, 3 parameters) --------
EL = O.LinearOracle(2, 4, O.PDE_ELASTIC2D)
EDX = [38.4615384615385, 35.7142857142857]


def test_stp_linear_2d_elasticity_golden():
    # ElasticityKernelTest2D.cpp:308-410: NCP only, dt = 0.813172798364530
    g = ELA["testSpaceTimePredictorLinear"]
    luh = np.concatenate([np.array(g["luh_IN"]).reshape(25, 9), np.array(g["param_IN"]).reshape(25, 3)], 1)
    r = EL.stp(luh.ravel(), EDX, 0.813172798364530, flags=O.LIN_NCP)
    Qh = r["lQhi"].reshape(25, 12)[:, :9]
    np.testing.assert_allclose(Qh.ravel(), g["lQhi_OUT"], rtol=1e-12, atol=1e-12 * np.abs(g["lQhi_OUT"]).max())
    Fh = r["lFhi"].reshape(3, 25, 9)[:2]
    np.testing.assert_allclose(Fh.ravel(), g["lFhi_OUT"], rtol=1e-12, atol=1e-12 * np.abs(g["lFhi_OUT"]).max())


def test_stp_linear_2d_elasticity_extrapolation_golden():
    g = ELA["testSpaceTimePredictorLinear"]
    luh = np.concatenate([np.array(g["luh_IN"]).reshape(25, 9), np.array(g["param_IN"]).reshape(25, 3)], 1)
    r = EL.stp(luh.ravel(), EDX, 0.813172798364530, flags=O.LIN_NCP)
    Qb = r["lQbnd"].reshape(4, 5, 12)[:, :, :9]
    Fb = r["lFbnd"].reshape(4, 5, 9)
    qo = np.array(g["lQbnd_OUT"]).reshape(4, 5, 9)
    fo = np.array(g["lFbnd_OUT"]).reshape(4, 5, 9)
    np.testing.assert_allclose(Qb, qo, rtol=1e-12, atol=1e-12 * np.abs(qo).max())
    np.testing.assert_allclose(Fb, fo, rtol=1e-12, atol=1e-12 * np.abs(fo).max())


def test_riemann_linear_2d_elasticity_golden():
    # ElasticityKernelTest2D.cpp:229-306: flux only, direction 1; the vectors
    # carry ~10 significant digits
    g = ELA["testRiemannSolverLinear"]
    QL = np.concatenate([np.array(g["qL_IN"]).reshape(5, 9), np.array(g["paramL_IN"]).reshape(5, 3)], 1)
    QR = np.concatenate([np.array(g["qR_IN"]).reshape(5, 9), np.array(g["paramR_IN"]).reshape(5, 3)], 1)
    FL, FR = EL.riemann(QL.ravel(), QR.ravel(), 1, flags=O.LIN_FLUX)
    scale = np.abs(g["FL_OUT"]).max()
    np.testing.assert_allclose(FL, g["FL_OUT"], rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(FR, g["FR_OUT"], rtol=0, atol=1e-9 * scale)


def test_volume_integral_linear_2d_elasticity_golden():
    g = ELA["testVolumeIntegralLinear"]
    lF = np.zeros((3, 25, 9))
    lF[:2] = np.array(g["lFhi_IN"]).reshape(2, 25, 9)
    lduh = EL.volume_integral(lF.ravel())
    np.testing.assert_allclose(lduh, g["lduh_OUT"], rtol=0, atol=1e-12 * np.abs(g["lduh_OUT"]).max())


def test_surface_integral_linear_2d_elasticity_golden():
    g = ELA["testSurfaceIntegralLinear"]
    lduh = EL.surface_integral(g["lduh_IN"], g["lFbnd_IN"], EDX)
    np.testing.assert_allclose(lduh, g["lduh_OUT"], rtol=0, atol=1e-12 * np.abs(g["lduh_OUT"]).max())


# ---- acoustic closed forms -------------------------------------------------
RHO, C = 1.3, 0.8


def plane_wave(order, dirn, coeffs, h, centre):
    """Nodal plane-wave state of degree len(coeffs)-1 in the cell; returns
    (luh[B^3][6], P polynomial, n, node coordinates)."""
    o = O.LinearOracle(3, order)
    B = order + 1
    x = np.array(o.t["x"])
    n = np.asarray(dirn, float) / np.linalg.norm(dirn)
    P = np.polynomial.Polynomial(coeffs)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    pts = np.stack([centre[0] + h * (X - 0.5), centre[1] + h * (Y - 0.5), centre[2] + h * (Z - 0.5)], -1).reshape(-1, 3)
    s = pts @ n
    luh = np.zeros((B ** 3, 6))
    luh[:, 0] = P(s)
    luh[:, 1:4] = -np.outer(P(s), n) / (RHO * C)
    luh[:, 4], luh[:, 5] = RHO, C
    return o, luh, P, n, pts


def exact_average(P, n, pts, dt):
    """(1/dt) int_0^dt of the plane wave at pts: p and u."""
    Pi = P.integ()
    s = pts @ n
    pav = (Pi(s) - Pi(s - C * dt)) / (C * dt)
    return np.concatenate([pav[:, None], -np.outer(pav, n) / (RHO * C)], 1)


@pytest.mark.parametrize("order", [2, 4])
def test_acoustic_plane_wave_exact_time_average(order):
    h, dt, centre = 0.25, 0.03, (0.3, -0.1, 0.7)
    rng = np.random.default_rng(order)
    o, luh, P, n, pts = plane_wave(order, [1.0, -0.6, 0.35], rng.normal(size=order + 1), h, centre)
    r = o.stp(luh.ravel(), [h] * 3, dt, want_levels=True)
    Qh = r["lQhi"].reshape(-1, 6)
    ex = exact_average(P, n, pts, dt)
    np.testing.assert_allclose(Qh[:, :4], ex, rtol=0, atol=1e-12 * np.abs(ex).max())
    np.testing.assert_allclose(Qh[:, 4:], luh[:, 4:], rtol=0, atol=0)
    # volume integral: lduh = W (Q(dt) - Q(0)) / dt
    B = order + 1
    w = np.array(o.t["w"])
    W = np.einsum("i,j,k->ijk", w, w, w).ravel()
    lduh = o.volume_integral(r["lFhi"]).reshape(-1, 4)
    s1 = pts @ n - C * dt
    q1 = np.concatenate([P(s1)[:, None], -np.outer(P(s1), n) / (RHO * C)], 1)
    np.testing.assert_allclose(lduh, W[:, None] * (q1 - luh[:, :4]) / dt, rtol=0, atol=1e-11 * np.abs(q1).max())
    # extrapolated face values: exact time average at the face points
    x = np.array(o.t["x"])
    lQb = r["lQbnd"].reshape(6, B * B, 6)
    for f in range(6):
        d, side = f // 2, f % 2
        a, b = np.meshgrid(x, x, indexing="ij")  # (slow, fast) remaining coordinates
        ref = np.zeros((B * B, 3))
        others = [e for e in range(3) if e != d]
        ref[:, others[0]] = b.ravel()
        ref[:, others[1]] = a.ravel()
        ref[:, d] = side
        fp = np.array(centre) + h * (ref - 0.5)
        np.testing.assert_allclose(lQb[f][:, :4], exact_average(P, n, fp, dt), rtol=0,
                                   atol=1e-11 * np.abs(ex).max())


def test_acoustic_riemann_is_exact_upwind():
    o = O.LinearOracle(3, 3)
    B2 = 16
    rng = np.random.default_rng(3)
    lam, irho = RHO * C * C, 1.0 / RHO
    for d in range(3):
        QL = np.zeros((B2, 6))
        QR = np.zeros((B2, 6))
        QL[:, 4:] = QR[:, 4:] = (RHO, C)
        QL[:, :4] = rng.normal(size=(B2, 4))
        QR[:, :4] = QL[:, :4]
        dq = rng.normal(size=(B2, 2))
        QR[:, 0] += dq[:, 0]
        QR[:, 1 + d] += dq[:, 1]
        # normal Jacobian on (p, u_n) of p_t = lam div u, u_t = grad p / rho written
        # as Q_t + A Q_x = 0:  A = [[0, -lam], [-1/rho, 0]]  (eigenvalues +-c)
        A = np.array([[0.0, -lam], [-irho, 0.0]])
        ev, R = np.linalg.eig(A)
        np.testing.assert_allclose(sorted(ev), [-C, C], rtol=1e-14)
        Am = R @ np.diag(np.minimum(ev, 0)) @ np.linalg.inv(R)
        Ap = R @ np.diag(np.maximum(ev, 0)) @ np.linalg.inv(R)
        FL, FR = o.riemann(QL.ravel(), QR.ravel(), d)
        FL, FR = FL.reshape(B2, 4), FR.reshape(B2, 4)
        np.testing.assert_allclose(FL[:, [0, 1 + d]], dq @ Am.T, rtol=0, atol=1e-13)
        np.testing.assert_allclose(FR[:, [0, 1 + d]], dq @ Ap.T, rtol=0, atol=1e-13)
        t = [e for e in range(3) if e != d]
        np.testing.assert_allclose(FL[:, [1 + e for e in t]], 0.0, atol=1e-15)
        # continuous state: no fluctuation at all
        FL0, FR0 = o.riemann(QL.ravel(), QL.ravel(), d)
        assert np.abs(FL0).max() == 0.0 and np.abs(FR0).max() == 0.0


def lagrange_eval(x, k, s):
    others = np.delete(x, k)
    return np.prod((s - others) / (x[k] - others))


def test_point_source_moments():
    order = 4
    o = O.LinearOracle(3, order)
    B = order + 1
    x, w = np.array(o.t["x"]), np.array(o.t["w"])
    h, t, dt, f0, t0, amp = 0.2, 0.013, 0.004, 5.0, 0.05, 2.5
    centre = np.array([0.5, 0.3, 0.1])
    xs = centre + h * (np.array([0.31, 0.62, 0.17]) - 0.5)
    PSi = o.psi(centre, h, t, dt, xs, amp, f0, t0).reshape(B + 1, B, B, B, 4)
    assert np.abs(PSi[..., 1:]).max() == 0.0
    tn = np.concatenate([[t], t + dt * x])
    xr = np.array([0.31, 0.62, 0.17])
    for n in range(B + 1):
        ricker = (1 - 2 * (np.pi * f0 * (tn[n] - t0)) ** 2) * np.exp(-(np.pi * f0 * (tn[n] - t0)) ** 2)
        for i in range(B):
            for j in range(B):
                for k in range(B):
                    phi = lagrange_eval(x, k, xr[0]) * lagrange_eval(x, j, xr[1]) * lagrange_eval(x, i, xr[2])
                    val = PSi[n, i, j, k, 0] * w[i] * w[j] * w[k] * h ** 3
                    assert abs(val - amp * ricker * phi) <= 1e-12 * amp


def test_point_source_taylor_terms():
    """Levels with a source minus the source-free CK of the previous level are
    the time derivatives of the interpolant of the source at the time nodes."""
    order = 4
    o = O.LinearOracle(3, order)
    B = order + 1
    x, w = np.array(o.t["x"]), np.array(o.t["w"])
    h, t, dt, f0, t0, amp = 0.2, 0.013, 0.004, 5.0, 0.02, 1.5
    centre = np.array([0.1, 0.1, 0.1])
    xs = centre + h * (np.array([0.4, 0.55, 0.7]) - 0.5)
    PSi = o.psi(centre, h, t, dt, xs, amp, f0, t0)
    rng = np.random.default_rng(1)
    luh = np.zeros((B ** 3, 6))
    luh[:, :4] = rng.normal(size=(B ** 3, 4))
    luh[:, 4:] = (RHO, C)
    r = o.stp(luh.ravel(), [h] * 3, dt, PSi=PSi, want_levels=True)
    L = r["levels"].reshape(B + 1, B ** 3, 6)
    S = PSi.reshape(B + 1, B ** 3, 4)
    # level 0: luh + dt * sum_t w_t S(t_t)
    np.testing.assert_allclose(L[0, :, :4], luh[:, :4] + dt * np.einsum("t,tnv->nv", w, S[1:]), rtol=0,
                               atol=1e-12 * np.abs(L[0]).max())
    # source = f(t) * spatial factor; interpolate f in time through the B time nodes
    spatial = S[1] / (amp * ((1 - 2 * (np.pi * f0 * (t + dt * x[0] - t0)) ** 2)
                             * np.exp(-(np.pi * f0 * (t + dt * x[0] - t0)) ** 2)))
    ftimes = np.array([(1 - 2 * (np.pi * f0 * (t + dt * xx - t0)) ** 2) * np.exp(-(np.pi * f0 * (t + dt * xx - t0)) ** 2)
                       for xx in x]) * amp
    poly = np.polynomial.Polynomial.fit(dt * x, ftimes, B - 1)
    for i in range(1, B + 1):
        ck = o.stp(np.ascontiguousarray(L[i - 1]).ravel(), [h] * 3, dt, want_levels=True)["levels"]
        ck = ck.reshape(B + 1, B ** 3, 6)[1, :, :4]
        src = L[i, :, :4] - ck
        if i == 1:
            expect = S[0]
        elif i < B:
            expect = poly.deriv(i - 1)(0.0) * spatial
        else:
            expect = np.zeros_like(src)  # the series stops at level B-1 for sources
        np.testing.assert_allclose(src, expect, rtol=0, atol=1e-7 * max(np.abs(expect).max(), 1.0)
                                   + 1e-10 * np.abs(L[i]).max())


def test_acoustic_design_order_convergence():
    """Sine plane wave on periodic meshes n=4, 8: error ratio ~ 2^(N+1)."""
    order = 2
    errs = []
    for n in (4, 8):
        o = O.LinearOracle(3, order)
        B = order + 1
        h = 1.0 / n
        x = np.array(o.t["x"])
        kvec = 2 * np.pi * np.array([1.0, 0.0, 0.0])
        k = np.linalg.norm(kvec)
        nn = kvec / k
        cells = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        loc = np.stack([X, Y, Z], -1).reshape(-1, 3)
        pts = (cells[:, None, :] + loc[None]) * h
        s = pts @ kvec

        def state(tt):
            p = np.sin(s - C * k * tt)
            u = -p[..., None] * nn / (RHO * C)
            return np.concatenate([p[..., None], u, np.full(p.shape + (1,), RHO), np.full(p.shape + (1,), C)], -1)

        luh = np.ascontiguousarray(state(0.0))
        dt = 0.9 * O.PNPM[order] * h / (3 * C)
        T = 0.1
        steps = int(np.ceil(T / dt))
        dt = T / steps
        o.run(luh, n, h, dt, steps, None, np.zeros((1, 3)))
        ex = state(T)
        errs.append(np.sqrt(np.mean((luh[..., :4] - ex[..., :4]) ** 2)))
    rate = np.log2(errs[0] / errs[1])
    assert rate > order + 0.6, (errs, rate)


# ---- adjoint ----------------------------------------------------------------
def adjoint_problem(n=3, order=2, steps=6, seed=0):
    o = O.LinearOracle(3, order)
    B = order + 1
    h = 1.0 / n
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.8, 1.2, size=n ** 3)
    luh = np.zeros((n ** 3, B ** 3, 6))
    luh[..., 0] = 0.1 * rng.normal(size=(n ** 3, B ** 3))
    luh[..., 1:4] = 0.1 * rng.normal(size=(n ** 3, B ** 3, 3))
    luh[..., 4] = 1.0
    luh[..., 5] = c[:, None]
    dt = 0.9 * O.PNPM[order] * h / (3 * 1.2)
    src = (np.array([0.45, 0.52, 0.38]), 1.0, 5.0, 0.03)
    rec = rng.uniform(0.05, 0.95, size=(5, 3))
    obs = rng.normal(scale=0.01, size=(steps, 5))
    return o, luh, c, h, dt, steps, src, rec, obs


def test_adjoint_gradient_matches_finite_differences():
    o, luh, c, h, dt, steps, src, rec, obs = adjoint_problem()
    n = round(1 / h)
    J, g = o.misfit_grad(luh, n, h, dt, steps, src, rec, obs)
    rng = np.random.default_rng(7)
    for trial in range(2):
        dc = rng.normal(size=c.size)
        eps = 1e-5
        up, dn = luh.copy(), luh.copy()
        up[..., 5] = (c + eps * dc)[:, None]
        dn[..., 5] = (c - eps * dc)[:, None]
        Jp, _ = o.misfit_grad(up, n, h, dt, steps, src, rec, obs, want_grad=False)
        Jm, _ = o.misfit_grad(dn, n, h, dt, steps, src, rec, obs, want_grad=False)
        fd = (Jp - Jm) / (2 * eps)
        assert abs(fd - g @ dc) <= 1e-6 * abs(fd), (fd, g @ dc)
    # single-cell entries too
    for cell in (0, 13, 26):
        up, dn = luh.copy(), luh.copy()
        eps = 1e-5
        up[cell, :, 5] += eps
        dn[cell, :, 5] -= eps
        Jp, _ = o.misfit_grad(up, n, h, dt, steps, src, rec, obs, want_grad=False)
        Jm, _ = o.misfit_grad(dn, n, h, dt, steps, src, rec, obs, want_grad=False)
        fd = (Jp - Jm) / (2 * eps)
        assert abs(fd - g[cell]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (cell, fd, g[cell])

"""The split Cauchy-Kowalevski oracle (optimised linear kernels,
KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_linear_split_ck_cpp.template)
pinned by closed forms.  No reference test or fixture runs this template.

3D acoustics with the MyLinearWaveSolver.cpp hooks.  The template keeps one
time-derivative level and forms
  L_{k+1} = A L_k (+ source),   A: p <- lambda sum_d d_d u_d,  u_d <- d_d p
(the material matrix multiplies the flux part only, so the NCP term d_d p
enters without 1/rho), lQhi = sum_{k=0}^{B} dt^k/(k+1)! L_k (one level more
than the generic kernel), lFhi_d from lQhi, lduh = -W sum_d lFhi_d.

Time-average factors: the template starts dtavFac at dt/2 and multiplies by
dt/(t+2) after level t+1 with t counted from 0 (:196, :353), so level k >= 1
is weighted dt^k/(2 k!) -- dt^2/4, dt^3/12, ... -- where the Taylor series of
the time average has dt^k/(k+1)! (the non-split template counts from 1 and
gets dt^2/6).  The oracle restates the template as written; the tests below
pin those factors and show that the Taylor ones fail, and the split mesh is
second order in time.

For a state that is a tensor polynomial of degree <= N per direction, every
level is again such a polynomial and the nodal derivative operator is exact,
so the expected arrays follow from exact polynomial algebra (numpy.polynomial)
evaluated at the nodes -- no code shared with the oracle.  rho != 1 pins the
missing 1/rho, unequal dx pins the direction handling, mixed-degree data pins
the level count (the extra level dt^B/(B+1)! L_B is non-zero), and the source
terms are the derivatives on the unit interval of the interpolant of f at the
Gauss time nodes.
"""
import math

import numpy as np
import numpy.polynomial.polynomial as npp
import pytest

import oracle as O

RHO, C = 2.0, 1.3
LAM = RHO * C * C


def nodes_xyz(o, dx):
    x = np.array(o.t["x"])
    Z, Y, X = np.meshgrid(x * dx[2], x * dx[1], x * dx[0], indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def ev(c, X, Y, Z):
    return npp.polyval3d(X, Y, Z, c)


def A(st):
    """One split CK level of polynomial state (p, u, v, w) in physical coordinates."""
    p, u, v, w = st
    d = lambda c, a: npp.polyder(c, axis=a)
    div = pad_add(pad_add(d(u, 0), d(v, 1)), d(w, 2))
    return [LAM * div, d(p, 0), d(p, 1), d(p, 2)]


def pad_add(a, b, sa=1.0, sb=1.0):
    shape = tuple(max(x, y) for x, y in zip(a.shape, b.shape))
    out = np.zeros(shape)
    out[tuple(slice(0, s) for s in a.shape)] += sa * a
    out[tuple(slice(0, s) for s in b.shape)] += sb * b
    return out


def taylor(B, dt):
    """Taylor factors of the time average, dt^k/(k+1)!"""
    return [1.0] + [dt ** k / math.factorial(k + 1) for k in range(1, B + 1)]


def template_factors(B, dt):
    """The split template's factors (dt/2, then times dt/(t+2) from t = 0): dt^k/(2 k!)"""
    return [1.0] + [dt ** k / (2 * math.factorial(k)) for k in range(1, B + 1)]


def face_values(c, d, o, dx, xn):
    """Polynomial c on face (d, o), face-node order x-face (z, y), y-face (z, x), z-face (y, x)."""
    B = len(xn)
    a, b = np.meshgrid(xn, xn, indexing="ij")
    a, b = a.ravel(), b.ravel()
    s = dx[d] * o
    if d == 0:
        return ev(c, np.full_like(a, s), b * dx[1], a * dx[2])
    if d == 1:
        return ev(c, b * dx[0], np.full_like(a, s), a * dx[2])
    return ev(c, b * dx[0], a * dx[1], np.full_like(a, s))


def weights3(o):
    w = np.array(o.t["w"])
    return np.einsum("i,j,k->ijk", w, w, w).ravel()


def random_poly(rng, B):
    return rng.normal(size=(B, B, B)) * 0.5


@pytest.mark.parametrize("order", [2, 3, 4])
def test_split_ck_polynomial_closed_form(order):
    o = O.LinearOracle(3, order)
    B = order + 1
    rng = np.random.default_rng(order)
    dx = np.array([0.7, 1.1, 0.9])
    dt = 0.35
    st = [random_poly(rng, B) for _ in range(4)]
    X, Y, Z = nodes_xyz(o, dx)
    luh = np.zeros((B ** 3, 6))
    for v in range(4):
        luh[:, v] = ev(st[v], X, Y, Z)
    luh[:, 4:] = (RHO, C)
    r = o.stp_split(luh.ravel(), dx, dt)
    # expected lQhi = sum_{k=0}^{B} a_k A^k st with the template's factors
    a, at = template_factors(B, dt), taylor(B, dt)
    lev = st
    qh = [c.copy() for c in st]
    qt = [c.copy() for c in st]
    for k in range(1, B + 1):
        lev = A(lev)
        qh = [pad_add(qh[v], lev[v], 1.0, a[k]) for v in range(4)]
        qt = [pad_add(qt[v], lev[v], 1.0, at[k]) for v in range(4)]
    exp_qh = np.stack([ev(qh[v], X, Y, Z) for v in range(4)], -1)
    got = r["lQhi"].reshape(B ** 3, 6)
    scale = np.abs(exp_qh).max()
    np.testing.assert_allclose(got[:, :4], exp_qh, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(got[:, 4:], luh[:, 4:], rtol=0, atol=0)
    # the extra level matters: dropping a_B L_B moves lQhi by far more than the tolerance
    drop = np.stack([ev(lev[v], X, Y, Z) for v in range(4)], -1) * a[B]
    assert np.abs(drop).max() > 1e-6 * scale
    # and the template's factors are not the Taylor series'
    taylor_qh = np.stack([ev(qt[v], X, Y, Z) for v in range(4)], -1)
    assert np.abs(got[:, :4] - taylor_qh).max() > 1e-4 * scale
    # lduh = W A(lQhi)
    aq = A(qh)
    W = weights3(o)
    exp_du = np.stack([ev(aq[v], X, Y, Z) for v in range(4)], -1) * W[:, None]
    np.testing.assert_allclose(r["lduh"].reshape(B ** 3, 4), exp_du, rtol=0, atol=1e-10 * np.abs(exp_du).max())
    # faces: lQbnd = lQhi on the faces (+ parameters); lFbnd = lFhi_d on the faces of d,
    # lFhi_d[p] = -lambda d_d qh_{u_d}, lFhi_d[u_d] = -d_d qh_p
    xn = np.array(o.t["x"])
    F2 = B * B
    Qb = r["lQbnd"].reshape(6, F2, 6)
    Fb = r["lFbnd"].reshape(6, F2, 4)
    for d in range(3):
        fp = -LAM * npp.polyder(qh[1 + d], axis=d)
        fu = -npp.polyder(qh[0], axis=d)
        for o_ in range(2):
            f = 2 * d + o_
            for v in range(4):
                np.testing.assert_allclose(Qb[f, :, v], face_values(qh[v], d, o_, dx, xn), rtol=0,
                                           atol=1e-11 * scale)
            np.testing.assert_allclose(Qb[f, :, 4:], np.tile([RHO, C], (F2, 1)), rtol=1e-14)
            np.testing.assert_allclose(Fb[f, :, 0], face_values(fp, d, o_, dx, xn), rtol=0,
                                       atol=1e-10 * np.abs(exp_du).max() / W.min())
            np.testing.assert_allclose(Fb[f, :, 1 + d], face_values(fu, d, o_, dx, xn), rtol=0,
                                       atol=1e-10 * np.abs(exp_du).max() / W.min())
            others = [v for v in (1, 2, 3) if v != 1 + d]
            assert np.abs(Fb[f, :, others]).max() == 0.0


def test_split_ck_point_source_closed_form():
    order = 3
    o = O.LinearOracle(3, order)
    B = order + 1
    rng = np.random.default_rng(7)
    dx = np.array([0.5, 0.5, 0.5])
    dt = 0.2
    X, Y, Z = nodes_xyz(o, dx)
    g = random_poly(rng, B)
    f = rng.normal(size=B)  # source amplitude at the Gauss time nodes
    luh = np.zeros((B ** 3, 6))
    luh[:, 4:] = (RHO, C)
    PSi = np.zeros((B, B ** 3, 4))
    gn = ev(g, X, Y, Z)
    for t in range(B):
        PSi[t, :, 0] = gn * f[t]
    r = o.stp_split(luh.ravel(), dx, dt, PSi=PSi.ravel())
    xt, w = np.array(o.t["x"]), np.array(o.t["w"])
    interp = np.polynomial.Polynomial.fit(xt, f, B - 1).convert()  # on the unit time interval
    s = [interp.deriv(k)(0.0) if k else interp(0.0) for k in range(B)]
    zero = np.zeros((1, 1, 1))
    sig = dt * float(np.dot(w, f))
    lev = [sig * g, zero, zero, zero]
    a = template_factors(B, dt)
    qh = [c.copy() for c in lev]
    for k in range(B):
        nxt = A(lev)
        nxt[0] = pad_add(nxt[0], g, 1.0, s[k])
        lev = nxt
        qh = [pad_add(qh[v], lev[v], 1.0, a[k + 1]) for v in range(4)]
    exp_qh = np.stack([ev(qh[v], X, Y, Z) for v in range(4)], -1)
    got = r["lQhi"].reshape(B ** 3, 6)[:, :4]
    np.testing.assert_allclose(got, exp_qh, rtol=0, atol=1e-11 * np.abs(exp_qh).max())
    # lduh = W (A(lQhi) - g sum_t w_t f_t e_p)
    aq = A(qh)
    aq[0] = pad_add(aq[0], g, 1.0, -float(np.dot(w, f)))
    W = weights3(o)
    exp_du = np.stack([ev(aq[v], X, Y, Z) for v in range(4)], -1) * W[:, None]
    np.testing.assert_allclose(r["lduh"].reshape(B ** 3, 4), exp_du, rtol=0, atol=1e-10 * np.abs(exp_du).max())
    # with unit-interval source derivatives (the template applies no 1/dt): a
    # physical-time derivative s_k / dt^k changes lQhi far beyond the tolerance
    assert abs(s[1] * (1 / dt - 1)) * a[2] * np.abs(gn).max() > 1e-6 * np.abs(exp_qh).max()


def lagrange_eval(x, k, s):
    r = 1.0
    for m in range(len(x)):
        if m != k:
            r *= (s - x[m]) / (x[k] - x[m])
    return r


def test_split_ck_source_moments():
    order = 4
    o = O.LinearOracle(3, order)
    B = order + 1
    x, w = np.array(o.t["x"]), np.array(o.t["w"])
    h, t, dt, f0, t0, amp = 0.2, 0.013, 0.004, 5.0, 0.05, 2.5
    centre = np.array([0.5, 0.3, 0.1])
    xr = np.array([0.31, 0.62, 0.17])
    xs = centre + h * (xr - 0.5)
    PSi = o.psi_split(centre, h, t, dt, xs, amp, f0, t0).reshape(B, B, B, B, 4)
    assert np.abs(PSi[..., 1:]).max() == 0.0
    for n in range(B):
        tn = t + dt * x[n]
        ricker = (1 - 2 * (np.pi * f0 * (tn - t0)) ** 2) * np.exp(-(np.pi * f0 * (tn - t0)) ** 2)
        for i in range(B):
            for j in range(B):
                for k in range(B):
                    phi = lagrange_eval(x, k, xr[0]) * lagrange_eval(x, j, xr[1]) * lagrange_eval(x, i, xr[2])
                    val = PSi[n, i, j, k, 0] * w[i] * w[j] * w[k] * h ** 3
                    assert abs(val - amp * ricker * phi) <= 1e-12 * amp


def test_split_ck_minus_generic_is_the_extra_level():
    """rho = 1 (both kernels then solve the same system): lQhi(split) - lQhi(generic)
    = sum_k (a'_k - a_k) L_k + a'_B L_B over the generic kernel's levels L_k
    (a: Taylor factors, a': the template's)."""
    order = 3
    o = O.LinearOracle(3, order)
    B = order + 1
    rng = np.random.default_rng(3)
    luh = np.zeros((B ** 3, 6))
    luh[:, :4] = rng.normal(size=(B ** 3, 4))
    luh[:, 4:] = (1.0, C)
    dx, dt = [0.3] * 3, 0.05
    sp = o.stp_split(luh.ravel(), dx, dt)
    ge = o.stp(luh.ravel(), dx, dt, want_levels=True)
    L = ge["levels"].reshape(B + 1, B ** 3, 6)[:, :, :4]
    diff = sp["lQhi"].reshape(B ** 3, 6)[:, :4] - ge["lQhi"].reshape(B ** 3, 6)[:, :4]
    a, at = template_factors(B, dt), taylor(B, dt)
    expect = sum((a[k] - at[k]) * L[k] for k in range(1, B)) + a[B] * L[B]
    assert np.abs(expect).max() > 1e-9
    np.testing.assert_allclose(diff, expect, rtol=0, atol=1e-13 * np.abs(luh).max() + 1e-9 * np.abs(expect).max())


@pytest.mark.parametrize("split", [False, True])
def test_split_ck_mesh_order(split):
    """Periodic sine plane wave (rho = 1) on n = 4, 8 at fixed CFL, order 4: the
    generic predictor converges at the design order N + 1 (measured 5.1); the
    split predictor with the template's factors at second order (measured 2.2;
    its dt^2 weight is dt^2/4)."""
    order = 4
    errs = []
    for n in (4, 8):
        o = O.LinearOracle(3, order, split_ck=split)
        B = order + 1
        h = 1.0 / n
        x = np.array(o.t["x"])
        kvec = 2 * np.pi * np.array([1.0, 1.0, 0.0])
        k = np.linalg.norm(kvec)
        nn = kvec / k
        cells = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
        Zg, Yg, Xg = np.meshgrid(x, x, x, indexing="ij")
        loc = np.stack([Xg, Yg, Zg], -1).reshape(-1, 3)
        pts = (cells[:, None, :] + loc[None]) * h
        sph = pts @ kvec

        def state(tt):
            p = np.sin(sph - C * k * tt)
            u = -p[..., None] * nn / C
            return np.concatenate([p[..., None], u, np.ones(p.shape + (1,)), np.full(p.shape + (1,), C)], -1)

        luh = np.ascontiguousarray(state(0.0))
        dt = 0.9 * O.PNPM[order] * h / (3 * C)
        T = 0.1
        steps = int(np.ceil(T / dt))
        dt = T / steps
        o.run(luh, n, h, dt, steps, None, np.zeros((1, 3)))
        errs.append(np.sqrt(np.mean((luh[..., :4] - state(T)[..., :4]) ** 2)))
    rate = np.log2(errs[0] / errs[1])
    if split:
        assert 1.6 < rate < 2.6, (errs, rate)
    else:
        assert rate > order + 0.6, (errs, rate)


def test_split_ck_has_no_adjoint():
    o = O.LinearOracle(3, 1, split_ck=True)
    luh = np.zeros((8, 8, 6))
    luh[..., 4:] = 1.0
    J, _ = o.misfit_grad(luh, 2, 0.5, 0.01, 2, None, np.full((1, 3), 0.3), np.zeros((2, 1)))
    assert math.isnan(J)

"""The AMR oracle (oracle/aderdg_amr_oracle.c): projections and the static
two-level mesh step.

Pins:
* constant preservation, as the reference's testFaceUnknownsProjection /
  testVolumeUnknownsProjection (GenericEulerKernelTest3D.cpp:707-855) assert,
  here for level differences 1 and 2 (the reference's face test loop is empty);
* prolongation reproduces any polynomial of degree <= N exactly at the fine
  nodes (positions (s + x_m) / 3^delta), evaluated independently with numpy;
* restriction of the fine representations of such a polynomial over all
  subfaces / subcells returns the coarse nodal values, and for arbitrary data
  preserves the weighted integral (it is the L2 projection);
* the mesh step preserves a uniform state, conserves mass on the periodic
  mesh, and reduces to the uniform driver when nothing / everything is refined.
"""
import itertools

import numpy as np
import pytest

import oracle as O


def nodes(order):
    return np.array(O.tables(order)["x"])


def weights(order):
    return np.array(O.tables(order)["w"])


def poly_values(coef, pts):
    """Tensor-product polynomial sum c[i,j,..] prod x_k^i_k at pts[..., K]."""
    K = pts.shape[-1]
    out = np.zeros(pts.shape[:-1])
    for idx in itertools.product(*[range(coef.shape[k]) for k in range(K)]):
        term = coef[idx]
        for k in range(K):
            term = term * pts[..., k] ** idx[k]
        out = out + term
    return out


def grid(order, K, offset=None, scale=1.0):
    """Node coordinates of a K-dim tensor grid, first coordinate fastest: [B^K, K]."""
    x = nodes(order)
    B = len(x)
    idx = np.arange(B ** K)
    pts = np.stack([x[(idx // B ** k) % B] for k in range(K)], -1)
    if offset is not None:
        pts = (np.asarray(offset) + pts) * scale
    return pts


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("delta", [1, 2])
def test_constant_preservation(dim, delta):
    o = O.AmrOracle(dim, 3, nvar=1, npar=1)
    K = dim - 1
    nsub = 3 ** delta
    one_q, one_f = np.ones(o.nface * 2), np.ones(o.nface)
    accQ, accF = np.zeros(o.nface * 2), np.zeros(o.nface)
    for sub in itertools.product(range(nsub), repeat=K):
        q, f = o.face_prolongation(one_q, one_f, list(sub) + [0], 0, delta)
        np.testing.assert_allclose(q, 1.0, rtol=0, atol=1e-12)
        np.testing.assert_allclose(f, 1.0, rtol=0, atol=1e-12)
        accQ, accF = o.face_restriction(accQ, accF, one_q, one_f, list(sub) + [0], 0, delta)
    np.testing.assert_allclose(accQ, 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(accF, 1.0, rtol=0, atol=1e-12)
    acc = np.zeros(o.nspace * 2)
    for sub in itertools.product(range(nsub), repeat=dim):
        v = o.volume_prolongation(np.ones(o.nspace * 2), list(sub), 0, delta)
        np.testing.assert_allclose(v, 1.0, rtol=0, atol=1e-12)
        acc = o.volume_restriction(acc, np.ones(o.nspace * 2), list(sub), 0, delta)
    np.testing.assert_allclose(acc, 1.0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim,order,delta", [(2, 3, 1), (3, 2, 1), (3, 4, 1), (3, 3, 2), (2, 5, 2)])
def test_polynomial_reproduction_and_projection(dim, order, delta):
    o = O.AmrOracle(dim, order, nvar=1, npar=0)
    rng = np.random.default_rng(order + 10 * delta)
    nsub = 3 ** delta
    for K, nb, prolong, restrict in (
            (dim - 1, o.nface, lambda c, s: o.face_prolongation(c, c, s, 0, delta)[0],
             lambda acc, f, s: o.face_restriction(acc, acc, f, f, s, 0, delta)[0]),
            (dim, o.nspace, lambda c, s: o.volume_prolongation(c, s, 0, delta),
             lambda acc, f, s: o.volume_restriction(acc, f, s, 0, delta))):
        coef = rng.normal(size=(order + 1,) * K)
        C = poly_values(coef, grid(order, K))
        acc = np.zeros(nb)
        for sub in itertools.product(range(nsub), repeat=K):
            s = list(sub) + [0] * (3 - K)
            Fx = prolong(C, s)
            exact = poly_values(coef, grid(order, K, offset=np.array(sub, float), scale=1.0 / nsub))
            np.testing.assert_allclose(Fx, exact, rtol=0, atol=1e-11 * np.abs(exact).max())
            acc = restrict(acc, exact, s)
        np.testing.assert_allclose(acc, C, rtol=0, atol=1e-11 * np.abs(C).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_restriction_preserves_integral(dim):
    order = 3
    o = O.AmrOracle(dim, order, nvar=1, npar=0)
    w = weights(order)
    rng = np.random.default_rng(4)
    for K, nb, restrict in ((dim - 1, o.nface, lambda acc, f, s: o.face_restriction(acc, acc, f, f, s)[0]),
                            (dim, o.nspace, lambda acc, f, s: o.volume_restriction(acc, f, s))):
        W = np.ones(1)
        for _ in range(K):
            W = np.kron(w, W)
        acc = np.zeros(nb)
        total = 0.0
        for sub in itertools.product(range(3), repeat=K):
            f = rng.normal(size=nb)
            total += W @ f / 3 ** K
            acc = restrict(acc, f, list(sub) + [0] * (3 - K))
        assert abs(W @ acc - total) <= 1e-13 * max(1.0, abs(total))


# ---- the two-level mesh -----------------------------------------------------
def leaf_state(order, n, refined, fn):
    """Nodal values fn(x, y, z) -> Q[..., 5] at every leaf (reference layout)."""
    level, coords = O.AmrOracle.amr_leaves(n, refined)
    x = nodes(order)
    B = order + 1
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    loc = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    h = np.where(level == 0, 1.0 / n, 1.0 / (3 * n))
    pts = (coords[:, None, :] + loc[None]) * h[:, None, None]
    return np.ascontiguousarray(fn(pts)), level, h


def euler_wave(pts, amp=0.2):
    rho = 1.0 + amp * np.sin(2 * np.pi * (pts[..., 0] + pts[..., 1] + pts[..., 2]))
    u, v, w, p = 0.3, -0.2, 0.1, 1.0
    return np.stack([rho, rho * u, rho * v, rho * w, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)], -1)


def centre_refined(n):
    r = np.zeros(n ** 3, dtype=np.uint8)
    c = n // 2
    r[(c * n + c) * n + c] = 1
    return r


def test_amr_uniform_state_is_preserved():
    order, n = 2, 3
    refined = centre_refined(n)
    o = O.AmrOracle(3, order)
    luh, level, h = leaf_state(order, n, refined, lambda p: euler_wave(p, amp=0.0))
    ref = luh.copy()
    o.amr_run(luh, n, refined, 3)
    np.testing.assert_allclose(luh, ref, rtol=0, atol=1e-13)


def test_amr_conserves_mass():
    order, n = 2, 3
    refined = centre_refined(n)
    o = O.AmrOracle(3, order)
    luh, level, h = leaf_state(order, n, refined, euler_wave)
    w = weights(order)
    W = np.einsum("i,j,k->ijk", w, w, w).ravel()

    def totals(q):
        return np.einsum("lnv,n,l->v", q.reshape(len(h), -1, 5), W, h ** 3)

    m0 = totals(luh)
    o.amr_run(luh, n, refined, 4)
    m1 = totals(luh)
    np.testing.assert_allclose(m1, m0, rtol=1e-13, atol=1e-14)
    assert np.abs(luh - leaf_state(order, n, refined, euler_wave)[0]).max() > 1e-4  # it moved


def test_amr_reduces_to_uniform_driver():
    order, n = 2, 3
    o = O.AmrOracle(3, order)
    # nothing refined
    none = np.zeros(n ** 3, dtype=np.uint8)
    luh, _, _ = leaf_state(order, n, none, euler_wave)
    ref = luh.copy()
    dt = o.amr_run(luh, n, none, 2)
    dt_ref, _ = o.run_uniform(ref, n, 1.0 / n, 2)
    np.testing.assert_allclose(luh, ref, rtol=0, atol=1e-13)
    assert abs(dt - dt_ref) <= 1e-15 * dt_ref
    # everything refined == uniform mesh of 3n cells (leaves reordered)
    allr = np.ones(n ** 3, dtype=np.uint8)
    luh, level, _ = leaf_state(order, n, allr, euler_wave)
    _, coords = O.AmrOracle.amr_leaves(n, allr)
    N = 3 * n
    perm = np.argsort((coords[:, 2] * N + coords[:, 1]) * N + coords[:, 0])
    ref = np.ascontiguousarray(luh[perm])
    o.amr_run(luh, n, allr, 2)
    o.run_uniform(ref, N, 1.0 / N, 2)
    np.testing.assert_allclose(luh[perm], ref, rtol=0, atol=1e-13)

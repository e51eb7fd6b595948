"""GPU parity of the linear acoustic path and its discrete adjoint.

Forward: the HIP steps (through the C ABI) against the oracle's restatement of
the reference's linear kernels (oracle/aderdg_linear_oracle.c) on seeded
heterogeneous meshes with a point source and receivers: <= 1e-12 relative L2
(FP64 bar).  Adjoint: against the oracle's FP64 discrete adjoint (<= 1e-10
relative), independence of the checkpoint interval (bitwise), and central
finite differences of the misfit to 1e-6 -- at small sizes and at the full
workload (54^3 cells, order 4, 300 steps, 64 receivers, checkpoints every 20),
the last one through a torch.autograd.Function wrapper.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


_CTX = {}


def ctx(order):
    if order not in _CTX:
        _CTX[order] = E.Context(3, order, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    return _CTX[order]


def small_problem(n, order, seed, steps, with_source=True, nrec=5):
    rng = np.random.default_rng(seed)
    B3 = (order + 1) ** 3
    c = rng.uniform(0.8, 1.2, size=n ** 3)
    rho = rng.uniform(0.9, 1.4, size=n ** 3)
    u = 0.1 * rng.normal(size=(n ** 3, B3, 4))
    luh = P.acoustic_state(n, order, c, rho, u)
    h = 1.0 / n
    dt = 0.9 * O.PNPM[order] * h / (3 * 1.2)
    src = (np.array([0.37, 0.52, 0.61]), 1.0, 5.0, 0.02) if with_source else None
    rec = rng.uniform(0.05, 0.95, size=(nrec, 3))
    obs = rng.normal(scale=0.01, size=(steps, nrec))
    return luh, c, h, dt, src, rec, obs


def gpu_acoustic(order, n, h, luh, dt, src, rec):
    ac = E.Acoustic(ctx(order), n, h)
    ac.set_solution(luh)
    ac.set_dt(dt)
    if src is not None:
        ac.set_source(src[0], src[1], src[2], src[3])
    ac.set_receivers(rec)
    return ac


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("order,n,steps", [(1, 4, 6), (2, 4, 8), (3, 3, 6), (4, 4, 8), (4, 5, 3), (5, 3, 5)])
def test_forward_matches_oracle(order, n, steps, variant):
    luh, c, h, dt, src, rec, _ = small_problem(n, order, seed=order, steps=steps)
    o = O.LinearOracle(3, order)
    ref = luh.copy()
    rec_ref, _ = o.run(ref, n, h, dt, steps, src, rec)
    ac = gpu_acoustic(order, n, h, luh, dt, src, rec)
    ac.set_predictor(variant)
    r = ac.run(steps)
    out = ac.get_solution().reshape(ref.shape)
    assert rel_l2(out, ref) <= 1e-12, rel_l2(out, ref)
    assert rel_l2(r, rec_ref) <= 1e-12, rel_l2(r, rec_ref)
    s, t, dtg = ac.time()
    assert s == steps and dtg == dt


def test_forward_without_source_and_graph_replay():
    order, n, steps = 4, 4, 6
    luh, c, h, dt, _, rec, _ = small_problem(n, order, seed=11, steps=steps, with_source=False)
    o = O.LinearOracle(3, order)
    ref = luh.copy()
    rec_ref, _ = o.run(ref, n, h, dt, steps, None, rec)
    ac = gpu_acoustic(order, n, h, luh, dt, None, rec)
    r = ac.run(steps, use_graph=True)
    assert rel_l2(ac.get_solution(), ref) <= 1e-12
    assert rel_l2(r, rec_ref) <= 1e-12
    # eager and graph steps are the same kernels: identical results
    src = (np.array([0.37, 0.52, 0.61]), 1.0, 5.0, 0.02)
    a1 = gpu_acoustic(order, n, h, luh, dt, src, rec)
    a2 = gpu_acoustic(order, n, h, luh, dt, src, rec)
    r1 = np.concatenate([a1.run(3), a1.run(4)])
    r2 = np.concatenate([a2.run(3, use_graph=True), a2.run(4, use_graph=True)])
    assert np.array_equal(r1, r2)
    assert np.array_equal(a1.get_solution(), a2.get_solution())


def test_state_roundtrip_and_material_rules():
    order, n = 2, 3
    luh, c, h, dt, _, rec, _ = small_problem(n, order, seed=3, steps=1)
    ac = gpu_acoustic(order, n, h, luh, dt, None, rec)
    assert np.array_equal(ac.get_solution(), luh.ravel())
    rho, cc = ac.get_material()
    assert np.array_equal(cc, c)
    bad = luh.copy()
    bad[0, 1, 5] += 0.1  # c varying inside a cell
    with pytest.raises(E.ExaError) as ei:
        ac.set_solution(bad)
    assert ei.value.code == E.EXA_ERR_UNSUPPORTED
    bad = luh.copy()
    bad[1, :, 4] = -1.0
    with pytest.raises(E.ExaError) as ei:
        ac.set_solution(bad)
    assert ei.value.code == E.EXA_ERR_ARG
    with pytest.raises(E.ExaError):
        E.Acoustic(ctx(order), 0, h)
    with pytest.raises(E.ExaError):
        ac.misfit_gradient(2, np.zeros((2, len(rec))), checkpoint_every=0)


@pytest.mark.parametrize("order,n,steps,every", [(2, 3, 6, 2), (3, 4, 7, 3), (4, 3, 5, 5)])
def test_gradient_matches_oracle_adjoint(order, n, steps, every):
    luh, c, h, dt, src, rec, obs = small_problem(n, order, seed=20 + order, steps=steps)
    o = O.LinearOracle(3, order)
    J_ref, g_ref = o.misfit_grad(luh, n, h, dt, steps, src, rec, obs)
    ac = gpu_acoustic(order, n, h, luh, dt, src, rec)
    J, g = ac.misfit_gradient(steps, obs, checkpoint_every=every)
    assert abs(J - J_ref) <= 1e-12 * abs(J_ref)
    assert rel_l2(g, g_ref) <= 1e-10, rel_l2(g, g_ref)
    # the state afterwards is the forward state after `steps` steps
    ref = luh.copy()
    o.run(ref, n, h, dt, steps, src, rec)
    assert rel_l2(ac.get_solution(), ref) <= 1e-12


def test_gradient_independent_of_checkpoint_interval():
    order, n, steps = 3, 4, 9
    luh, c, h, dt, src, rec, obs = small_problem(n, order, seed=5, steps=steps)
    grads = []
    for every in (1, 4, 9, 20):
        ac = gpu_acoustic(order, n, h, luh, dt, src, rec)
        J, g = ac.misfit_gradient(steps, obs, checkpoint_every=every)
        grads.append((J, g))
    for J, g in grads[1:]:
        assert J == grads[0][0]
        assert np.array_equal(g, grads[0][1])


def fd_directional(ac_factory, steps, obs, c, dc, eps):
    def J_of(cv):
        ac = ac_factory()
        ac.set_material(c=cv)
        rec = ac.run(steps)
        return 0.5 * np.sum((rec - obs) ** 2)
    return (J_of(c + eps * dc) - J_of(c - eps * dc)) / (2 * eps)


def test_gradient_matches_finite_differences_gpu():
    order, n, steps = 4, 6, 30
    luh, c, h, dt, src, rec, obs = small_problem(n, order, seed=9, steps=steps, nrec=8)
    ac = gpu_acoustic(order, n, h, luh, dt, src, rec)
    J, g = ac.misfit_gradient(steps, obs, checkpoint_every=7)
    rng = np.random.default_rng(1)
    factory = lambda: gpu_acoustic(order, n, h, luh, dt, src, rec)  # noqa: E731
    for trial in range(2):
        dc = rng.normal(size=c.size)
        fd = fd_directional(factory, steps, obs, c, dc, 1e-5)
        assert abs(fd - g @ dc) <= 1e-6 * abs(fd), (fd, g @ dc)


# ---- the full workload: 54^3, order 4, 300 steps, 64 receivers, every 20 ----
FULL_N, FULL_ORDER, FULL_STEPS, FULL_EVERY = 54, 4, 300, 20


def full_problem():
    n, order = FULL_N, FULL_ORDER
    c_true = P.acoustic_wave_speed(n, seed=0, scale=1.0)
    c0 = P.acoustic_wave_speed(n, seed=0, scale=0.5)
    rec = P.acoustic_receivers(seed=0, nrec=64)

    def make(c):
        ac = E.Acoustic(ctx(order), n, 1.0 / n)
        ac.set_material(rho=np.ones(n ** 3), c=c)
        ac.set_dt(0.0)  # CFL of the true model
        ac.set_source(np.array(P.ACOUSTIC_SOURCE), P.ACOUSTIC_AMP, P.ACOUSTIC_F0, P.ACOUSTIC_T0)
        ac.set_receivers(rec)
        return ac

    truth = make(c_true)
    dt = truth.time()[2]
    obs = truth.run(FULL_STEPS)
    truth.close()

    def make_fixed(c):
        ac = make(c)
        ac.set_dt(dt)
        return ac

    return c0, obs, make_fixed, dt


def test_full_workload_gradient_torch_autograd_vs_finite_differences():
    torch = pytest.importorskip("torch")
    c0, obs, make, dt = full_problem()
    assert np.abs(obs).max() > 0  # the receivers see the source

    class AcousticMisfit(torch.autograd.Function):
        """J(c) through the C ABI; backward returns the adjoint gradient."""

        @staticmethod
        def forward(fctx, c):
            ac = make(c.detach().cpu().numpy())
            J, g = ac.misfit_gradient(FULL_STEPS, obs, checkpoint_every=FULL_EVERY)
            fctx.save_for_backward(torch.from_numpy(g))
            ac.close()
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(fctx, gJ):
            (g,) = fctx.saved_tensors
            return gJ * g

    c = torch.tensor(c0, dtype=torch.float64, requires_grad=True)
    J = AcousticMisfit.apply(c)
    J.backward()
    g = c.grad.numpy()
    assert J.item() > 0 and np.isfinite(g).all() and np.abs(g).max() > 0
    # smooth perturbation direction (a Gaussian around the source/receiver region)
    n = FULL_N
    ar = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(ar, ar, ar, indexing="ij")
    dc = np.exp(-((X - 0.55) ** 2 + (Y - 0.47) ** 2 + (Z - 0.5) ** 2) / (2 * 0.08 ** 2)).ravel()

    def J_of(cv):
        ac = make(cv)
        r = ac.run(FULL_STEPS)
        ac.close()
        return 0.5 * np.sum((r - obs) ** 2)

    eps = 1e-5
    fd = (J_of(c0 + eps * dc) - J_of(c0 - eps * dc)) / (2 * eps)
    assert abs(fd - g @ dc) <= 1e-6 * abs(fd), (fd, g @ dc)


# ---- per-cell compatibility entry points of the linear path -------------------
def cell_state(order, seed):
    rng = np.random.default_rng(seed)
    N = (order + 1) ** 3
    luh = np.zeros((N, 6))
    luh[:, :4] = rng.normal(size=(N, 4))
    luh[:, 4] = rng.uniform(0.8, 1.5, size=N)  # parameters may vary per node here
    luh[:, 5] = rng.uniform(0.7, 1.3, size=N)
    return luh.ravel()


@pytest.mark.parametrize("order", [1, 2, 4, 5])
def test_cell_predictor_matches_oracle(order):
    c = ctx(order)
    o = O.LinearOracle(3, order)
    luh = cell_state(order, order)
    dx = [0.1, 0.12, 0.09]
    dt = 0.004
    r = c.fusedSpaceTimePredictorVolumeIntegral(luh, dx, 0.0, dt, centre=[0.5, 0.5, 0.5])
    ref = o.stp(luh, dx, dt)
    assert rel_l2(r["lduh"], o.volume_integral(ref["lFhi"])) <= 1e-12
    assert rel_l2(r["lQhbnd"], ref["lQbnd"]) <= 1e-12
    assert rel_l2(r["lFhbnd"], ref["lFbnd"]) <= 1e-12
    # addVolumeIntegralResultToUpdate = false: luh += dt/W * volume integral
    r2 = c.fusedSpaceTimePredictorVolumeIntegral(luh, dx, 0.0, dt, addVolumeIntegralResultToUpdate=False,
                                                 centre=[0.5, 0.5, 0.5])
    assert rel_l2(r2["luh"], o.solution_update(luh, r["lduh"], dt)) <= 1e-13


def test_cell_predictor_with_point_source_matches_oracle():
    order = 4
    c = E.Context(3, order, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    o = O.LinearOracle(3, order)
    h, t, dt = 0.1, 0.013, 0.003
    centre = np.array([0.45, 0.55, 0.35])
    xs = centre + h * (np.array([0.3, 0.61, 0.47]) - 0.5)
    c.set_point_source(xs, 2.0, 5.0, 0.02)
    luh = cell_state(order, 3)
    r = c.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * 3, t, dt, centre=centre)
    PSi = o.psi(centre, h, t, dt, xs, 2.0, 5.0, 0.02)
    ref = o.stp(luh, [h] * 3, dt, PSi=PSi)
    assert rel_l2(r["lduh"], o.volume_integral(ref["lFhi"])) <= 1e-12
    assert rel_l2(r["lQhbnd"], ref["lQbnd"]) <= 1e-12
    assert rel_l2(r["lFhbnd"], ref["lFbnd"]) <= 1e-12
    # a cell that does not contain the source sees none
    far = c.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * 3, t, dt, centre=centre + 0.3)
    plain = o.stp(luh, [h] * 3, dt)
    assert rel_l2(far["lduh"], o.volume_integral(plain["lFhi"])) <= 1e-12


@pytest.mark.parametrize("order", [2, 4])
def test_cell_riemann_face_update_dt_match_oracle(order):
    c = ctx(order)
    o = O.LinearOracle(3, order)
    rng = np.random.default_rng(order + 30)
    F = (order + 1) ** 2
    QL = np.zeros((F, 6))
    QR = np.zeros((F, 6))
    QL[:, :4], QR[:, :4] = rng.normal(size=(F, 4)), rng.normal(size=(F, 4))
    QL[:, 4:], QR[:, 4:] = (1.1, 0.9), (1.3, 1.2)
    for d in range(3):
        FL, FR = c.riemannSolver(np.zeros(F * 4), np.zeros(F * 4), QL.ravel(), QR.ravel(), d)
        rFL, rFR = o.riemann(QL.ravel(), QR.ravel(), d)
        assert rel_l2(FL, rFL) <= 1e-13 and rel_l2(FR, rFR) <= 1e-13
    N = (order + 1) ** 3
    lduh = rng.normal(size=N * 4)
    face = rng.normal(size=F * 4)
    dx = [0.1, 0.12, 0.09]
    for d in range(3):
        for orient in range(2):
            got = c.faceIntegral(lduh, face, d, orient, dx)
            assert rel_l2(got, o.face_integral(lduh, face, d, orient, dx)) <= 1e-13
    lFb = rng.normal(size=6 * F * 4)
    assert rel_l2(c.surfaceIntegral(lduh, lFb, dx), o.surface_integral(lduh, lFb, dx)) <= 1e-13
    luh = cell_state(order, 5)
    dtv = c.stableTimeStepSize(luh, dx)
    cmax = np.abs(luh.reshape(-1, 6)[:, 5])
    expect = np.min(0.9 * O.PNPM[order] / (cmax / dx[0] + cmax / dx[1] + cmax / dx[2]))
    assert abs(dtv - expect) <= 1e-14 * expect

"""GPU parity of the split Cauchy-Kowalevski dialect (EXA_DIALECT_SPLIT_CK) of
the linear path, through the C ABI, against the oracle's restatement of
KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_linear_split_ck_cpp.template
(pinned by closed forms in tests/test_oracle_split_ck.py).  FP64 bar: 1e-12
relative L2.

* per-cell predictor + volume integral (lduh, lQhbnd, lFhbnd), orders 1-5,
  per-node material parameters, unequal dx; with a point source in the cell;
* acoustic meshes with a point source and receivers (split predictor, generic
  Riemann / face / update kernels), eager and graph-replayed steps;
* the split dialect differs from the generic one (one more Taylor level, the
  template's averaging factors, no 1/rho on the NCP);
* refused: the node-form mesh predictor, the adjoint, the dialect on a
  nonlinear context and the Picard dialects on a linear one.
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


def split_ctx(order):
    c = E.Context(3, order, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    c.set_dialect(E.DIALECT_SPLIT_CK)
    return c


def cell_state(order, seed):
    rng = np.random.default_rng(seed)
    N = (order + 1) ** 3
    luh = np.zeros((N, 6))
    luh[:, :4] = rng.normal(size=(N, 4))
    luh[:, 4] = rng.uniform(0.8, 1.5, size=N)
    luh[:, 5] = rng.uniform(0.7, 1.3, size=N)
    return luh.ravel()


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_cell_split_predictor_matches_oracle(order):
    print(f"[split cell] order={order}", flush=True)
    c = split_ctx(order)
    o = O.LinearOracle(3, order)
    luh = cell_state(order, 40 + order)
    dx, dt = [0.1, 0.12, 0.09], 0.004
    r = c.fusedSpaceTimePredictorVolumeIntegral(luh, dx, 0.0, dt, centre=[0.5, 0.5, 0.5])
    ref = o.stp_split(luh, dx, dt)
    for k, rk in (("lduh", "lduh"), ("lQhbnd", "lQbnd"), ("lFhbnd", "lFbnd")):
        assert rel_l2(r[k], ref[rk]) <= TOL, (k, rel_l2(r[k], ref[rk]))
    # not the generic kernel
    gen = o.stp(luh, dx, dt)
    assert rel_l2(o.volume_integral(gen["lFhi"]), ref["lduh"]) > 1e-6
    # addVolumeIntegralResultToUpdate = false: luh += dt/W * volume integral
    r2 = c.fusedSpaceTimePredictorVolumeIntegral(luh, dx, 0.0, dt, addVolumeIntegralResultToUpdate=False,
                                                 centre=[0.5, 0.5, 0.5])
    assert rel_l2(r2["luh"], o.solution_update(luh, ref["lduh"], dt)) <= 1e-13
    c.close()


def test_cell_split_predictor_with_point_source_matches_oracle():
    order = 4
    c = split_ctx(order)
    o = O.LinearOracle(3, order)
    h, t, dt = 0.1, 0.013, 0.003
    centre = np.array([0.45, 0.55, 0.35])
    xs = centre + h * (np.array([0.3, 0.61, 0.47]) - 0.5)
    c.set_point_source(xs, 2.0, 5.0, 0.02)
    luh = cell_state(order, 3)
    r = c.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * 3, t, dt, centre=centre)
    PSi = o.psi_split(centre, h, t, dt, xs, 2.0, 5.0, 0.02)
    ref = o.stp_split(luh, [h] * 3, dt, PSi=PSi)
    for k, rk in (("lduh", "lduh"), ("lQhbnd", "lQbnd"), ("lFhbnd", "lFbnd")):
        assert rel_l2(r[k], ref[rk]) <= TOL, (k, rel_l2(r[k], ref[rk]))
    plain = o.stp_split(luh, [h] * 3, dt)
    assert rel_l2(plain["lduh"], ref["lduh"]) > 1e-6  # the source matters
    far = c.fusedSpaceTimePredictorVolumeIntegral(luh, [h] * 3, t, dt, centre=centre + 0.3)
    assert rel_l2(far["lduh"], plain["lduh"]) <= TOL
    c.close()


def mesh_problem(n, order, seed):
    rng = np.random.default_rng(seed)
    B3 = (order + 1) ** 3
    cc = rng.uniform(0.8, 1.2, size=n ** 3)
    rho = rng.uniform(0.9, 1.4, size=n ** 3)
    u = 0.1 * rng.normal(size=(n ** 3, B3, 4))
    luh = P.acoustic_state(n, order, cc, rho, u)
    h = 1.0 / n
    dt = 0.9 * O.PNPM[order] * h / (3 * 1.2)
    src = (np.array([0.37, 0.52, 0.61]), 1.0, 5.0, 0.02)
    rec = rng.uniform(0.05, 0.95, size=(5, 3))
    return luh, h, dt, src, rec


@pytest.mark.parametrize("order,n,steps", [(1, 4, 6), (2, 4, 5), (3, 3, 6), (4, 4, 4), (5, 3, 3)])
def test_split_mesh_matches_oracle(order, n, steps):
    print(f"[split mesh] order={order} n={n} steps={steps}", flush=True)
    luh, h, dt, src, rec = mesh_problem(n, order, 60 + order)
    o = O.LinearOracle(3, order, split_ck=True)
    ref = luh.copy()
    rec_ref, _ = o.run(ref, n, h, dt, steps, src, rec)
    c = split_ctx(order)
    ac = E.Acoustic(c, n, h)
    ac.set_solution(luh)
    ac.set_dt(dt)
    ac.set_source(*src)
    ac.set_receivers(rec)
    r = ac.run(steps)
    out = ac.get_solution()
    assert rel_l2(out, ref) <= TOL, rel_l2(out, ref)
    assert rel_l2(r, rec_ref) <= TOL, rel_l2(r, rec_ref)
    # generic dialect on the same problem: a different scheme
    gen = luh.copy()
    O.LinearOracle(3, order).run(gen, n, h, dt, steps, src, rec)
    assert rel_l2(out - luh.ravel(), gen.ravel() - luh.ravel()) > 1e-8
    ac.close()
    c.close()


def test_split_mesh_graph_replay_equals_eager():
    order, n = 3, 4
    luh, h, dt, src, rec = mesh_problem(n, order, 7)
    c = split_ctx(order)
    runs = []
    for graph in (False, True):
        ac = E.Acoustic(c, n, h)
        ac.set_solution(luh)
        ac.set_dt(dt)
        ac.set_source(*src)
        ac.set_receivers(rec)
        r = np.concatenate([ac.run(3, use_graph=graph), ac.run(4, use_graph=graph)])
        runs.append((r, ac.get_solution()))
        ac.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    c.close()


def test_split_refusals():
    order, n = 2, 3
    luh, h, dt, src, rec = mesh_problem(n, order, 1)
    c = split_ctx(order)
    ac = E.Acoustic(c, n, h)
    ac.set_solution(luh)
    ac.set_dt(dt)
    ac.set_receivers(rec)
    with pytest.raises(E.ExaError) as ei:
        ac.set_predictor(1)
    assert ei.value.code == E.EXA_ERR_UNSUPPORTED
    with pytest.raises(E.ExaError) as ei:
        ac.misfit_gradient(2, np.zeros((2, len(rec))), checkpoint_every=1)
    assert ei.value.code == E.EXA_ERR_UNSUPPORTED
    ac.close()
    c.close()
    lin = E.Context(3, order, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    for d in (E.DIALECT_OPTIMISED, E.DIALECT_OPTIMISED_CERK):
        with pytest.raises(E.ExaError):
            lin.set_dialect(d)
    lin.close()
    nl = E.Context(3, 3, 5, 0)
    with pytest.raises(E.ExaError) as ei:
        nl.set_dialect(E.DIALECT_SPLIT_CK)
    assert ei.value.code == E.EXA_ERR_UNSUPPORTED
    nl.close()

"""Non-cubic cells through the per-cell entry points (GPU vs oracle).

The reference passes dx[D] to the predictor and the time-step size
(ExaHyPE/exahype/solvers/ADERDGSolver.h:2542-2553; stableTimeStepSize.cpph:14-62
sums max |lambda_d| / dx_d).  A cell with unequal edges runs the node-column
predictor with 1/dx per direction (stiffness, CERK divergences, volume
integral) and the stable dt with dx per direction.  FP64 bar 1e-12 relative L2
with equal Picard counts; the FP32 form within 1e-4 of the oracle (lduh of a
near-constant state cancels to ~2e-5 relative in FP32).
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def cell_state(D, order, seed):
    rng = np.random.default_rng(seed)
    B = order + 1
    n = B ** D
    rho = rng.uniform(0.9, 1.1, n)
    u = rng.uniform(-0.2, 0.2, (n, 3))
    if D == 2:
        u[:, 2] = 0
    p = rng.uniform(0.9, 1.1, n)
    Q = np.stack([rho, rho * u[:, 0], rho * u[:, 1], rho * u[:, 2], p / 0.4 + 0.5 * rho * (u ** 2).sum(1)], 1)
    return np.ascontiguousarray(Q.ravel())


@pytest.mark.parametrize("D,order,prec", [(3, 3, E.PREC_FP64), (3, 5, E.PREC_FP64), (2, 4, E.PREC_FP64),
                                          (3, 5, E.PREC_FP32)])
def test_noncubic_cell_predictor_vs_oracle(D, order, prec):
    dx = [0.1, 0.07, 0.13][:D]
    luh = cell_state(D, order, 10 * D + order)
    o = O.Oracle(D, order)
    dt = 0.5 * o.stable_dt(luh, dx)
    ref = o.fused(luh, dx, dt)
    ctx = E.Context(D, order, 5, 0, precision=prec, picard_max=0 if prec == E.PREC_FP64 else 3)
    got = ctx.fusedSpaceTimePredictorVolumeIntegral(luh, dx, 0.0, dt)
    ctx.close()
    tol = 1e-12 if prec == E.PREC_FP64 else 1e-4
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        assert rel_l2(got[k], ref[k]) <= tol, (k, rel_l2(got[k], ref[k]))
    if prec == E.PREC_FP64:
        assert got["iters"] == ref["iters"]
    # the anisotropy matters: the cubic cell with dx[0] differs
    cub = o.fused(luh, [dx[0]] * D, dt)
    assert rel_l2(cub["lduh"], ref["lduh"]) > 1e-3


@pytest.mark.parametrize("D", [2, 3])
def test_noncubic_stable_dt_vs_oracle(D):
    order = 3
    dx = [0.1, 0.07, 0.13][:D]
    luh = cell_state(D, order, D)
    ctx = E.Context(D, order)
    got = ctx.stableTimeStepSize(luh, dx)
    ctx.close()
    want = O.Oracle(D, order).stable_dt(luh, dx)
    assert abs(got / want - 1) <= 1e-14

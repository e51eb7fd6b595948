"""The CPU oracle against the reference's own golden vectors.

Fixtures: tests/golden/reference_testdata.json, extracted from
ExaHyPE/exahype/tests/kernels/testdata/generic_euler_testdata.cpp by
tests/golden/extract_reference_fixtures.py.  Test set-ups mirror
ExaHyPE/exahype/tests/kernels/c/GenericEulerKernelTest{2D,3D}.cpp
(NumberOfVariables=5, NumberOfParameters=1, Order=3; GenericEulerKernelTest.h:32-38).
The reference compares with eps = 1e-10 (GenericEulerKernelTest.h:96); the
same bound is used here.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_testdata.json")))["data"]
EPS = 1e-10
V, P, ND, B = 5, 1, 6, 4


def widen(a, nodes):
    """nVar-strided golden data -> nData-strided (parameter slot zero)."""
    out = np.zeros((nodes, ND))
    out[:, :V] = np.asarray(a).reshape(nodes, V)
    return out.ravel()


def narrow(a):
    return np.asarray(a).reshape(-1, ND)[:, :V].ravel()


@pytest.mark.parametrize("dim", [2, 3])
def test_pde_fluxes(dim):
    # GenericEulerKernelTest{2D:127-149, 3D:137-163}
    o = O.Oracle(dim, 3, nvar=5, npar=1)
    F = o.flux(np.array([1.0, 0.1, 0.2, 0.3, 3.5, 0.0]))
    g = GOLD[f"dim{dim}"]["testPDEFluxes"]
    for d, key in enumerate("fgh"[:dim]):
        np.testing.assert_allclose(F[d], g[key], atol=EPS, rtol=0)


def test_stp_nonlinear_3d():
    # GenericEulerKernelTest3D.cpp:584-705
    g = GOLD["dim3"]["testSpaceTimePredictorNonlinear"]
    o = O.Oracle(3, 3, nvar=5, npar=1)
    r = o.stp(widen(g["luh"], B ** 3), [0.05] * 3, 1.083937460199773e-3)
    np.testing.assert_allclose(narrow(r["lQi"]), g["lQi"], atol=EPS, rtol=0)
    lFi = r["lFi"].reshape(B, B, B, B, 4, V)[:, :, :, :, :3, :].ravel()
    np.testing.assert_allclose(lFi, g["lFi"], atol=EPS, rtol=0)
    np.testing.assert_allclose(narrow(r["lQhi"]), g["lQhi"], atol=EPS, rtol=0)
    np.testing.assert_allclose(r["lFhi"][: 3 * B ** 3 * V], g["lFhi"], atol=EPS, rtol=0)
    np.testing.assert_allclose(narrow(r["lQhbnd"]), g["lQhbnd"], atol=EPS, rtol=0)
    np.testing.assert_allclose(r["lFhbnd"], g["lFhbnd"], atol=EPS, rtol=0)
    # uniform input is already the Picard fixed point: 3D returns the 0-based break index
    assert r["iters"] == 0


def test_stp_nonlinear_2d():
    # GenericEulerKernelTest2D.cpp:876-985
    g = GOLD["dim2"]["testSpaceTimePredictorNonlinear"]
    o = O.Oracle(2, 3, nvar=5, npar=1)
    r = o.stp(widen(g["luh"], B ** 2), [0.05] * 2, 1.686854344081342e-3)
    np.testing.assert_allclose(narrow(r["lQi"]), g["lQi"], atol=EPS, rtol=0)
    lFi = r["lFi"].reshape(B, B, B, 3, V)[:, :, :, :2, :].ravel()
    np.testing.assert_allclose(lFi, g["lFi"], atol=EPS, rtol=0)
    np.testing.assert_allclose(narrow(r["lQhi"]), g["lQhi"], atol=EPS, rtol=0)
    np.testing.assert_allclose(r["lFhi"][: 2 * B ** 2 * V], g["lFhi"], atol=EPS, rtol=0)
    np.testing.assert_allclose(narrow(r["lQhbnd"]), g["lQhbnd"], atol=EPS, rtol=0)
    np.testing.assert_allclose(r["lFhbnd"], g["lFhbnd"], atol=EPS, rtol=0)
    # 2D convention on convergence: MaxIterations + 2 (c/2d/spaceTimePredictorNonlinear.cpph:301-306)
    assert r["iters"] == 2 * (3 + 1) + 2


def test_riemann_3d():
    # GenericEulerKernelTest3D.cpp:398-458
    g = GOLD["dim3"]
    o = O.Oracle(3, 3, nvar=5, npar=1)
    FL = np.arange(80) + 1.0
    FR = np.arange(80) - 1.0
    FL, FR = o.riemann(FL, FR, widen(g["testRiemannSolver"]["QL"], 16), widen(g["testRiemannSolver"]["QR"], 16), 1)
    np.testing.assert_allclose(FL, g["testRiemannSolverNonlinear"]["FL"], atol=EPS, rtol=0)
    np.testing.assert_allclose(FR, g["testRiemannSolverNonlinear"]["FR"], atol=EPS, rtol=0)


@pytest.mark.parametrize("case,direction", [(1, 0), (2, 1)])
def test_riemann_2d(case, direction):
    # GenericEulerKernelTest2D.cpp:512-618 (tests 2 and 3)
    R = GOLD["dim2"]["testRiemannSolverNonlinear"]
    o = O.Oracle(2, 3, nvar=5, npar=1)
    FL, FR = o.riemann(R[f"FL_{case}_in"], R[f"FR_{case}_in"], widen(R[f"QL_{case}_in"], B),
                       widen(R[f"QR_{case}_in"], B), direction)
    np.testing.assert_allclose(FL, R[f"F_{case}_out"], atol=EPS, rtol=0)
    np.testing.assert_allclose(FR, R[f"F_{case}_out"], atol=EPS, rtol=0)


def test_riemann_2d_symmetric_state():
    # GenericEulerKernelTest2D.cpp:474-509: equal states give FL == FR
    o = O.Oracle(2, 3, nvar=5, npar=1)
    q = np.tile([1.0, 0, 0, 0, 2.5], B)
    FL, FR = o.riemann(np.zeros(20), np.zeros(20), widen(q, B), widen(q, B), 0)
    np.testing.assert_array_equal(FL, FR)


def _seed_faces(dim):
    n = 80 if dim == 3 else 20
    lF = np.zeros((2 * dim, n))
    for i in range(0, n, 5):
        lF[0, i + 1] = lF[1, i + 1] = 1.0
        lF[2, i + 2] = lF[3, i + 2] = 1.0
        if dim == 3:
            lF[4, i + 3] = lF[5, i + 3] = 1.0
    return lF.ravel()


def test_surface_integral_3d():
    # GenericEulerKernelTest3D.cpp:277-319
    o = O.Oracle(3, 3, nvar=5, npar=1)
    out = o.surface_integral(np.arange(320) / 10.0, _seed_faces(3), [0.1] * 3)
    np.testing.assert_allclose(out, GOLD["dim3"]["testSurfaceIntegralNonlinear"]["lduh"], atol=EPS, rtol=0)


def test_surface_integral_2d():
    # GenericEulerKernelTest2D.cpp:266-339 (tests 1 and 2)
    o = O.Oracle(2, 3, nvar=5, npar=1)
    S = GOLD["dim2"]["testSurfaceIntegral"]
    SN = GOLD["dim2"]["testSurfaceIntegralNonlinear"]
    out = o.surface_integral(np.array(S["lduh_in"]), _seed_faces(2), [0.1] * 2)
    np.testing.assert_allclose(out, SN["lduh_1"], atol=EPS, rtol=0)
    out = o.surface_integral(np.array(S["lduh_in"]), np.array(S["lFhbnd_in"]), [0.1] * 2)
    np.testing.assert_allclose(out, SN["lduh_2"], atol=EPS, rtol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_face_integrals_sum_to_surface_integral(dim):
    # faceIntegralNonlinear over the 2D faces == surfaceIntegralNonlinear
    o = O.Oracle(dim, 3, nvar=5, npar=1)
    rng = np.random.default_rng(1)
    nf = B ** (dim - 1) * V
    lF = rng.standard_normal(2 * dim * nf)
    lduh0 = rng.standard_normal(B ** dim * V)
    dx = [0.1, 0.2, 0.3][:dim]
    ref = o.surface_integral(lduh0, lF, dx)
    out = lduh0.copy()
    for d in range(dim):
        for orient in range(2):
            out = o.face_integral(out, lF[(2 * d + orient) * nf:(2 * d + orient + 1) * nf], d, orient, dx)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-13)


def test_solution_update_3d():
    # GenericEulerKernelTest3D.cpp:467-503 (the reference compares only 16 nodes; all are checked)
    o = O.Oracle(3, 3, nvar=5, npar=1)
    luh = np.zeros((64, ND))
    luh[:, 0], luh[:, 4] = 1.0, 2.5
    out = o.solution_update(luh.ravel(), np.arange(320, dtype=float), 1.40831757919882352703e-03)
    np.testing.assert_allclose(narrow(out), GOLD["dim3"]["testSolutionUpdate"]["luh"], atol=EPS, rtol=0)


def test_solution_update_2d():
    # GenericEulerKernelTest2D.cpp:156-191
    o = O.Oracle(2, 3, nvar=5, npar=1)
    luh = np.zeros((16, ND))
    luh[:, 0], luh[:, 4] = 1.0, 2.5
    U = GOLD["dim2"]["testSolutionUpdate"]
    out = o.solution_update(luh.ravel(), np.array(U["lduh"]), 1.40831757919882352703e-03)
    np.testing.assert_allclose(narrow(out), U["luh"], atol=EPS, rtol=0)

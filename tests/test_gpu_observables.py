"""GPU: global observables of the mesh solution (exa_mesh_global_observables).

The reference's reset / map / merge cycle with the Euler example's
observables (dgcells = 1, fvcells = 0 per cell, merged by +=,
ApplicationExamples/Euler/Euler/EulerSolver_ADERDG.cpp:607-628) and per
quantity the domain integral (Gauss weights times h^dim), nodal minimum and
maximum.  Pinned by closed forms (a constant state integrates to its value
times the unit volume), by a host numpy reduction of the same nodal data
(min / max bit for bit), by mass conservation of periodic steps and by
bitwise reproducibility of the fixed-order reduction.
"""
import ctypes

import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P

pytestmark = pytest.mark.gpu


def host_observables(luh, order, D, n):
    w = np.array(E.tables(order)["w"])
    W = np.einsum("i,j,k->ijk", w, w, w).ravel() if D == 3 else np.einsum("i,j->ij", w, w).ravel()
    B = order + 1
    u = luh.reshape(n ** D, B ** D, 5)
    integral = np.einsum("cnk,n->k", u, W) * (1.0 / n) ** D
    return integral, u.reshape(-1, 5).min(0), u.reshape(-1, 5).max(0)


@pytest.mark.parametrize("D,order,n", [(3, 5, 6), (3, 3, 9), (2, 4, 12)])
def test_observables_match_host_reduction(D, order, n):
    print(f"[observables] D={D} order={order} n={n}", flush=True)
    ctx = E.Context(D, order)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    luh = P.density_wave_3d(n, order) if D == 3 else P.density_wave_2d(n, order)
    mesh.set_solution(luh)
    obs = mesh.global_observables()
    assert obs["dgcells"] == n ** D and obs["fvcells"] == 0
    integral, mn, mx = host_observables(luh, order, D, n)
    np.testing.assert_allclose(obs["integral"], integral, rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(obs["min"], mn)
    np.testing.assert_array_equal(obs["max"], mx)
    # fixed reduction order: bitwise reproducible
    again = mesh.global_observables()
    for k in ("integral", "min", "max"):
        np.testing.assert_array_equal(obs[k], again[k])
    # mass, momentum and energy of periodic steps are conserved
    mesh.init_dt(0.0)
    mesh.run(3)
    mesh.sync()
    after = mesh.global_observables()
    np.testing.assert_allclose(after["integral"], obs["integral"], rtol=1e-12, atol=1e-13)
    assert not np.array_equal(after["min"], obs["min"]) or not np.array_equal(after["max"], obs["max"])
    mesh.close()
    ctx.close()


def test_constant_state_closed_form():
    print("[observables] constant state", flush=True)
    D, order, n = 3, 4, 5
    ctx = E.Context(D, order)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    q = np.array([1.3, 0.2, -0.4, 0.1, 3.0])
    mesh.set_solution(np.tile(q, n ** D * (order + 1) ** D))
    obs = mesh.global_observables()
    assert obs["dgcells"] == n ** D
    np.testing.assert_allclose(obs["integral"], q, rtol=1e-13)  # unit volume
    np.testing.assert_array_equal(obs["min"], q)
    np.testing.assert_array_equal(obs["max"], q)
    # too small an output array is an argument error
    out = np.zeros(2 + 3 * 5 - 1)
    rc = E._lib.exa_mesh_global_observables(mesh.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size)
    assert rc != 0
    mesh.close()
    ctx.close()

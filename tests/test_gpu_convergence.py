"""Design order of convergence of the 3D step on the GPU (through the C ABI).

The exact solution is the density wave rho = 1 + 0.2 sin(2 pi (x + y + z - 3 t)),
u = v = w = 1, p = 1 (a smooth Euler solution); every mesh runs to T = 0.05
with a fixed dt = T / ceil(T / dt_CFL).  The error is the quadrature L2 error
of rho at the Gauss nodes, and the observed order between meshes n_i, n_j is
log(e_i / e_j) / log(n_j / n_i), the reference's convergence method
(Miscellaneous/ConvergenceAnalysis/README.md:9-29).  ADER-DG of order N is
of design order N + 1; the bar is N + 0.5 on the finest pair and N - 0.5 on
every pair (the coarsest meshes are pre-asymptotic).

* FP64 generic form (trivial guess, adaptive Picard), N = 3, 4 on 3^3, 9^3,
  27^3 cells; N = 5 on 3^3, 6^3, 9^3 cells: the absolute 1e-7 Picard
  tolerance leaves an error floor near 4e-10 at N = 5 and 27^3 cells (the
  iteration stops that far from its fixed point), so 27^3 would measure the
  stopping rule, not the discretisation.
* The headline FP32 form (CERK guess, one collapsed Picard iteration with the
  Picard guard; N = 5 on the x-line kernel, N = 3, 4 on the node-column
  kernel) until its FP32 error floor near 1.3e-7: N = 5 on 3^3, 4^3, 5^3
  cells, N = 4 on 4^3, 6^3, 9^3, N = 3 on 3^3, 6^3, 9^3.
"""
import math

import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P

pytestmark = pytest.mark.gpu
T = 0.05


def rho_error(luh, n, order, t):
    B = order + 1
    pts = P.node_points(n, order, 3)
    exact = 1.0 + 0.2 * np.sin(2 * np.pi * (pts.sum(-1) - 3.0 * t))
    w = np.array(E.tables(order)["w"])
    W = np.einsum("i,j,k->ijk", w, w, w).ravel()
    e = luh.reshape(n ** 3, B ** 3, 5)[..., 0] - exact
    return math.sqrt(np.sum(e ** 2 * W[None, :]) / n ** 3)


def run(n, order, precision, picard, guess):
    ctx = E.Context(3, order, 5, 0, precision=precision, picard_max=picard)
    ctx.set_initial_guess(guess)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.adjust_solution(P.density_wave_params(3, 0, 0))  # no Fourier modes: the exact wave
    mesh.init_dt(0.0)
    steps = int(math.ceil(T / mesh.time()[1]))
    dt = T / steps
    for _ in range(steps):
        mesh.init_dt(dt)
        mesh.run(1)
    mesh.sync()
    err = rho_error(mesh.get_solution(), n, order, T)
    mesh.close()
    ctx.close()
    return err


def orders(ns, errs):
    return [math.log(errs[i] / errs[i + 1]) / math.log(ns[i + 1] / ns[i]) for i in range(len(ns) - 1)]


@pytest.mark.parametrize("order,ns", [(3, (3, 9, 27)), (4, (3, 9, 27)), (5, (3, 6, 9))])
def test_fp64_design_order(order, ns):
    errs = [run(n, order, E.PREC_FP64, 0, E.GUESS_TRIVIAL) for n in ns]
    r = orders(ns, errs)
    print(f"[convergence fp64] N={order} n={ns} errors {['%.3e' % e for e in errs]} orders {['%.2f' % x for x in r]}")
    assert r[-1] >= order + 0.5, (errs, r)
    assert min(r) >= order - 0.5, (errs, r)


@pytest.mark.parametrize("order,ns", [(5, (3, 4, 5)), (4, (4, 6, 9)), (3, (3, 6, 9))])
def test_fp32_headline_form_design_order(order, ns):
    errs = [run(n, order, E.PREC_FP32, 1, E.GUESS_CERK) for n in ns]
    r = orders(ns, errs)
    print(f"[convergence fp32] N={order} n={ns} errors {['%.3e' % e for e in errs]} orders {['%.2f' % x for x in r]}")
    assert r[-1] >= order + 0.5, (errs, r)
    assert min(r) >= order - 0.5, (errs, r)

"""The headline FP32 form against the FP64 oracle on the bench workload's state
(GPU, through the C ABI).

The form the bench measures: FP32 line kernel, CERK initial guess, one Picard
iteration evaluated in collapsed form, the Picard guard on, FP32 faces in the
fused two-kernel step.  The state is the bench's density wave with 8 seeded
Fourier modes, generated on the device, at 18^3 cells (5832 cells, 5 steps) so
that the oracle (the reference's generic FP64 algorithm: trivial guess,
adaptive Picard) finishes in seconds.  Bars, per variable:
* state: relative L2 <= 1e-6 (the stated FP32 accuracy target; measured
  3.7e-8 .. 5.0e-8);
* update over the 5 steps (state - initial state): relative L2 <= 1e-4
  (measured 8.7e-6 .. 2.9e-5), i.e. the FP32 form resolves the change of the
  state, not only the background;
* dt of the next step within 1e-6 relative.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_fp32_headline_form_vs_oracle_on_the_bench_state():
    n, order, steps = 18, 5, 5
    ctx = E.Context(3, order, 5, 0, precision=E.PREC_FP32, picard_max=1)
    ctx.set_initial_guess(E.GUESS_CERK)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    mesh.adjust_solution(P.density_wave_params(3, 0, 8))
    mesh.sync()
    u0 = mesh.get_solution()
    mesh.init_dt(0.0)
    mesh.run(steps)
    mesh.sync()
    got = mesh.get_solution()
    _, dt_next = mesh.time()
    recomputed = mesh.picard_guard_recomputed()
    mesh.close()
    ctx.close()
    ref = u0.copy()
    dt_o, _ = O.Oracle(3, order).run_uniform(ref, n, 1.0 / n, steps, 0.0)
    g, r, s0 = got.reshape(-1, 5), ref.reshape(-1, 5), u0.reshape(-1, 5)
    es = [rel_l2(g[:, v], r[:, v]) for v in range(5)]
    eu = [rel_l2(g[:, v] - s0[:, v], r[:, v] - s0[:, v]) for v in range(5)]
    print(f"[fp32 headline] state {['%.2e' % e for e in es]} update {['%.2e' % e for e in eu]} "
          f"guard recomputed {recomputed}", flush=True)
    assert max(es) <= 1e-6, es
    assert max(eu) <= 1e-4, eu
    assert abs(dt_next / dt_o - 1) <= 1e-6

"""The BF16 experiment (EXA_PREC_BF16) on the GPU, through the C ABI.

The tile kernel's contractions take BF16 operands (operator entries and flux /
divergence values rounded to BF16, nearest even) on v_mfma_f32_16x16x16_bf16
with FP32 accumulation; flux evaluations, the iterate and the post-Picard phase
stay FP32.  BF16's 8-bit significand rounds each operand by up to 2^-9
relative, so the results sit near 1e-3 of the FP64 oracle rather than the FP32
form's 1e-7.  Stated bounds (measured: per cell 3.6e-4 .. 5.9e-4; meshes
1.6e-4 .. 5.5e-4 per variable): 2e-3 relative L2, and at least 1e-5 so that a
kernel silently running FP32 operands fails.
"""
import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P
import oracle as O

pytestmark = pytest.mark.gpu
ORDER = 5


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_bf16_cell_vs_oracle():
    cell = P.density_wave_3d(1, ORDER, modes=2)
    h, dt = [1.0 / 81] * 3, 1e-4
    o = O.Oracle(3, ORDER).fused(cell, h, dt)
    out = {}
    for prec in (E.PREC_BF16, E.PREC_FP32):
        ctx = E.Context(3, ORDER, 5, 0, precision=prec, picard_max=3)
        ctx.set_predictor_variant(4)
        out[prec] = ctx.fusedSpaceTimePredictorVolumeIntegral(cell, h, 0.0, dt)
        ctx.close()
    for k in ("lduh", "lQhbnd", "lFhbnd"):
        e16, e32 = rel_l2(out[E.PREC_BF16][k], o[k]), rel_l2(out[E.PREC_FP32][k], o[k])
        print(f"[bf16 cell] {k}: bf16 {e16:.2e}, fp32 tile {e32:.2e}")
        assert 1e-5 <= e16 <= 2e-3, (k, e16)
        assert e32 < 1e-4


@pytest.mark.parametrize("n,steps", [(9, 5), (27, 3)])
def test_bf16_mesh_vs_fp64(n, steps):
    prm = P.density_wave_params(3, 0, 8)

    def run(prec, picard):
        ctx = E.Context(3, ORDER, 5, 0, precision=prec, picard_max=picard)
        if prec != E.PREC_FP64:
            ctx.set_predictor_variant(4)
        mesh = E.Mesh(ctx, n, 1.0 / n)
        mesh.adjust_solution(prm)
        mesh.init_dt(0.0)
        mesh.run(steps)
        mesh.sync()
        out = mesh.get_solution()
        mesh.close()
        ctx.close()
        return out.reshape(-1, 5)

    ref = run(E.PREC_FP64, 0)
    got = run(E.PREC_BF16, 3)
    errs = [rel_l2(got[:, v], ref[:, v]) for v in range(5)]
    print(f"[bf16 mesh] n={n} steps={steps} per-variable rel L2 vs FP64: {['%.2e' % e for e in errs]}")
    assert all(1e-5 <= e <= 2e-3 for e in errs), errs


def test_bf16_argument_errors():
    with pytest.raises(E.ExaError):  # not compiled for this shape
        E.Context(3, 4, 5, 0, precision=E.PREC_BF16, picard_max=3)
    with pytest.raises(E.ExaError):  # BF16 iterates cannot meet the adaptive 1e-7 rule
        E.Context(3, ORDER, 5, 0, precision=E.PREC_BF16, picard_max=0)
    ctx = E.Context(3, ORDER, 5, 0, precision=E.PREC_BF16, picard_max=3)
    ctx.set_initial_guess(E.GUESS_CERK)  # the tile kernel has no CERK guess
    with pytest.raises(E.ExaError):
        ctx.fusedSpaceTimePredictorVolumeIntegral(P.density_wave_3d(1, ORDER), [0.01] * 3, 0.0, 1e-4)
    ctx.close()

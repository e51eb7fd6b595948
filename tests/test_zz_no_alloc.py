"""exa_batch_predictor_noalloc (GPU, through the C ABI): the allocation-free
variant of exa_batch_predictor, whose Picard-guard cell list lives in a static
device array.  It must give bit for bit what exa_batch_predictor gives, on two
calls in a row, the second with fewer cells (the shared array keeps the first
call's list), for the configuration that uses the list (FP32, CERK guess, one
guarded iteration, on a state where the guard flags some cells) and for one
that does not (FP64); more cells than the array holds is an argument error.
Device arrays come from the HIP runtime through ctypes.
"""
import ctypes

import numpy as np
import pytest

import exahype_amd as E
from exahype_amd import problems as P

pytestmark = pytest.mark.gpu

_hip = ctypes.CDLL("libamdhip64.so")
_H2D, _D2H = 1, 2
_vp = ctypes.c_void_p
_lib = E._lib
_lib.exa_batch_predictor_noalloc.argtypes = [_vp, _vp, ctypes.c_long] + [_vp] * 7
_lib.exa_batch_predictor_noalloc.restype = ctypes.c_int
_lib.exa_batch_predictor.argtypes = [_vp, _vp, ctypes.c_long] + [_vp] * 7
_lib.exa_batch_predictor.restype = ctypes.c_int
MAX_CELLS = 1 << 20


class Dev:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = _vp()
        assert _hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.host.nbytes, 8))) == 0
        assert _hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(_vp), ctypes.c_size_t(self.host.nbytes), _H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert _hip.hipDeviceSynchronize() == 0
        assert _hip.hipMemcpy(out.ctypes.data_as(_vp), self.ptr, ctypes.c_size_t(out.nbytes), _D2H) == 0
        return out

    def free(self):
        _hip.hipFree(self.ptr)


def noisy_state(n, seed):
    """Density wave + modes with nodal noise in rho of per-cell amplitude 10^u,
    u uniform in [-6, -1): some cells reach the guard's tolerance."""
    rng = np.random.default_rng(seed)
    luh = P.density_wave_3d(n, 5, modes=8).reshape(n ** 3, 216, 5).copy()
    luh[..., 0] += (10.0 ** rng.uniform(-6, -1, n ** 3))[:, None] * rng.normal(size=(n ** 3, 216))
    return np.ascontiguousarray(luh.ravel())


def one_call(ctx, fn, n, luh, dt):
    nc = n ** 3
    a = dict(luh=Dev(luh), h=Dev(np.full(nc, 1.0 / n)), dt=Dev(np.array([dt])),
             lduh=Dev(np.zeros(nc * 216 * 5)), qb=Dev(np.zeros(nc * 6 * 36 * 5)), fb=Dev(np.zeros(nc * 6 * 36 * 5)),
             pic=Dev(np.zeros(nc, dtype=np.int32)))
    rc = fn(ctx.h, None, nc, a["luh"].ptr, a["h"].ptr, a["dt"].ptr, a["lduh"].ptr, a["qb"].ptr, a["fb"].ptr,
            a["pic"].ptr)
    assert rc == E.EXA_OK, E._lib.exa_last_error()
    out = {k: a[k].get() for k in ("lduh", "qb", "fb", "pic")}
    for d in a.values():
        d.free()
    return out


@pytest.mark.parametrize("precision", [E.PREC_FP32, E.PREC_FP64])
def test_noalloc_equals_batch_predictor_twice(precision):
    ctx = E.Context(3, 5, 5, 0, precision=precision, picard_max=1 if precision == E.PREC_FP32 else 0)
    if precision == E.PREC_FP32:
        ctx.set_initial_guess(E.GUESS_CERK)
    for n, seed in ((9, 7), (4, 8)):  # the second call has fewer cells
        print(f"[noalloc] precision={precision} n={n}", flush=True)
        luh = noisy_state(n, seed)
        dt = 0.9 * 0.05 / n / 6.0
        ref = one_call(ctx, _lib.exa_batch_predictor, n, luh, dt)
        got = one_call(ctx, _lib.exa_batch_predictor_noalloc, n, luh, dt)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (n, k)
        if precision == E.PREC_FP32:
            assert 0 < np.sum(got["pic"] == 3) < n ** 3  # the list was used and not every cell flagged
    ctx.close()


def test_noalloc_rejects_more_cells_than_the_array_holds():
    ctx = E.Context(3, 5, 5, 0, precision=E.PREC_FP32, picard_max=1)
    ctx.set_initial_guess(E.GUESS_CERK)
    d = Dev(np.zeros(8))
    rc = _lib.exa_batch_predictor_noalloc(ctx.h, None, MAX_CELLS + 1, *([d.ptr] * 7))
    assert rc == E.EXA_ERR_ARG
    assert b"1048576" in E._lib.exa_last_error()
    d.free()
    ctx.close()

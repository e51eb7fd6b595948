"""CPU-side checks of the product library (no GPU needed).

* the library loads and exports every function include/exahype_amd.h declares
* the operator tables it derives from closed forms agree with the values the
  reference generator printed (ExaHyPE/kernels/generated/GaussLegendreBasis.csnippet,
  extracted to tests/golden/legendre_tables.json) and with identities the
  mathematics fixes
* argument validation returns error codes before touching a device
"""
import ctypes

import numpy as np
import pytest

import exahype_amd as E
import oracle as O


def test_library_exports_every_declared_symbol():
    names = E.declared_symbols()
    assert len(names) >= 25
    lib = ctypes.CDLL(E.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("order", list(range(0, 10)))
def test_tables_match_reference_generator(order):
    ours = E.tables(order)
    ref = O.tables(order)
    # values printed with ~64 significant digits; a long double derivation
    # rounded to double must agree to a few ulp
    for k in ("w", "x", "F0", "F1"):
        np.testing.assert_allclose(ours[k], ref[k], rtol=0, atol=4e-16 * max(1.0, np.abs(ref[k]).max()), err_msg=k)
    for k in ("Kxi", "iK1", "dudx", "equidistant", "fgp"):
        ok = "egp" if k == "equidistant" else k
        scale = max(1.0, np.abs(ref[k]).max())
        np.testing.assert_allclose(ours[ok], ref[k], rtol=0, atol=8e-16 * scale, err_msg=k)


@pytest.mark.parametrize("order", [1, 3, 5, 7])
def test_table_identities(order):
    t = E.tables(order)
    B = order + 1
    # sum of weights is 1; sum_n phi_n' = 0 => columns of Kxi sum to zero
    assert abs(t["w"].sum() - 1.0) < 1e-15
    np.testing.assert_allclose(t["Kxi"].sum(axis=0), 0.0, atol=1e-13)
    # K1 * 1 = F0 (integration by parts in time) => iK1 @ F0 = 1 (constant states are Picard fixed points)
    np.testing.assert_allclose(t["iK1"] @ t["F0"], np.ones(B), atol=1e-13)
    # dudx differentiates polynomials of degree <= N exactly
    x = t["x"]
    for deg in range(B):
        np.testing.assert_allclose(t["dudx"] @ x ** deg, deg * x ** max(deg - 1, 0) if deg else 0 * x, atol=1e-11)
    # Gauss quadrature with B nodes is exact to degree 2B-1 on [0,1]
    for deg in range(2 * B):
        assert abs(t["w"] @ x ** deg - 1.0 / (deg + 1)) < 1e-14
    # projector rows: coarse basis evaluated at fine nodes reproduces polynomials
    for sub in range(3):
        xf = (sub + x) / 3.0
        np.testing.assert_allclose(t["fgp"][sub].T @ x ** order, xf ** order, atol=1e-12)


def test_invalid_arguments_return_error_codes():
    h = ctypes.c_void_p()
    lib = E._lib
    assert lib.exa_ctx_create(ctypes.byref(h), 4, 3, 5, 0, 0, 0, 0, 0.9) == E.EXA_ERR_ARG
    assert "dim" in lib.exa_last_error().decode()
    assert lib.exa_ctx_create(ctypes.byref(h), 3, 3, 5, 0, 7, 0, 0, 0.9) == E.EXA_ERR_ARG
    assert lib.exa_ctx_create(ctypes.byref(h), 3, 3, 5, 0, 0, 0, 0, -1.0) == E.EXA_ERR_ARG
    assert lib.exa_ctx_create(ctypes.byref(h), 3, 12, 5, 0, 0, 0, 0, 0.9) == E.EXA_ERR_UNSUPPORTED
    assert lib.exa_ctx_create(None, 3, 3, 5, 0, 0, 0, 0, 0.9) == E.EXA_ERR_ARG
    assert lib.exa_tables(-1, *([None] * 9)) == E.EXA_ERR_ARG
    with pytest.raises(E.ExaError):
        E.tables(40)


def test_probe_ascii_format(tmp_path):
    """The probe writer follows the reference's probe::ascii layout
    (plotters/ADERDG2ProbeAscii.cpp: header, then plot time, time stamp, values)."""
    from exahype_amd.plotters import ProbeAscii

    class FakeMesh:
        def __init__(self):
            self.t = 0.0

        def time(self):
            return self.t, 0.01

        def probe(self, pts):
            return np.array([[1.0, 0.5, 1e-5, 0.0, 2.5 + self.t]])

    m = FakeMesh()
    w = ProbeAscii(str(tmp_path / "rho"), [0.3, 0.4, 0.5], 5)
    w.plot(m)
    m.t = 0.125
    w.plot(m)
    w.close()
    lines = (tmp_path / "rho.probe").read_text().splitlines()
    assert lines[0] == "# plot-time, real-time,Q0,Q1,Q2,Q3,Q4"
    assert lines[1] == "0, 0, 1, 0.5, 1e-05, 0, 2.5"
    assert lines[2] == "0.125, 0.125, 1, 0.5, 1e-05, 0, 2.625"

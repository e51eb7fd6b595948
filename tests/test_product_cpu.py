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


def _null_context(dim=3, order=3):
    """A Context whose handle is NULL: every entry point rejects it with a
    status code before touching the device, so calling the wrappers checks the
    ctypes argument conversion of each binding on the CPU."""
    c = E.Context.__new__(E.Context)
    c.h = None
    c.dim, c.order, c.nvar, c.npar = dim, order, 5, 0
    c.B = order + 1
    return c


@pytest.mark.parametrize("dim", [2, 3])
def test_context_bindings_convert_their_arguments(dim):
    c = _null_context(dim)
    B, NF, NS = c.B, c.nface, c.nspace
    z = np.zeros
    calls = [
        lambda: c.fusedSpaceTimePredictorVolumeIntegral(z(NS * 5), dx=[0.1] * dim, t=0.0, dt=0.01),
        lambda: c.riemannSolver(z(NF * 5), z(NF * 5), z(NF * 5), z(NF * 5), 0, dx=[0.1] * dim),
        lambda: c.faceIntegral(z(NS * 5), z(NF * 5), 0, 1, [0.1] * dim),
        lambda: c.faceIntegral(z(NS * 5), z(NF * 5), 0, 1, [0.1] * dim, subfaceIndex=[0] * (dim - 1), levelDelta=1),
        lambda: c.faceUnknownsProlongation(z(NF * 5), z(NF * 5), [0] * (dim - 1)),
        lambda: c.faceUnknownsRestriction(z(NF * 5), z(NF * 5), z(NF * 5), z(NF * 5), [0] * (dim - 1)),
        lambda: c.volumeUnknownsProlongation(z(NS * 5), [0] * dim),
        lambda: c.volumeUnknownsRestriction(z(NS * 5), z(NS * 5), [0] * dim),
        lambda: c.boundaryConditions(z(NF * 5), z(NF * 5), [0.5] * dim, [0.1] * dim, 0.0, 0.01, 0, 0),
        lambda: c.set_boundary_values(E.BC_WALL),
    ]
    for call in calls:
        with pytest.raises(E.ExaError):
            call()


def test_mesh_bindings_convert_their_arguments():
    m = E.Mesh.__new__(E.Mesh)
    m.h, m.ctx, m.n, m.ncells = None, _null_context(3), 2, 8
    calls = [
        lambda: m.set_solution(np.zeros(8 * m.ctx.nspace * 5)),
        lambda: m.get_solution(),
        lambda: m.adjust_solution([0.0], ic=E.IC_ENTROPY_WAVE),
        lambda: m.init_dt(0.0),
        lambda: m.run(1),
        lambda: m.time(),
        lambda: m.set_boundary([E.BC_WALL] * 6),
        lambda: m.set_time(0.5),
        lambda: m.probe([[0.1, 0.2, 0.3]]),
        lambda: m.step_predict(),
        lambda: m.step_correct(),
        lambda: m.step_advance(),
        lambda: m.halo_size(),
        lambda: m.halo_pack(4, 0, 0),
        lambda: m.halo_unpack(4, 0, 0),
        lambda: m.dt_next(),
        lambda: m.set_dt_next(0.1),
        lambda: m.set_fused(1),
    ]
    for call in calls:
        with pytest.raises(E.ExaError):
            call()


def test_amr_ensemble_group_acoustic_bindings_convert_their_arguments():
    ctx = _null_context(3, 4)
    a = E.Amr.__new__(E.Amr)
    a.h, a.ctx, a.n, a.nleaves, a.refined = None, ctx, 3, 27, np.zeros(27, np.uint8)
    en = E.Ensemble.__new__(E.Ensemble)
    en.h, en.ctx, en.sizes, en.ninst = None, _null_context(2, 4), np.array([9, 10], np.int32), 2
    g = E.Group.__new__(E.Group)
    g.h, g.members = None, []
    ac = E.Acoustic.__new__(E.Acoustic)
    ac.h, ac.ctx, ac.n, ac.ncells, ac.nrec = None, _null_context(3, 4), 2, 8, 0
    NSa = a.ctx.nspace * 5
    calls = [
        lambda: a.leaves(), lambda: a.set_solution(np.zeros(27 * NSa)), lambda: a.get_solution(),
        lambda: a.init_dt(0.0), lambda: a.run(1), lambda: a.sync(), lambda: a.time(), lambda: a.picard(),
        lambda: a.autotune(),
        lambda: en.init_dt(0.0), lambda: en.run(1), lambda: en.sync(), lambda: en.time(0), lambda: en.autotune(),
        lambda: g.run(1), lambda: g.reset(),
        lambda: ac.set_dt(0.0), lambda: ac.set_step(0), lambda: ac.time(), lambda: ac.run(1),
        lambda: ac.set_source(np.full(3, 0.5), 1.0, 5.0, 0.1), lambda: ac.set_receivers(np.full((2, 3), 0.5)),
        lambda: ac.set_timing(True), lambda: ac.set_predictor(0), lambda: ac.phase_times(), lambda: ac.last_times(),
        lambda: ac.get_material(),
    ]
    for call in calls:
        with pytest.raises(E.ExaError):
            call()

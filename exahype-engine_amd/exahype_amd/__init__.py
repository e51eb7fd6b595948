"""ctypes binding of libexahype_amd.so (the C ABI in include/exahype_amd.h).

The product path: every call goes to the HIP library.  There is no CPU
fallback -- importing this module raises if the library is missing, and every
entry point raises ``ExaError`` with the library's message on failure.

Method names mirror the reference's ADER-DG kernel API
(exahype::solvers::ADERDGSolver, ExaHyPE/exahype/solvers/ADERDGSolver.h:2392-2660).
"""
import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ENGINE_DIR = os.path.dirname(PKG_DIR)
REPO_DIR = os.path.dirname(ENGINE_DIR)
LIB_PATH = os.path.join(ENGINE_DIR, "libexahype_amd.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "exahype_amd.h")

PDE_EULER, PDE_ACOUSTIC, PDE_ADVECTION, PDE_USER = 0, 1, 2, 3
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
DIALECT_GENERIC, DIALECT_OPTIMISED, DIALECT_OPTIMISED_CERK, DIALECT_SPLIT_CK = 0, 1, 2, 3
GUESS_TRIVIAL, GUESS_CERK = 0, 1
CONVERTERS = ("Qhbnd", "Fhbnd", "QFace", "FFace", "Qhi", "Fhi", "PSi", "Lduh")
EXA_OK, EXA_ERR_ARG, EXA_ERR_HIP, EXA_ERR_UNSUPPORTED, EXA_ERR_NOMEM = 0, 1, 2, 3, 4

if not os.path.exists(LIB_PATH):
    raise ImportError(f"exahype_amd: {LIB_PATH} is missing -- build it (python -c 'import __graft_entry__ as g; g.build()')")

_lib = ctypes.CDLL(LIB_PATH)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_cp = ctypes.c_void_p  # opaque handles


class ExaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"exahype_amd error {code}: {msg}")
        self.code = code


def _sig(name, args, res=ctypes.c_int):
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("exa_last_error", [], ctypes.c_char_p)
_sig("exa_ctx_create", [ctypes.POINTER(_cp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                        ctypes.c_int, ctypes.c_int, ctypes.c_double])
_sig("exa_ctx_destroy", [_cp])
_sig("exa_ctx_set_advection", [_cp, _dp])
_sig("exa_ctx_set_cell_dx", [_cp, ctypes.c_int])
_sig("exa_ctx_set_pde_params", [_cp, _dp, ctypes.c_int])
_sig("exa_ctx_set_point_source", [_cp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double])
_sig("exa_ctx_set_predictor_variant", [_cp, ctypes.c_int])
_sig("exa_ctx_set_dialect", [_cp, ctypes.c_int, ctypes.c_int])
_sig("exa_ctx_set_initial_guess", [_cp, ctypes.c_int])
_sig("exa_ctx_set_picard_guard", [_cp, ctypes.c_double, ctypes.c_int])
_sig("exa_ctx_padded_size", [_cp, ctypes.c_int, _ip])
_sig("exa_opt_fusedSpaceTimePredictorVolumeIntegral",
     [_cp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _ip])
for _name in CONVERTERS:
    _sig(f"exa_converter_{_name}_optimised2generic", [_cp, _dp, _dp])
    _sig(f"exa_converter_{_name}_generic2optimised", [_cp, _dp, _dp])
_sig("exa_tables", [ctypes.c_int] + [_dp] * 9)
_sig("exa_ctx_tables", [_cp] + [_dp] * 8)
_sig("exa_fusedSpaceTimePredictorVolumeIntegral",
     [_cp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip])
_sig("exa_riemannSolver", [_cp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int])
_sig("exa_faceIntegral", [_cp, _dp, _dp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _dp, ctypes.c_double,
                          ctypes.c_int])
_sig("exa_surfaceIntegral", [_cp, _dp, _dp, _dp])
_sig("exa_addUpdateToSolution", [_cp, _dp, _dp, _dp, ctypes.c_double])
_sig("exa_stableTimeStepSize", [_cp, _dp, _dp, _dp])
_sig("exa_batch_predictor", [_cp, _vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp])
_sig("exa_batch_riemann", [_cp, _vp, ctypes.c_long, _vp, _vp, _vp])
_sig("exa_batch_corrector", [_cp, _vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp])
_sig("exa_batch_stable_dt", [_cp, _vp, ctypes.c_long, _vp, _vp, _vp])
_sig("exa_mesh_create_uniform", [ctypes.POINTER(_cp), _cp, ctypes.c_int, ctypes.c_double])
_sig("exa_mesh_destroy", [_cp])
_sig("exa_mesh_ncells", [_cp], ctypes.c_long)
_sig("exa_mesh_set_solution", [_cp, _dp])
_sig("exa_mesh_get_solution", [_cp, _dp])
_sig("exa_mesh_device_solution", [_cp], ctypes.c_void_p)
_sig("exa_mesh_adjust_solution", [_cp, ctypes.c_int, _dp, ctypes.c_int])
IC_DENSITY_WAVE, IC_ENTROPY_WAVE = 1, 2
BC_PERIODIC, BC_WALL, BC_DIRICHLET_ENTROPY_WAVE = 0, 1, 2
_sig("exa_ctx_set_boundary_values", [_cp, ctypes.c_int])
_sig("exa_boundaryConditions", [_cp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int])
_sig("exa_mesh_set_boundary", [_cp, _ip])
_sig("exa_mesh_set_time", [_cp, ctypes.c_double])
_sig("exa_mesh_probe", [_cp, ctypes.c_int, _dp, _dp])
_sig("exa_mesh_global_observables", [_cp, _dp, ctypes.c_int])
BC_EXCHANGE = 3
_sig("exa_mesh_step_predict", [_cp])
_sig("exa_mesh_step_correct", [_cp])
_sig("exa_mesh_step_advance", [_cp])
_sig("exa_mesh_halo_size", [_cp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)])
_sig("exa_mesh_halo_pack", [_cp, ctypes.c_int, _vp, _vp])
_sig("exa_mesh_halo_unpack", [_cp, ctypes.c_int, _vp, _vp])
_sig("exa_mesh_device_state", [_cp], ctypes.c_void_p)
_sig("exa_mesh_halo_pack_host", [_cp, ctypes.c_int, _dp, _dp])
_sig("exa_mesh_halo_unpack_host", [_cp, ctypes.c_int, _dp, _dp])
_sig("exa_mesh_dt_next", [_cp, _dp])
_sig("exa_mesh_set_dt_next", [_cp, ctypes.c_double])
_sig("exa_mesh_init_dt", [_cp, ctypes.c_double])
_sig("exa_mesh_run", [_cp, ctypes.c_int, ctypes.c_int])
_sig("exa_mesh_sync", [_cp])
_sig("exa_mesh_time", [_cp, _dp, _dp])
_sig("exa_mesh_picard", [_cp, _ip])
_sig("exa_mesh_picard_guard", [_cp, ctypes.POINTER(ctypes.c_long)])
_sig("exa_mesh_stream", [_cp], ctypes.c_void_p)
_sig("exa_mesh_set_timing", [_cp, ctypes.c_int])
_sig("exa_mesh_set_fused", [_cp, ctypes.c_int])
_sig("exa_mesh_set_overlap", [_cp, ctypes.c_int])
_sig("exa_mesh_phase_times", [_cp, _dp, _dp, _dp, _ip])
_I, _D = ctypes.c_int, ctypes.c_double
_u8p = ctypes.POINTER(ctypes.c_ubyte)
_sig("exa_faceUnknownsProlongation", [_cp, _dp, _dp, _dp, _dp, _I, _I, _ip])
_sig("exa_faceUnknownsRestriction", [_cp, _dp, _dp, _dp, _dp, _I, _I, _ip])
_sig("exa_volumeUnknownsProlongation", [_cp, _dp, _dp, _I, _I, _ip])
_sig("exa_volumeUnknownsRestriction", [_cp, _dp, _dp, _I, _I, _ip])
_sig("exa_amr_create", [ctypes.POINTER(_cp), _cp, _I, _u8p])
_sig("exa_amr_destroy", [_cp])
_sig("exa_amr_nleaves", [_cp], ctypes.c_long)
_sig("exa_amr_leaves", [_cp, _ip, _ip])
_sig("exa_amr_set_solution", [_cp, _dp])
_sig("exa_amr_get_solution", [_cp, _dp])
_sig("exa_amr_init_dt", [_cp, _D])
_sig("exa_amr_run", [_cp, _I, _I])
_sig("exa_amr_sync", [_cp])
_sig("exa_amr_time", [_cp, _dp, _dp])
_sig("exa_amr_autotune", [_cp, _ip, _dp])
_sig("exa_amr_picard", [_cp, _ip])
_sig("exa_ensemble_create", [ctypes.POINTER(_cp), _cp, _I, _ip])
_sig("exa_ensemble_destroy", [_cp])
_sig("exa_ensemble_ncells", [_cp, _I], ctypes.c_long)
_sig("exa_ensemble_set_solution", [_cp, _I, _dp])
_sig("exa_ensemble_get_solution", [_cp, _I, _dp])
_sig("exa_ensemble_init_dt", [_cp, _D])
_sig("exa_ensemble_run", [_cp, _I, _I])
_sig("exa_ensemble_sync", [_cp])
_sig("exa_ensemble_time", [_cp, _I, _dp, _dp])
_sig("exa_ensemble_autotune", [_cp, _ip, _dp])
_sig("exa_ensemble_picard", [_cp, _ip])
_sig("exa_group_create", [ctypes.POINTER(_cp), _I, ctypes.POINTER(_cp), _I, ctypes.POINTER(_cp)])
_sig("exa_group_destroy", [_cp])
_sig("exa_group_run", [_cp, _I, _I])
_sig("exa_group_reset", [_cp])
_sig("exa_acoustic_create", [ctypes.POINTER(_cp), _cp, _I, _D])
_sig("exa_acoustic_destroy", [_cp])
_sig("exa_acoustic_ncells", [_cp], ctypes.c_long)
_sig("exa_acoustic_set_solution", [_cp, _dp])
_sig("exa_acoustic_get_solution", [_cp, _dp])
_sig("exa_acoustic_set_material", [_cp, _dp, _dp])
_sig("exa_acoustic_get_material", [_cp, _dp, _dp])
_sig("exa_acoustic_set_dt", [_cp, _D])
_sig("exa_acoustic_set_step", [_cp, _I])
_sig("exa_acoustic_time", [_cp, _ip, _dp, _dp])
_sig("exa_acoustic_set_source", [_cp, _dp, _D, _D, _D])
_sig("exa_acoustic_set_receivers", [_cp, _I, _dp])
_sig("exa_acoustic_run", [_cp, _I, _I, _dp])
_sig("exa_acoustic_misfit_gradient", [_cp, _I, _I, _dp, _dp, _dp])
_sig("exa_acoustic_last_times", [_cp, _dp, _dp, _dp])
_sig("exa_acoustic_set_timing", [_cp, _I])
_sig("exa_acoustic_set_predictor", [_cp, _I])
_sig("exa_acoustic_phase_times", [_cp, _dp, _dp, _ip])


def declared_symbols():
    """Names of all functions declared in include/exahype_amd.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exa_\w+)\s*\(", text)))


def _check(rc):
    if rc != EXA_OK:
        msg = _lib.exa_last_error()
        raise ExaError(rc, msg.decode() if msg else "")


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected a C-contiguous float64 array"
    return a.ctypes.data_as(_dp)


def tables(order):
    """Operator tables derived by the library (host only, no device needed)."""
    B = order + 1
    out = {k: np.zeros(n) for k, n in (("w", B), ("x", B), ("Kxi", B * B), ("iK1", B * B), ("dudx", B * B),
                                       ("F0", B), ("F1", B), ("fgp", 3 * B * B), ("egp", B * B))}
    _check(_lib.exa_tables(order, *[_p(out[k]) for k in ("w", "x", "Kxi", "iK1", "dudx", "F0", "F1", "fgp", "egp")]))
    for k in ("Kxi", "iK1", "dudx", "egp"):
        out[k] = out[k].reshape(B, B)
    out["fgp"] = out["fgp"].reshape(3, B, B)
    return out


class Context:
    """One solver configuration; methods follow the reference kernel API."""

    def __init__(self, dim, order, nvar=5, npar=0, pde=PDE_EULER, precision=PREC_FP64, picard_max=0, cfl=0.9):
        h = _cp()
        _check(_lib.exa_ctx_create(ctypes.byref(h), dim, order, nvar, npar, pde, precision, picard_max, cfl))
        self.h = h
        self.dim, self.order, self.nvar, self.npar = dim, order, nvar, npar
        self.B = order + 1

    def close(self):
        if self.h:
            _lib.exa_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nd(self):
        return self.nvar + self.npar

    @property
    def nspace(self):
        return self.B ** self.dim

    @property
    def nface(self):
        return self.B ** (self.dim - 1)

    def set_predictor_variant(self, variant):
        """0: fastest kernel for the shape (default), 1: generic node-column kernel."""
        _check(_lib.exa_ctx_set_predictor_variant(self.h, variant))

    def set_dialect(self, dialect, vect_size=1):
        """Generic (0) or optimised-kernel (1; 2 with the CERK guess) Picard rule and padding width
        (nonlinear contexts); DIALECT_SPLIT_CK (3): the split Cauchy-Kowalevski predictor (linear)."""
        _check(_lib.exa_ctx_set_dialect(self.h, dialect, vect_size))

    def set_initial_guess(self, guess):
        """GUESS_TRIVIAL (luh at every time node) or GUESS_CERK."""
        _check(_lib.exa_ctx_set_initial_guess(self.h, guess))

    def set_picard_guard(self, tol=1e-6, recompute_count=0):
        """Relative proxy-residual guard of the FP32 one-iteration CERK form (tol <= 0: off)."""
        _check(_lib.exa_ctx_set_picard_guard(self.h, tol, recompute_count))

    def padded_size(self, which):
        """nVarPad (which = 0) or nDataPad (which = 1)."""
        n = ctypes.c_int(0)
        _check(_lib.exa_ctx_padded_size(self.h, which, ctypes.byref(n)))
        return n.value

    def opt_fusedSpaceTimePredictorVolumeIntegral(self, luh, inverseDx, dt):
        """The optimised kernels' predictor + volume integral on padded layouts."""
        D = self.dim
        VP, NDP = self.padded_size(0), self.padded_size(1)
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        lduh = np.full(self.nspace * VP, np.nan)
        lQhbnd = np.full(2 * D * self.nface * NDP, np.nan)
        lFhbnd = np.full(2 * D * self.nface * VP, np.nan)
        pic = ctypes.c_int(0)
        _check(_lib.exa_opt_fusedSpaceTimePredictorVolumeIntegral(self.h, _p(lduh), _p(lQhbnd), _p(lFhbnd), _p(luh),
                                                                   inverseDx, dt, ctypes.byref(pic)))
        return dict(iters=pic.value, lduh=lduh, lQhbnd=lQhbnd, lFhbnd=lFhbnd)

    def convert(self, name, direction, src, out_size):
        """exa_converter_<name>_<direction> (direction "optimised2generic" or
        "generic2optimised"); the output starts as NaN so untouched entries show."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        out = np.full(out_size, np.nan)
        _check(getattr(_lib, f"exa_converter_{name}_{direction}")(self.h, _p(src), _p(out)))
        return out

    def set_point_source(self, x, amp, f0, t0):
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        _check(_lib.exa_ctx_set_point_source(self.h, _p(xs), amp, f0, t0))

    def set_pde_params(self, prm):
        """Parameters of the user PDE hooks (at most 16)."""
        p = np.ascontiguousarray(prm, dtype=np.float64)
        _check(_lib.exa_ctx_set_pde_params(self.h, _p(p), p.size))

    def set_advection(self, a):
        _check(_lib.exa_ctx_set_advection(self.h, _p(np.ascontiguousarray(a, dtype=np.float64))))

    def set_cell_dx(self, per_direction):
        """Batched entry points: h is [ncells] (False) or the dx vector per cell [ncells][dim] (True)."""
        _check(_lib.exa_ctx_set_cell_dx(self.h, int(bool(per_direction))))

    # -- per-cell compatibility API ---------------------------------------
    def fusedSpaceTimePredictorVolumeIntegral(self, luh, dx, t, dt, addVolumeIntegralResultToUpdate=True,
                                              centre=None):
        D, V, ND = self.dim, self.nvar, self.nd
        luh = np.array(luh, dtype=np.float64)
        lduh = np.zeros(self.nspace * V)
        lQhbnd = np.zeros(2 * D * self.nface * ND)
        lFhbnd = np.zeros(2 * D * self.nface * V)
        centre = np.zeros(D) if centre is None else np.ascontiguousarray(centre, dtype=np.float64)
        pic = ctypes.c_int(0)
        _check(_lib.exa_fusedSpaceTimePredictorVolumeIntegral(
            self.h, _p(lduh), _p(lQhbnd), None, _p(lFhbnd), _p(luh), _p(centre),
            _p(np.array(dx, dtype=np.float64)), t, dt, int(bool(addVolumeIntegralResultToUpdate)), ctypes.byref(pic)))
        return dict(iters=pic.value, lduh=lduh, lQhbnd=lQhbnd, lFhbnd=lFhbnd, luh=luh)

    def riemannSolver(self, FL, FR, QL, QR, direction, dx=None, t=0.0, dt=0.0):
        FL = np.array(FL, dtype=np.float64)
        FR = np.array(FR, dtype=np.float64)
        dxa = np.array(dx if dx is not None else [1.0] * self.dim, dtype=np.float64)
        _check(_lib.exa_riemannSolver(self.h, _p(FL), _p(FR), _p(np.ascontiguousarray(QL, dtype=np.float64)),
                                      _p(np.ascontiguousarray(QR, dtype=np.float64)), t, dt, _p(dxa), direction, 0, 0))
        return FL, FR

    def set_boundary_values(self, kind):
        """boundaryValues hook of boundaryConditions: BC_WALL or BC_DIRICHLET_ENTROPY_WAVE."""
        _check(_lib.exa_ctx_set_boundary_values(self.h, kind))

    def boundaryConditions(self, fluxIn, stateIn, centre, dx, t, dt, direction, orientation):
        """The reference's boundaryConditions for one face of one cell: returns the new fluxIn."""
        F = np.array(fluxIn, dtype=np.float64)
        _check(_lib.exa_boundaryConditions(self.h, _p(F), _p(np.ascontiguousarray(stateIn, dtype=np.float64)), None,
                                           None, _p(np.ascontiguousarray(centre, dtype=np.float64)),
                                           _p(np.ascontiguousarray(dx, dtype=np.float64)), t, dt, direction,
                                           orientation))
        return F

    def faceIntegral(self, out, lFhbnd_face, direction, orientation, dx, dt=0.0, addToUpdate=True,
                     subfaceIndex=None, levelDelta=0):
        out = np.array(out, dtype=np.float64)
        sub = None if subfaceIndex is None else np.ascontiguousarray(subfaceIndex, dtype=np.int32)
        _check(_lib.exa_faceIntegral(self.h, _p(out), _p(np.ascontiguousarray(lFhbnd_face, dtype=np.float64)),
                                     direction, orientation, _ipv(sub), levelDelta,
                                     _p(np.array(dx, dtype=np.float64)), dt, int(bool(addToUpdate))))
        return out

    # AMR projections (reference amrRoutines.cpph); prolongations return the
    # fine arrays, restrictions return coarse + restricted fine
    def faceUnknownsProlongation(self, lQhbndCoarse, lFhbndCoarse, subfaceIndex, coarseGridLevel=0,
                                 fineGridLevel=1):
        qf = np.zeros(self.nface * self.nd)
        ff = np.zeros(self.nface * self.nvar)
        sub = np.ascontiguousarray(subfaceIndex, dtype=np.int32)
        _check(_lib.exa_faceUnknownsProlongation(self.h, _p(qf), _p(ff),
                                                 _p(np.ascontiguousarray(lQhbndCoarse, dtype=np.float64)),
                                                 _p(np.ascontiguousarray(lFhbndCoarse, dtype=np.float64)),
                                                 coarseGridLevel, fineGridLevel, _ipv(sub)))
        return qf, ff

    def faceUnknownsRestriction(self, lQhbndCoarse, lFhbndCoarse, lQhbndFine, lFhbndFine, subfaceIndex,
                                coarseGridLevel=0, fineGridLevel=1):
        qc = np.array(lQhbndCoarse, dtype=np.float64)
        fc = np.array(lFhbndCoarse, dtype=np.float64)
        sub = np.ascontiguousarray(subfaceIndex, dtype=np.int32)
        _check(_lib.exa_faceUnknownsRestriction(self.h, _p(qc), _p(fc),
                                                _p(np.ascontiguousarray(lQhbndFine, dtype=np.float64)),
                                                _p(np.ascontiguousarray(lFhbndFine, dtype=np.float64)),
                                                coarseGridLevel, fineGridLevel, _ipv(sub)))
        return qc, fc

    def volumeUnknownsProlongation(self, luhCoarse, subcellIndex, coarseGridLevel=0, fineGridLevel=1):
        out = np.zeros(self.nspace * self.nd)
        sub = np.ascontiguousarray(subcellIndex, dtype=np.int32)
        _check(_lib.exa_volumeUnknownsProlongation(self.h, _p(out), _p(np.ascontiguousarray(luhCoarse, dtype=np.float64)),
                                                   coarseGridLevel, fineGridLevel, _ipv(sub)))
        return out

    def volumeUnknownsRestriction(self, luhCoarse, luhFine, subcellIndex, coarseGridLevel=0, fineGridLevel=1):
        out = np.array(luhCoarse, dtype=np.float64)
        sub = np.ascontiguousarray(subcellIndex, dtype=np.int32)
        _check(_lib.exa_volumeUnknownsRestriction(self.h, _p(out), _p(np.ascontiguousarray(luhFine, dtype=np.float64)),
                                                  coarseGridLevel, fineGridLevel, _ipv(sub)))
        return out

    def surfaceIntegral(self, lduh, lFhbnd, dx):
        lduh = np.array(lduh, dtype=np.float64)
        _check(_lib.exa_surfaceIntegral(self.h, _p(lduh), _p(np.ascontiguousarray(lFhbnd, dtype=np.float64)),
                                        _p(np.array(dx, dtype=np.float64))))
        return lduh

    def addUpdateToSolution(self, luh, lduh, dt):
        luh = np.array(luh, dtype=np.float64)
        _check(_lib.exa_addUpdateToSolution(self.h, _p(luh), _p(luh.copy()),
                                            _p(np.ascontiguousarray(lduh, dtype=np.float64)), dt))
        return luh

    def stableTimeStepSize(self, luh, dx):
        out = ctypes.c_double(0)
        _check(_lib.exa_stableTimeStepSize(self.h, _p(np.ascontiguousarray(luh, dtype=np.float64)),
                                           _p(np.array(dx, dtype=np.float64)), ctypes.byref(out)))
        return out.value

    # -- batched device API (device pointers as ints, e.g. torch data_ptr()) --
    def batch_predictor(self, stream, ncells, luh, h, dt, lduh, lQhbnd, lFhbnd, picard=None):
        _check(_lib.exa_batch_predictor(self.h, stream, ncells, luh, h, dt, lduh, lQhbnd, lFhbnd, picard))

    def batch_riemann(self, stream, nfaces, faces, lQhbnd, lFhbnd):
        _check(_lib.exa_batch_riemann(self.h, stream, nfaces, faces, lQhbnd, lFhbnd))

    def batch_corrector(self, stream, ncells, luh, lduh, lFhbnd, h, dt, dt_next):
        _check(_lib.exa_batch_corrector(self.h, stream, ncells, luh, lduh, lFhbnd, h, dt, dt_next))

    def batch_stable_dt(self, stream, ncells, luh, h, dt_out):
        _check(_lib.exa_batch_stable_dt(self.h, stream, ncells, luh, h, dt_out))


class Mesh:
    """Uniform periodic n^D mesh with device-resident patch arrays."""

    def __init__(self, ctx, n, h):
        m = _cp()
        _check(_lib.exa_mesh_create_uniform(ctypes.byref(m), ctx.h, n, h))
        self.h = m
        self.ctx = ctx
        self.n = n
        self.ncells = _lib.exa_mesh_ncells(m)

    def close(self):
        if self.h:
            _lib.exa_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solution(self, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        assert luh.size == self.ncells * self.ctx.nspace * self.ctx.nd
        _check(_lib.exa_mesh_set_solution(self.h, _p(luh)))

    def get_solution(self):
        out = np.zeros(self.ncells * self.ctx.nspace * self.ctx.nd)
        _check(_lib.exa_mesh_get_solution(self.h, _p(out)))
        return out

    def adjust_solution(self, params, ic=IC_DENSITY_WAVE):
        """Evaluate a compiled initial condition on the device (adjustSolution at t = 0)."""
        prm = np.ascontiguousarray(params, dtype=np.float64)
        _check(_lib.exa_mesh_adjust_solution(self.h, ic, _p(prm), prm.size))

    def init_dt(self, dt=0.0):
        _check(_lib.exa_mesh_init_dt(self.h, dt))

    def set_fused(self, enabled):
        """Two-kernel fused step (default where compiled) or the four-kernel form."""
        _check(_lib.exa_mesh_set_fused(self.h, int(bool(enabled))))

    def set_overlap(self, slabs):
        """Slab-pipelined fused step: `slabs` (>= 3) slabs of z planes, the corrector
        of a slab beside the predictors of the later ones; 0 switches it off.
        3D fused periodic meshes with the one-iteration FP32 CERK predictor."""
        _check(_lib.exa_mesh_set_overlap(self.h, int(slabs)))

    def set_boundary(self, kinds):
        """Boundary kind per domain face [2*d + o] (BC_*)."""
        k = np.ascontiguousarray(kinds, dtype=np.int32)
        assert k.size == 2 * self.ctx.dim
        _check(_lib.exa_mesh_set_boundary(self.h, k.ctypes.data_as(_ip)))

    def set_time(self, t):
        _check(_lib.exa_mesh_set_time(self.h, t))

    # ---- partitioned stepping with halo messages ----
    def step_predict(self):
        _check(_lib.exa_mesh_step_predict(self.h))

    def step_correct(self):
        _check(_lib.exa_mesh_step_correct(self.h))

    def step_advance(self):
        _check(_lib.exa_mesh_step_advance(self.h))

    def global_observables(self):
        """dict: dgcells, fvcells, integral[nData], min[nData], max[nData] of the current solution."""
        nd = self.ctx.nd
        out = np.zeros(2 + 3 * nd)
        _check(_lib.exa_mesh_global_observables(self.h, _p(out), out.size))
        return {"dgcells": out[0], "fvcells": out[1], "integral": out[2:2 + nd], "min": out[2 + nd:2 + 2 * nd],
                "max": out[2 + 2 * nd:]}

    def halo_size(self):
        a, b = ctypes.c_long(0), ctypes.c_long(0)
        _check(_lib.exa_mesh_halo_size(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def halo_pack(self, face, d_msg_q, d_msg_f):
        """Device pointers (ints) of the two message buffers."""
        _check(_lib.exa_mesh_halo_pack(self.h, face, ctypes.c_void_p(d_msg_q), ctypes.c_void_p(d_msg_f)))

    def halo_unpack(self, face, d_msg_q, d_msg_f):
        _check(_lib.exa_mesh_halo_unpack(self.h, face, ctypes.c_void_p(d_msg_q), ctypes.c_void_p(d_msg_f)))

    def halo_pack_host(self, face):
        """The two messages of a domain face as host arrays (lQhbnd faces, lFhbnd faces)."""
        nq, nf = self.halo_size()
        q, f = np.zeros(nq), np.zeros(nf)
        _check(_lib.exa_mesh_halo_pack_host(self.h, face, _p(q), _p(f)))
        return q, f

    def halo_unpack_host(self, face, msg_q, msg_f):
        _check(_lib.exa_mesh_halo_unpack_host(self.h, face, _p(np.ascontiguousarray(msg_q, dtype=np.float64)),
                                              _p(np.ascontiguousarray(msg_f, dtype=np.float64))))

    def dt_next(self):
        v = ctypes.c_double(0)
        _check(_lib.exa_mesh_dt_next(self.h, ctypes.byref(v)))
        return v.value

    def set_dt_next(self, v):
        _check(_lib.exa_mesh_set_dt_next(self.h, v))

    def probe(self, points):
        """Solution values [npoints][nData] at points [npoints][dim] (Lagrange interpolation)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, self.ctx.dim)
        out = np.zeros((pts.shape[0], self.ctx.nd))
        _check(_lib.exa_mesh_probe(self.h, pts.shape[0], _p(pts), _p(out)))
        return out

    def run(self, nsteps, use_graph=False):
        _check(_lib.exa_mesh_run(self.h, nsteps, int(bool(use_graph))))

    def sync(self):
        _check(_lib.exa_mesh_sync(self.h))

    def time(self):
        t, dt = ctypes.c_double(0), ctypes.c_double(0)
        _check(_lib.exa_mesh_time(self.h, ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def picard(self):
        out = np.zeros(self.ncells, dtype=np.int32)
        _check(_lib.exa_mesh_picard(self.h, out.ctypes.data_as(_ip)))
        return out

    def picard_guard_recomputed(self):
        """Cells the Picard guard recomputed since the mesh was created."""
        v = ctypes.c_long(0)
        _check(_lib.exa_mesh_picard_guard(self.h, ctypes.byref(v)))
        return v.value

    @property
    def stream(self):
        return _lib.exa_mesh_stream(self.h)

    def set_timing(self, enabled):
        _check(_lib.exa_mesh_set_timing(self.h, int(bool(enabled))))

    def phase_times(self):
        a, b, c, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
        _check(_lib.exa_mesh_phase_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(n)))
        return dict(predictor_ms=a.value, riemann_ms=b.value, corrector_ms=c.value, steps=n.value)


def _ipv(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


class Amr:
    """Static two-level 3D mesh (refined coarse cells split into 3^3 children)."""

    def __init__(self, ctx, n, refined):
        self.refined = np.ascontiguousarray(refined, dtype=np.uint8)
        assert self.refined.size == n ** 3
        m = _cp()
        _check(_lib.exa_amr_create(ctypes.byref(m), ctx.h, n, self.refined.ctypes.data_as(_u8p)))
        self.h = m
        self.ctx = ctx
        self.n = n
        self.nleaves = _lib.exa_amr_nleaves(m)

    def close(self):
        if self.h:
            _lib.exa_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def leaves(self):
        level = np.zeros(self.nleaves, dtype=np.int32)
        coords = np.zeros((self.nleaves, 3), dtype=np.int32)
        _check(_lib.exa_amr_leaves(self.h, _ipv(level), _ipv(coords)))
        return level, coords

    def set_solution(self, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        assert luh.size == self.nleaves * self.ctx.nspace * self.ctx.nd
        _check(_lib.exa_amr_set_solution(self.h, _p(luh)))

    def get_solution(self):
        out = np.zeros(self.nleaves * self.ctx.nspace * self.ctx.nd)
        _check(_lib.exa_amr_get_solution(self.h, _p(out)))
        return out

    def init_dt(self, dt=0.0):
        _check(_lib.exa_amr_init_dt(self.h, dt))

    def run(self, nsteps, use_graph=False):
        _check(_lib.exa_amr_run(self.h, nsteps, 1 if use_graph else 0))

    def sync(self):
        _check(_lib.exa_amr_sync(self.h))

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        _check(_lib.exa_amr_time(self.h, ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def picard(self):
        out = np.zeros(self.nleaves, dtype=np.int32)
        _check(_lib.exa_amr_picard(self.h, _ipv(out)))
        return out

    def autotune(self):
        cfg = np.zeros(3, dtype=np.int32)
        ms = np.zeros(3)
        _check(_lib.exa_amr_autotune(self.h, _ipv(cfg), _p(ms)))
        return dict(predictor_grid=int(cfg[0]), riemann_waves=int(cfg[1]), corrector_grid=int(cfg[2]),
                    ms=ms.tolist())


class Ensemble:
    """Independent uniform periodic meshes of n[i] cells per side, stepped together."""

    def __init__(self, ctx, n):
        self.sizes = np.ascontiguousarray(n, dtype=np.int32)
        m = _cp()
        _check(_lib.exa_ensemble_create(ctypes.byref(m), ctx.h, len(self.sizes), _ipv(self.sizes)))
        self.h = m
        self.ctx = ctx
        self.ninst = len(self.sizes)

    def close(self):
        if self.h:
            _lib.exa_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ncells(self, inst=-1):
        return _lib.exa_ensemble_ncells(self.h, inst)

    def set_solution(self, inst, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        assert luh.size == self.ncells(inst) * self.ctx.nspace * self.ctx.nd
        _check(_lib.exa_ensemble_set_solution(self.h, inst, _p(luh)))

    def get_solution(self, inst):
        out = np.zeros(self.ncells(inst) * self.ctx.nspace * self.ctx.nd)
        _check(_lib.exa_ensemble_get_solution(self.h, inst, _p(out)))
        return out

    def init_dt(self, dt=0.0):
        _check(_lib.exa_ensemble_init_dt(self.h, dt))

    def run(self, nsteps, use_graph=False):
        _check(_lib.exa_ensemble_run(self.h, nsteps, 1 if use_graph else 0))

    def sync(self):
        _check(_lib.exa_ensemble_sync(self.h))

    def time(self, inst):
        t, dt = ctypes.c_double(), ctypes.c_double()
        _check(_lib.exa_ensemble_time(self.h, inst, ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def picard(self):
        out = np.zeros(self.ncells(-1), dtype=np.int32)
        _check(_lib.exa_ensemble_picard(self.h, _ipv(out)))
        return out

    def autotune(self):
        cfg = np.zeros(3, dtype=np.int32)
        ms = np.zeros(3)
        _check(_lib.exa_ensemble_autotune(self.h, _ipv(cfg), _p(ms)))
        return dict(predictor_grid=int(cfg[0]), riemann_waves=int(cfg[1]), corrector_grid=int(cfg[2]),
                    ms=ms.tolist())


class Group:
    """Two-level meshes and ensembles stepped by one captured graph per step."""

    def __init__(self, amrs=(), ensembles=()):
        self.members = list(amrs) + list(ensembles)  # keep them alive
        A = (_cp * max(len(amrs), 1))(*[a.h for a in amrs])
        En = (_cp * max(len(ensembles), 1))(*[e.h for e in ensembles])
        g = _cp()
        _check(_lib.exa_group_create(ctypes.byref(g), len(amrs), A, len(ensembles), En))
        self.h = g

    def close(self):
        if self.h:
            _lib.exa_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps, use_graph=True):
        _check(_lib.exa_group_run(self.h, nsteps, 1 if use_graph else 0))

    def reset(self):
        _check(_lib.exa_group_reset(self.h))


class Acoustic:
    """Linear acoustics (p, u, v, w; per-cell rho, c) on a periodic n^3 mesh,
    forward ADER-DG steps and the misfit gradient with respect to c."""

    def __init__(self, ctx, n, h):
        a = _cp()
        _check(_lib.exa_acoustic_create(ctypes.byref(a), ctx.h, n, h))
        self.h = a
        self.ctx = ctx
        self.n = n
        self.ncells = _lib.exa_acoustic_ncells(a)
        self.nrec = 0

    def close(self):
        if self.h:
            _lib.exa_acoustic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solution(self, luh):
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        assert luh.size == self.ncells * self.ctx.nspace * 6
        _check(_lib.exa_acoustic_set_solution(self.h, _p(luh)))

    def get_solution(self):
        out = np.zeros(self.ncells * self.ctx.nspace * 6)
        _check(_lib.exa_acoustic_get_solution(self.h, _p(out)))
        return out

    def set_material(self, rho=None, c=None):
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        cc = None if c is None else np.ascontiguousarray(c, dtype=np.float64)
        for a in (r, cc):
            assert a is None or a.size == self.ncells
        _check(_lib.exa_acoustic_set_material(self.h, _p(r), _p(cc)))

    def get_material(self):
        r, c = np.zeros(self.ncells), np.zeros(self.ncells)
        _check(_lib.exa_acoustic_get_material(self.h, _p(r), _p(c)))
        return r, c

    def set_dt(self, dt=0.0):
        _check(_lib.exa_acoustic_set_dt(self.h, dt))

    def set_step(self, step):
        _check(_lib.exa_acoustic_set_step(self.h, step))

    def time(self):
        s, t, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _check(_lib.exa_acoustic_time(self.h, ctypes.byref(s), ctypes.byref(t), ctypes.byref(dt)))
        return s.value, t.value, dt.value

    def set_source(self, x, amp, f0, t0):
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        _check(_lib.exa_acoustic_set_source(self.h, _p(xs), amp, f0, t0))

    def set_receivers(self, xr):
        xr = np.ascontiguousarray(xr, dtype=np.float64).reshape(-1, 3)
        _check(_lib.exa_acoustic_set_receivers(self.h, len(xr), _p(xr)))
        self.nrec = len(xr)

    def run(self, nsteps, use_graph=False):
        """Advance nsteps; returns p at the receivers, shape (nsteps, nrec)."""
        rec = np.zeros((nsteps, max(self.nrec, 1)))
        _check(_lib.exa_acoustic_run(self.h, nsteps, 1 if use_graph else 0, _p(rec)))
        return rec[:, :self.nrec]

    def misfit_gradient(self, nsteps, obs, checkpoint_every=20):
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        assert obs.size == nsteps * self.nrec
        J = np.zeros(1)
        g = np.zeros(self.ncells)
        _check(_lib.exa_acoustic_misfit_gradient(self.h, nsteps, checkpoint_every, _p(obs), _p(J), _p(g)))
        return float(J[0]), g

    def last_times(self):
        """(forward, recompute, adjoint) milliseconds of the last run / gradient."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(_lib.exa_acoustic_last_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_timing(self, enabled):
        _check(_lib.exa_acoustic_set_timing(self.h, 1 if enabled else 0))

    def set_predictor(self, variant):
        """Forward predictor form: 0 x-line kernel (default), 1 node kernel."""
        _check(_lib.exa_acoustic_set_predictor(self.h, int(variant)))

    def phase_times(self):
        """Per-kernel milliseconds summed over the last eager run: predictor, corrector, steps."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(_lib.exa_acoustic_phase_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return dict(predictor_ms=a.value, corrector_ms=b.value, steps=n.value)

"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker / CPU baseline.  The
operator tables come from ``tests/golden/legendre_tables.json`` -- the values
the reference generator printed (ExaHyPE/kernels/generated/GaussLegendreBasis.csnippet).
"""
import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "liborc.so")
TABLES_PATH = os.path.join(ROOT, "tests", "golden", "legendre_tables.json")

MAXB = 16
PDE_EULER, PDE_ACOUSTIC, PDE_ADVECTION, PDE_ELASTIC2D, PDE_ADV_NCP_SRC = 0, 1, 2, 3, 4
LIN_FLUX, LIN_NCP, LIN_MM = 1, 2, 4

# PNPM per order (Toolkit/exahype/toolkit/solverController.py:17-23).
PNPM = [1.0, 0.33, 0.17, 0.1, 0.069, 0.045, 0.038, 0.03, 0.02, 0.015]


class Basis(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int),
        ("w", ctypes.c_double * MAXB),
        ("x", ctypes.c_double * MAXB),
        ("Kxi", ctypes.c_double * (MAXB * MAXB)),
        ("iK1", ctypes.c_double * (MAXB * MAXB)),
        ("dudx", ctypes.c_double * (MAXB * MAXB)),
        ("F0", ctypes.c_double * MAXB),
        ("F1", ctypes.c_double * MAXB),
        ("fgp", ctypes.c_double * (3 * MAXB * MAXB)),
    ]


class Solver(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("nvar", ctypes.c_int),
        ("npar", ctypes.c_int),
        ("pde", ctypes.c_int),
        ("use_max_picard", ctypes.c_int),
        ("max_picard", ctypes.c_int),
        ("cfl", ctypes.c_double),
        ("pnpm", ctypes.c_double),
        ("adv", ctypes.c_double * 3),
        ("b", Basis),
        ("cerk_guess", ctypes.c_int),
        ("opt_iterations", ctypes.c_int),
        ("ncp_b", ctypes.c_double * 3),
        ("kappa", ctypes.c_double),
        ("split_ck", ctypes.c_int),
    ]


_tables = None


def tables(order):
    """Reference operator tables for ``order`` as numpy arrays."""
    global _tables
    if _tables is None:
        with open(TABLES_PATH) as f:
            _tables = json.load(f)["tables"]
    t = _tables[str(order)]
    return {
        "w": np.array(t["weights"]),
        "x": np.array(t["nodes"]),
        "Kxi": np.array(t["Kxi"]),
        "iK1": np.array(t["iK1"]),
        "dudx": np.array(t["dudx"]),
        "F0": np.array(t["FCoeff"][0]),
        "F1": np.array(t["FCoeff"][1]),
        "fgp": np.array(t["fineGridProjector"]),
        "equidistant": np.array(t["equidistantGridProjector"]),
    }


def _lib():
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])
    lib = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sp = ctypes.POINTER(Solver)
    lib.orc_flux.argtypes = [sp, dp, dp]
    lib.orc_eigenvalues.argtypes = [sp, dp, ctypes.c_int, dp]
    lib.orc_stp_nonlinear.argtypes = [sp, dp, dp, ctypes.c_double] + [dp] * 7
    lib.orc_stp_nonlinear.restype = ctypes.c_int
    lib.orc_volume_integral_nonlinear.argtypes = [sp, dp, dp, dp]
    lib.orc_riemann_nonlinear.argtypes = [sp, dp, dp, dp, dp, ctypes.c_int]
    lib.orc_face_integral_nonlinear.argtypes = [sp, dp, dp, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_surface_integral_nonlinear.argtypes = [sp, dp, dp, dp]
    lib.orc_solution_update.argtypes = [sp, dp, dp, dp, ctypes.c_double]
    lib.orc_stable_dt.argtypes = [sp, dp, dp]
    lib.orc_stable_dt.restype = ctypes.c_double
    lib.orc_fused_stp_vi.argtypes = [sp, dp, dp, dp, dp, dp, ctypes.c_double]
    lib.orc_fused_stp_vi.restype = ctypes.c_int
    lib.orc_run_uniform.argtypes = [sp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, ip, ctypes.c_int]
    lib.orc_run_uniform.restype = ctypes.c_int
    I, D = ctypes.c_int, ctypes.c_double
    lib.orc_lin_stp.argtypes = [sp, I, dp, dp, D, dp, dp, dp, dp, dp, dp]
    lib.orc_lin_volume.argtypes = [sp, dp, dp]
    lib.orc_lin_face.argtypes = [sp, dp, dp, I, I, dp]
    lib.orc_lin_surface.argtypes = [sp, dp, dp, dp]
    lib.orc_lin_riemann.argtypes = [sp, I, dp, dp, dp, dp, I]
    lib.orc_ricker.argtypes = [D, D, D]
    lib.orc_ricker.restype = D
    lib.orc_lin_psi.argtypes = [sp, dp, D, D, D, dp, D, D, D, dp]
    lib.orc_lin_psi_split.argtypes = [sp, dp, D, D, D, dp, D, D, D, dp]
    lib.orc_lin_stp_split.argtypes = [sp, I, dp, dp, D, dp, dp, dp, dp, dp]
    lib.orc_lin_run.argtypes = [sp, dp, I, D, D, I, dp, D, D, D, I, dp, dp, dp, I]
    lib.orc_lin_run.restype = I
    lib.orc_lin_misfit_grad.argtypes = [sp, dp, I, D, D, I, dp, D, D, D, I, dp, dp, dp, I]
    lib.orc_lin_misfit_grad.restype = D
    u8p = ctypes.POINTER(ctypes.c_ubyte)
    lib.orc_face_prolongation.argtypes = [sp, dp, dp, dp, dp, I, I, ip]
    lib.orc_face_restriction.argtypes = [sp, dp, dp, dp, dp, I, I, ip]
    lib.orc_face_flux_restriction.argtypes = [sp, dp, dp, ip, I]
    lib.orc_volume_prolongation.argtypes = [sp, dp, dp, I, I, ip]
    lib.orc_volume_restriction.argtypes = [sp, dp, dp, I, I, ip]
    lib.orc_amr_nleaves.argtypes = [I, u8p]
    lib.orc_amr_nleaves.restype = I
    lib.orc_amr_leaves.argtypes = [I, u8p, ip, ip]
    lib.orc_amr_run.argtypes = [sp, I, u8p, dp, I, dp, I]
    lib.orc_amr_run.restype = I
    lib.orc_entropy_wave.argtypes = [sp, dp, D, dp]
    lib.orc_boundary_values.argtypes = [sp, I, dp, D, D, I, I, dp, dp, dp, dp]
    lib.orc_boundary_conditions.argtypes = [sp, I, dp, dp, dp, dp, D, D, I, I]
    lib.orc_run_bc.argtypes = [sp, dp, I, D, I, dp, dp, ip, I]
    lib.orc_run_bc.restype = I
    lib.orc_run_box.argtypes = [sp, dp, ip, D, I, dp, dp, ip, I]
    lib.orc_run_box.restype = I
    return lib


LIB = None


def lib():
    global LIB
    if LIB is None:
        LIB = _lib()
    return LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


BC_PERIODIC, BC_WALL, BC_DIRICHLET_ENTROPY_WAVE = 0, 1, 2


class Oracle:
    """One solver configuration (dimension, order, variables, PDE)."""

    def __init__(self, dim, order, nvar=5, npar=0, pde=PDE_EULER, cfl=0.9, pnpm=None,
                 max_picard=None, adv=(1.0, 0.5, 0.25), table_override=None, dialect=None,
                 ncp_b=(0.0, 0.0, 0.0), kappa=0.0):
        """dialect: None (generic kernels), "cerk" (generic Picard rule with the
        CERK initial guess) or "optimised" / "optimised-cerk" (the
        KernelGenerator kernels' Picard rule, trivial or CERK initial guess)."""
        self.dim, self.order, self.nvar, self.npar = dim, order, nvar, npar
        self.B = order + 1
        s = Solver()
        s.dim, s.nvar, s.npar, s.pde = dim, nvar, npar, pde
        s.opt_iterations = 1 if dialect in ("optimised", "optimised-cerk") else 0
        s.cerk_guess = 1 if dialect in ("optimised-cerk", "cerk") else 0
        s.use_max_picard = 1 if max_picard is not None else 0
        s.max_picard = max_picard if max_picard is not None else order + 1
        s.cfl = cfl
        s.pnpm = PNPM[order] if pnpm is None else pnpm
        for i in range(3):
            s.adv[i] = adv[i]
            s.ncp_b[i] = ncp_b[i]
        s.kappa = kappa
        t = dict(tables(order))
        if table_override:
            t.update(table_override)
        self.t = t
        b = s.b
        b.order = order
        B = self.B
        for i in range(B):
            b.w[i] = t["w"][i]
            b.x[i] = t["x"][i]
            b.F0[i] = t["F0"][i]
            b.F1[i] = t["F1"][i]
            for j in range(B):
                b.Kxi[i * B + j] = t["Kxi"][i][j]
                b.iK1[i * B + j] = t["iK1"][i][j]
                b.dudx[i * B + j] = t["dudx"][i][j]
                for sub in range(3):
                    b.fgp[(sub * B + i) * B + j] = t["fgp"][sub][i][j]
        self.s = s
        self.ps = ctypes.pointer(s)

    # sizes -----------------------------------------------------------------
    @property
    def nd(self):
        return self.nvar + self.npar

    @property
    def nspace(self):
        return self.B ** self.dim

    @property
    def nface(self):
        return self.B ** (self.dim - 1)

    # kernels ---------------------------------------------------------------
    def flux(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        F = np.zeros(self.dim * self.nvar)
        lib().orc_flux(self.ps, _p(Q), _p(F))
        return F.reshape(self.dim, self.nvar)

    def stp(self, luh, dx, dt):
        B, D, V, ND = self.B, self.dim, self.nvar, self.nd
        st = self.nspace * B
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        lQi = np.zeros(st * ND)
        rhs = np.zeros(st * ND)
        lFi = np.zeros(st * (D + 1) * V)
        lQhi = np.zeros(self.nspace * ND)
        lFhi = np.zeros(self.nspace * (D + 1) * V)
        lQhbnd = np.zeros(2 * D * self.nface * ND)
        lFhbnd = np.zeros(2 * D * self.nface * V)
        invdx = np.array([1.0 / d for d in dx], dtype=np.float64)
        it = lib().orc_stp_nonlinear(self.ps, _p(luh), _p(invdx), dt, _p(lQi), _p(rhs), _p(lFi),
                                     _p(lQhi), _p(lFhi), _p(lQhbnd), _p(lFhbnd))
        return dict(iters=it, lQi=lQi, lFi=lFi, lQhi=lQhi, lFhi=lFhi, lQhbnd=lQhbnd, lFhbnd=lFhbnd)

    def volume_integral(self, lFhi, dx):
        lFhi = np.ascontiguousarray(lFhi, dtype=np.float64)
        lduh = np.zeros(self.nspace * self.nvar)
        lib().orc_volume_integral_nonlinear(self.ps, _p(lduh), _p(lFhi), _p(np.array(dx, dtype=np.float64)))
        return lduh

    def fused(self, luh, dx, dt):
        D, V, ND = self.dim, self.nvar, self.nd
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        lduh = np.zeros(self.nspace * V)
        lQhbnd = np.zeros(2 * D * self.nface * ND)
        lFhbnd = np.zeros(2 * D * self.nface * V)
        it = lib().orc_fused_stp_vi(self.ps, _p(lduh), _p(lQhbnd), _p(lFhbnd), _p(luh),
                                    _p(np.array(dx, dtype=np.float64)), dt)
        return dict(iters=it, lduh=lduh, lQhbnd=lQhbnd, lFhbnd=lFhbnd)

    def riemann(self, FL, FR, QL, QR, direction):
        FL = np.array(FL, dtype=np.float64)
        FR = np.array(FR, dtype=np.float64)
        QL = np.ascontiguousarray(QL, dtype=np.float64)
        QR = np.ascontiguousarray(QR, dtype=np.float64)
        lib().orc_riemann_nonlinear(self.ps, _p(FL), _p(FR), _p(QL), _p(QR), direction)
        return FL, FR

    def face_integral(self, lduh, lFhbnd_face, direction, orientation, dx):
        lduh = np.array(lduh, dtype=np.float64)
        lib().orc_face_integral_nonlinear(self.ps, _p(lduh), _p(np.ascontiguousarray(lFhbnd_face, dtype=np.float64)),
                                          direction, orientation, _p(np.array(dx, dtype=np.float64)))
        return lduh

    def surface_integral(self, lduh, lFhbnd, dx):
        lduh = np.array(lduh, dtype=np.float64)
        lib().orc_surface_integral_nonlinear(self.ps, _p(lduh), _p(np.ascontiguousarray(lFhbnd, dtype=np.float64)),
                                             _p(np.array(dx, dtype=np.float64)))
        return lduh

    def solution_update(self, luh, lduh, dt):
        luh = np.array(luh, dtype=np.float64)
        lib().orc_solution_update(self.ps, _p(luh), _p(luh), _p(np.ascontiguousarray(lduh, dtype=np.float64)), dt)
        return luh

    def stable_dt(self, luh, dx):
        return lib().orc_stable_dt(self.ps, _p(np.ascontiguousarray(luh, dtype=np.float64)),
                                   _p(np.array(dx, dtype=np.float64)))

    def run_uniform(self, luh, n, h, steps, dt=0.0, threads=0, want_picard=False):
        """Advance a uniform periodic mesh in place; returns (dt_next, picard)."""
        assert luh.dtype == np.float64 and luh.flags["C_CONTIGUOUS"]
        dt_io = np.array([dt], dtype=np.float64)
        pic = np.zeros(luh.size // (self.nspace * self.nd), dtype=np.int32) if want_picard else None
        rc = lib().orc_run_uniform(self.ps, _p(luh), n, h, steps, _p(dt_io),
                                   _ip(pic) if pic is not None else None, threads)
        if rc != 0:
            raise MemoryError("oracle run_uniform failed to allocate")
        return float(dt_io[0]), pic

    # ---- boundary conditions (Euler hooks of the reference's Euler example) ----
    def entropy_wave(self, x, t):
        Q = np.zeros(self.nvar, dtype=np.float64)
        lib().orc_entropy_wave(self.ps, _p(np.ascontiguousarray(x, dtype=np.float64)), t, _p(Q))
        return Q

    def boundary_values(self, kind, x, t, dt, face_index, direction, FIn, QIn):
        FOut = np.zeros(self.nvar, dtype=np.float64)
        QOut = np.zeros(self.nd, dtype=np.float64)
        lib().orc_boundary_values(self.ps, kind, _p(np.ascontiguousarray(x, dtype=np.float64)), t, dt, face_index,
                                  direction, _p(np.ascontiguousarray(FIn, dtype=np.float64)),
                                  _p(np.ascontiguousarray(QIn, dtype=np.float64)), _p(FOut), _p(QOut))
        return FOut, QOut

    def boundary_conditions(self, kind, FIn, QIn, centre, dx, t, dt, direction, orientation):
        """The glue boundaryConditions for one face: returns the new FIn."""
        F = np.ascontiguousarray(FIn, dtype=np.float64).copy()
        lib().orc_boundary_conditions(self.ps, kind, _p(F), _p(np.ascontiguousarray(QIn, dtype=np.float64)),
                                      _p(np.ascontiguousarray(centre, dtype=np.float64)),
                                      _p(np.ascontiguousarray(dx, dtype=np.float64)), t, dt, direction,
                                      orientation)
        return F

    def run_bc(self, luh, n, h, steps, bc, dt=0.0, t=0.0, threads=0):
        """Advance a uniform mesh with a boundary kind per domain face; returns (t, dt_next)."""
        assert luh.dtype == np.float64 and luh.flags["C_CONTIGUOUS"]
        dt_io = np.array([dt], dtype=np.float64)
        t_io = np.array([t], dtype=np.float64)
        rc = lib().orc_run_bc(self.ps, _p(luh), n, h, steps, _p(dt_io), _p(t_io),
                              _ip(np.ascontiguousarray(bc, dtype=np.int32)), threads)
        if rc == -2:
            raise ValueError("a direction must be periodic on both faces or on neither")
        if rc != 0:
            raise MemoryError("oracle run_bc failed to allocate")
        return float(t_io[0]), float(dt_io[0])

    def run_box(self, luh, ncells, h, steps, bc, dt=0.0, t=0.0, threads=0):
        """run_bc on a box of ncells = (nx, ny[, nz]) cells (x fastest); returns (t, dt_next)."""
        assert luh.dtype == np.float64 and luh.flags["C_CONTIGUOUS"]
        nd = np.ones(3, dtype=np.int32)
        nd[:len(ncells)] = ncells
        dt_io = np.array([dt], dtype=np.float64)
        t_io = np.array([t], dtype=np.float64)
        rc = lib().orc_run_box(self.ps, _p(luh), _ip(nd), h, steps, _p(dt_io), _p(t_io),
                               _ip(np.ascontiguousarray(bc, dtype=np.int32)), threads)
        if rc == -2:
            raise ValueError("a direction must be periodic on both faces or on neither")
        if rc != 0:
            raise MemoryError("oracle run_box failed to allocate")
        return float(t_io[0]), float(dt_io[0])


def _u8(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))


def _iv(v):
    return _ip(np.ascontiguousarray(v, dtype=np.int32))


def _opt(a):
    return _p(a) if a is not None else None


class AmrOracle(Oracle):
    """AMR projections (any dimension) and the static two-level 3D mesh."""

    def face_prolongation(self, lQc, lFc, subface, level_coarse=0, level_fine=1):
        lQf = np.zeros(self.nface * self.nd)
        lFf = np.zeros(self.nface * self.nvar)
        sub = np.ascontiguousarray(subface, dtype=np.int32)
        lib().orc_face_prolongation(self.ps, _p(lQf), _p(lFf), _p(np.ascontiguousarray(lQc, dtype=np.float64)),
                                    _p(np.ascontiguousarray(lFc, dtype=np.float64)), level_coarse, level_fine, _ip(sub))
        return lQf, lFf

    def face_restriction(self, lQc, lFc, lQf, lFf, subface, level_coarse=0, level_fine=1):
        lQc = np.array(lQc, dtype=np.float64)
        lFc = np.array(lFc, dtype=np.float64)
        sub = np.ascontiguousarray(subface, dtype=np.int32)
        lib().orc_face_restriction(self.ps, _p(lQc), _p(lFc), _p(np.ascontiguousarray(lQf, dtype=np.float64)),
                                   _p(np.ascontiguousarray(lFf, dtype=np.float64)), level_coarse, level_fine, _ip(sub))
        return lQc, lFc

    def face_flux_restriction(self, lFf, subface, level_delta=1):
        out = np.zeros(self.nface * self.nvar)
        sub = np.ascontiguousarray(subface, dtype=np.int32)
        lib().orc_face_flux_restriction(self.ps, _p(out), _p(np.ascontiguousarray(lFf, dtype=np.float64)), _ip(sub),
                                        level_delta)
        return out

    def volume_prolongation(self, luhC, subcell, level_coarse=0, level_fine=1):
        out = np.zeros(self.nspace * self.nd)
        sub = np.ascontiguousarray(subcell, dtype=np.int32)
        lib().orc_volume_prolongation(self.ps, _p(out), _p(np.ascontiguousarray(luhC, dtype=np.float64)),
                                      level_coarse, level_fine, _ip(sub))
        return out

    def volume_restriction(self, luhC, luhF, subcell, level_coarse=0, level_fine=1):
        luhC = np.array(luhC, dtype=np.float64)
        sub = np.ascontiguousarray(subcell, dtype=np.int32)
        lib().orc_volume_restriction(self.ps, _p(luhC), _p(np.ascontiguousarray(luhF, dtype=np.float64)),
                                     level_coarse, level_fine, _ip(sub))
        return luhC

    @staticmethod
    def amr_leaves(n, refined):
        r = np.ascontiguousarray(refined, dtype=np.uint8)
        L = lib().orc_amr_nleaves(n, _u8(r))
        level = np.zeros(L, dtype=np.int32)
        coords = np.zeros((L, 3), dtype=np.int32)
        lib().orc_amr_leaves(n, _u8(r), _ip(level), _ip(coords))
        return level, coords

    def amr_run(self, luh, n, refined, steps, dt=0.0, threads=0):
        """Advance the leaves' state in place; returns the dt of the next step."""
        assert luh.dtype == np.float64 and luh.flags["C_CONTIGUOUS"]
        r = np.ascontiguousarray(refined, dtype=np.uint8)
        dt_io = np.array([dt], dtype=np.float64)
        rc = lib().orc_amr_run(self.ps, n, _u8(r), _p(luh), steps, _p(dt_io), threads)
        if rc != 0:
            raise RuntimeError(f"orc_amr_run failed ({rc})")
        return float(dt_io[0])


class LinearOracle(Oracle):
    """Linear (Cauchy-Kowalevski) kernels: acoustic 3D (4 variables, parameters
    rho, c) or the 2D elasticity test PDE (9 variables, lambda, mu, rho)."""

    def __init__(self, dim=3, order=4, pde=PDE_ACOUSTIC, cfl=0.9, split_ck=False):
        nvar, npar = (4, 2) if pde == PDE_ACOUSTIC else (9, 3)
        super().__init__(dim, order, nvar=nvar, npar=npar, pde=pde, cfl=cfl)
        self.s.split_ck = 1 if split_ck else 0

    def stp_split(self, luh, dx, dt, flags=LIN_FLUX | LIN_NCP | LIN_MM, PSi=None):
        """Split Cauchy-Kowalevski predictor + volume integral of the optimised
        linear kernels (fusedSPTVI_linear_split_ck_cpp.template); PSi [B][B^D][V]
        at the Gauss time nodes (psi_split) or None."""
        ND, V, D = self.nd, self.nvar, self.dim
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        lduh = np.zeros(self.nspace * V)
        lQhi = np.zeros(self.nspace * ND)
        lQbnd = np.zeros(2 * D * self.nface * ND)
        lFbnd = np.zeros(2 * D * self.nface * V)
        ps = np.ascontiguousarray(PSi, dtype=np.float64) if PSi is not None else None
        lib().orc_lin_stp_split(self.ps, flags, _p(luh), _p(np.array(dx, dtype=np.float64)), dt, _opt(ps),
                                _p(lduh), _p(lQhi), _p(lQbnd), _p(lFbnd))
        return dict(lduh=lduh, lQhi=lQhi, lQbnd=lQbnd, lFbnd=lFbnd)

    def psi_split(self, centre, h, t, dt, xs, amp, f0, t0):
        out = np.zeros(self.B * self.nspace * self.nvar)
        lib().orc_lin_psi_split(self.ps, _p(np.array(centre, dtype=np.float64)), h, t, dt,
                                _p(np.array(xs, dtype=np.float64)), amp, f0, t0, _p(out))
        return out

    def stp(self, luh, dx, dt, flags=LIN_FLUX | LIN_NCP | LIN_MM, PSi=None, want_levels=False):
        B, D, V, ND = self.B, self.dim, self.nvar, self.nd
        luh = np.ascontiguousarray(luh, dtype=np.float64)
        lQhi = np.zeros(self.nspace * ND)
        lFhi = np.zeros((D + 1) * self.nspace * V)
        lQbnd = np.zeros(2 * D * self.nface * ND)
        lFbnd = np.zeros(2 * D * self.nface * V)
        lev = np.zeros((B + 1) * self.nspace * ND) if want_levels else None
        ps = np.ascontiguousarray(PSi, dtype=np.float64) if PSi is not None else None
        lib().orc_lin_stp(self.ps, flags, _p(luh), _p(np.array(dx, dtype=np.float64)), dt, _opt(ps),
                          _p(lQhi), _p(lFhi), _p(lQbnd), _p(lFbnd), _opt(lev))
        return dict(lQhi=lQhi, lFhi=lFhi, lQbnd=lQbnd, lFbnd=lFbnd, levels=lev)

    def volume_integral(self, lFhi, dx=None):
        lduh = np.zeros(self.nspace * self.nvar)
        lib().orc_lin_volume(self.ps, _p(lduh), _p(np.ascontiguousarray(lFhi, dtype=np.float64)))
        return lduh

    def face_integral(self, lduh, lFhbnd_face, direction, orientation, dx):
        lduh = np.array(lduh, dtype=np.float64)
        lib().orc_lin_face(self.ps, _p(lduh), _p(np.ascontiguousarray(lFhbnd_face, dtype=np.float64)),
                           direction, orientation, _p(np.array(dx, dtype=np.float64)))
        return lduh

    def surface_integral(self, lduh, lFhbnd, dx):
        lduh = np.array(lduh, dtype=np.float64)
        lib().orc_lin_surface(self.ps, _p(lduh), _p(np.ascontiguousarray(lFhbnd, dtype=np.float64)),
                              _p(np.array(dx, dtype=np.float64)))
        return lduh

    def riemann(self, QL, QR, direction, flags=LIN_FLUX | LIN_NCP | LIN_MM):
        n = self.nface * self.nvar
        FL, FR = np.zeros(n), np.zeros(n)
        lib().orc_lin_riemann(self.ps, flags, _p(FL), _p(FR), _p(np.ascontiguousarray(QL, dtype=np.float64)),
                              _p(np.ascontiguousarray(QR, dtype=np.float64)), direction)
        return FL, FR

    def psi(self, centre, h, t, dt, xs, amp, f0, t0):
        out = np.zeros((self.B + 1) * self.nspace * self.nvar)
        lib().orc_lin_psi(self.ps, _p(np.array(centre, dtype=np.float64)), h, t, dt,
                          _p(np.array(xs, dtype=np.float64)), amp, f0, t0, _p(out))
        return out

    @staticmethod
    def ricker(t, f0, t0):
        return lib().orc_ricker(t, f0, t0)

    def run(self, luh, n, h, dt, nsteps, src, receivers, threads=0, want_states=False):
        """Forward acoustic run in place; src = (xs, amp, f0, t0) or None.
        Returns (rec[nsteps][nrec], states or None)."""
        assert luh.dtype == np.float64 and luh.flags["C_CONTIGUOUS"]
        xr = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 3)
        rec = np.zeros((nsteps, max(len(xr), 1)))
        st = np.zeros((nsteps,) + luh.shape) if want_states else None
        xs, amp, f0, t0 = src if src is not None else (None, 0.0, 0.0, 0.0)
        rc = lib().orc_lin_run(self.ps, _p(luh), n, h, dt, nsteps,
                               _opt(np.array(xs, dtype=np.float64) if xs is not None else None), amp, f0, t0,
                               len(xr), _p(xr), _p(rec), _opt(st), threads)
        if rc != 0:
            raise MemoryError("oracle lin_run failed to allocate")
        return rec[:, :len(xr)], st

    def misfit_grad(self, u0, n, h, dt, nsteps, src, receivers, obs, threads=0, want_grad=True):
        """J and dJ/dc per cell (FP64 discrete adjoint, all states kept)."""
        xr = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 3)
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        gc = np.zeros(n ** 3) if want_grad else None
        xs, amp, f0, t0 = src if src is not None else (None, 0.0, 0.0, 0.0)
        J = lib().orc_lin_misfit_grad(self.ps, _p(np.ascontiguousarray(u0, dtype=np.float64)), n, h, dt, nsteps,
                                      _opt(np.array(xs, dtype=np.float64) if xs is not None else None),
                                      amp, f0, t0, len(xr), _p(xr), _p(obs), _opt(gc), threads)
        return J, gc

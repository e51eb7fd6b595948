/*
 * ADER-DG oracle: a plain-C FP64 restatement of the reference's generic
 * ("generic::c") ADER-DG kernels, used ONLY as test infrastructure.
 *
 * Who may use this: the tests under tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py -- and only as the checker or the CPU
 * baseline, never as the product path.  The product (the HIP library behind
 * include/exahype_amd.h) never links, loads or calls anything in oracle/.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference root (ExaHyPE-Engine).
 *
 * Parity pins (what fixes each function independently of this code; the
 * test that checks it in brackets):
 *   tables          legendre_tables.json (values printed by the reference
 *                   generator, GaussLegendreBasis.csnippet) [test_product_cpu]
 *   euler flux      golden testPDEFluxes (2D/3D) [test_oracle_golden]
 *   stp_nonlinear   golden testSpaceTimePredictorNonlinear (2D/3D, uniform
 *                   state: layout + extrapolation) [test_oracle_golden]; the
 *                   contractions through the design-order convergence of the
 *                   3D step on the exact density wave [test_oracle_pins]
 *   volume integral the reference's own test is disabled; pinned by the exact
 *                   integral of phi_k' times polynomial fluxes of degree <= N+2,
 *                   a different polynomial per direction and variable and an
 *                   anisotropic dx (a transposed lFhi_y / lFhi_z block or a
 *                   swapped dx fails) [test_oracle_pins]
 *   riemann         golden testRiemannSolverNonlinear (2D/3D) [test_oracle_golden]
 *   face/surface    golden testSurfaceIntegralNonlinear (2D/3D) [test_oracle_golden]
 *   solution update golden testSolutionUpdate (2D/3D) [test_oracle_golden]
 *   stable dt       closed form CFL*PNPM/sum_d((|u_d|+c)/dx_d) on uniform
 *                   states with anisotropic dx, and the minimum of it over
 *                   the nodes of a non-uniform state [test_oracle_pins]
 *   run_uniform     (3D) observed L2 order >= N + 0.5 on the exact density
 *                   wave, 3^3 -> 6^3 cells, orders 3 and 5 [test_oracle_pins]
 *   linear kernels  golden 2D elasticity fixtures (testSpaceTimePredictorLinear,
 *                   testRiemannSolverLinear, testVolumeIntegralLinear,
 *                   testSurfaceIntegralLinear; disabled in the reference but
 *                   matching the generic kernels) + closed-form acoustic plane
 *                   wave (exact time average, exact upwind fluctuations)
 *                   [test_oracle_linear]
 *   point sources   moments against the Lagrange basis at the source point,
 *                   the Taylor terms of the time interpolant [test_oracle_linear];
 *                   the lFhi term sum_t w_t PSi[1+t] / D in every direction
 *                   [test_oracle_pins]
 *   lin_stp_split   the split Cauchy-Kowalevski template (no reference test
 *                   runs it): exact polynomial algebra of the levels, time
 *                   average (the template's factors dt^k/(2 k!)), volume
 *                   integral and faces for tensor-polynomial data; point-source
 *                   levels = derivatives of the unit-interval interpolant;
 *                   split - generic = the factor differences and the extra
 *                   level; the mesh is second order in time, the generic one
 *                   of design order [test_oracle_split_ck]
 *   linear adjoint  finite differences + dot-product (transpose) identity
 *                   [test_oracle_linear]
 *   amr projections constant preservation (the reference's
 *                   testFaceUnknownsProjection / testVolumeUnknownsProjection)
 *                   + exact reproduction of polynomials (prolongation) and
 *                   exact L2 projection (restriction) at level differences 1, 2
 *                   [test_oracle_amr]
 *   amr mesh step   free-stream preservation, conservation, reduction to the
 *                   uniform driver for an unrefined / fully refined mesh
 *                   [test_oracle_amr]
 *   boundary        wall: ghost = mirror state (normal momentum negated), zero
 *                   mass/energy wall flux for consistent data, the wall's
 *                   momentum flux from the closed-form Rusanov formula;
 *                   Dirichlet: Gauss time average vs adaptive quadrature of
 *                   the exact entropy wave; mesh: rest state in a walled box,
 *                   mirror symmetry of a walled box, reduction to the periodic
 *                   driver, and convergence to the exact entropy wave
 *                   [test_oracle_bc]
 *   cerk_guess /    the optimised kernels' initial guess and Picard rule
 *   picard rule     (no reference test runs them): the CERK expansion is the
 *                   exact space-time solution of linear advection with data
 *                   quadratic in space (one iteration, lQi = u(x - a xi dt));
 *                   constant states are fixed points; CERK and trivial guesses
 *                   converge to the same predictor; the trivial-guess arrays
 *                   equal the generic kernel's at the same iteration count;
 *                   the cap max(nDof - 3, 1) is reached on rough data
 *                   [test_oracle_dialect]
 */
#ifndef EXAHYPE_AMD_ORACLE_H
#define EXAHYPE_AMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXB 16

/* Operator tables for one order N (B = N+1 nodes), laid out as the
 * reference's kernels::legendre::*[order] arrays (row-major [i][j]). */
typedef struct orc_basis {
  int order;
  double w[ORC_MAXB];                 /* weights on [0,1] */
  double x[ORC_MAXB];                 /* Gauss-Legendre nodes on [0,1] */
  double Kxi[ORC_MAXB * ORC_MAXB];    /* Kxi[i][j] */
  double iK1[ORC_MAXB * ORC_MAXB];    /* iK1[i][j] */
  double dudx[ORC_MAXB * ORC_MAXB];   /* dudx[i][j] */
  double F0[ORC_MAXB];                /* FCoeff[order][0] = phi(0) */
  double F1[ORC_MAXB];                /* FCoeff[order][1] = phi(1) */
  double fgp[3 * ORC_MAXB * ORC_MAXB];/* fineGridProjector[order][sub][i][j] */
} orc_basis;

enum { ORC_PDE_EULER = 0, ORC_PDE_ACOUSTIC = 1, ORC_PDE_ADVECTION = 2, ORC_PDE_ELASTIC2D = 3,
       ORC_PDE_ADV_NCP_SRC = 4 };

typedef struct orc_solver {
  int dim;            /* 2 or 3 */
  int nvar;           /* NumberOfVariables */
  int npar;           /* NumberOfParameters */
  int pde;            /* ORC_PDE_* */
  int use_max_picard; /* UseMaxPicardIterations */
  int max_picard;     /* MaxPicardIterations */
  double cfl;         /* CFL */
  double pnpm;        /* PNPM */
  double adv[3];      /* advection velocity (ORC_PDE_ADVECTION only) */
  orc_basis b;
  /* optimised-kernel dialect (KernelGenerator/kernelgenerator/templates/aderdg/
   * fusedSPTVI_nonlinear_cpp.template): cerk_guess = Continuous Extension
   * Runge-Kutta initial guess (:121-173); opt_iterations = that kernel's
   * MaxIterations (:177: max(nDof-3, 1) with CERK, else 2 nDof + 1), tolerance
   * exit always on (:307-309), return min(iter + 1, MaxIterations) (:607). */
  int cerk_guess;
  int opt_iterations;
  /* ORC_PDE_ADV_NCP_SRC (test PDE with a non-conservative product and an
   * algebraic source, dQ/dt + div(a Q) + b . grad Q = -kappa Q per variable):
   * flux a Q (adv), ncp b . grad Q (ncp_b), source -kappa Q (kappa) */
  double ncp_b[3];
  double kappa;
  /* linear path: 1 = the split Cauchy-Kowalevski predictor of the optimised
   * kernels (fusedSPTVI_linear_split_ck_cpp.template) in orc_lin_run */
  int split_ck;
} orc_solver;

/* ---- PDE hooks ---------------------------------------------------------- */
void orc_flux(const orc_solver* s, const double* Q, double* F /* [dim][nvar] */);
void orc_eigenvalues(const orc_solver* s, const double* Q, int d, double* lambda);
/* nonConservativeProduct(Q, gradQ[dim][nvar], BgradQ) and algebraicSource(Q, S)
 * of ORC_PDE_ADV_NCP_SRC (the hooks the predictor calls at :349-362 and the
 * Rusanov solver at riemannSolverNonlinear.cpph:98-107) */
void orc_ncp(const orc_solver* s, const double* Q, const double* gradQ, double* BgradQ);
void orc_source(const orc_solver* s, const double* Q, double* S);

/* ---- nonlinear cell kernels (reference generic c/{2d,3d}) --------------- */
int orc_stp_nonlinear(const orc_solver* s, const double* luh, const double* invDx,
                      double dt, double* lQi, double* rhs, double* lFi,
                      double* lQhi, double* lFhi, double* lQhbnd, double* lFhbnd);
void orc_volume_integral_nonlinear(const orc_solver* s, double* lduh,
                                   const double* lFhi, const double* dx);
void orc_riemann_nonlinear(const orc_solver* s, double* FL, double* FR,
                           const double* QL, const double* QR, int direction);
void orc_face_integral_nonlinear(const orc_solver* s, double* lduh, const double* lFhbnd,
                                 int direction, int orientation, const double* dx);
void orc_surface_integral_nonlinear(const orc_solver* s, double* lduh,
                                    const double* lFhbnd, const double* dx);
void orc_solution_update(const orc_solver* s, double* luh, const double* luhOld,
                         const double* lduh, double dt);
double orc_stable_dt(const orc_solver* s, const double* luh, const double* dx);

/* Glue: fusedSpaceTimePredictorVolumeIntegral with addToUpdate == true. */
int orc_fused_stp_vi(const orc_solver* s, double* lduh, double* lQhbnd, double* lFhbnd,
                     const double* luh, const double* dx, double dt);

/* ---- uniform periodic mesh driver (time loop semantics) ----------------- */
/* luh: [ncells][B^D][nData], cells ordered x fastest.  Runs `steps` global
 * time steps in place.  dt_io: in = first dt (<=0: compute from luh),
 * out = dt of the next step.  picard_out (nullable): per-cell Picard
 * counts of the last step.  threads: OpenMP threads (<=0: default). */
int orc_run_uniform(const orc_solver* s, double* luh, int n_per_dim, double h,
                    int steps, double* dt_io, int* picard_out, int threads);

/* ---- linear kernels (aderdg_linear_oracle.c) ----------------------------- */
/* flags: 1 flux, 2 non-conservative product, 4 material matrix */
void orc_lin_stp(const orc_solver* s, int flags, const double* luh, const double* dx, double dt,
                 const double* PSi, double* lQhi, double* lFhi, double* lQbnd, double* lFbnd,
                 double* levels);
void orc_lin_volume(const orc_solver* s, double* lduh, const double* lFhi);
void orc_lin_face(const orc_solver* s, double* lduh, const double* lFhbnd, int direction,
                  int orientation, const double* dx);
void orc_lin_surface(const orc_solver* s, double* lduh, const double* lFhbnd, const double* dx);
void orc_lin_riemann(const orc_solver* s, int flags, double* FL, double* FR, const double* QL,
                     const double* QR, int direction);
double orc_ricker(double t, double f0, double t0);
/* split Cauchy-Kowalevski predictor + volume integral of the optimised linear
 * kernels (fusedSPTVI_linear_split_ck_cpp.template), its point-source moments
 * PSi[B][B^D][V] at the Gauss time nodes (:112-155) */
void orc_lin_stp_split(const orc_solver* s, int flags, const double* luh, const double* dx, double dt,
                       const double* PSi, double* lduh, double* lQhi_out, double* lQbnd, double* lFbnd);
void orc_lin_psi_split(const orc_solver* s, const double* centre, double h, double t, double dt,
                       const double* xs, double amp, double f0, double t0, double* PSi);
void orc_lin_psi(const orc_solver* s, const double* centre, double h, double t, double dt,
                 const double* xs, double amp, double f0, double t0, double* PSi);
int orc_lin_run(const orc_solver* s, double* luh, int n, double h, double dt, int nsteps,
                const double* src_x, double amp, double f0, double t0, int nrec, const double* xr,
                double* rec, double* states, int threads);
double orc_lin_misfit_grad(const orc_solver* s, const double* u0, int n, double h, double dt,
                           int nsteps, const double* src_x, double amp, double f0, double t0,
                           int nrec, const double* xr, const double* obs, double* gc, int threads);

/* ---- AMR projections and the static two-level mesh (aderdg_amr_oracle.c) -- */
void orc_face_prolongation(const orc_solver* s, double* lQf, double* lFf, const double* lQc, const double* lFc,
                           int level_coarse, int level_fine, const int* subface);
void orc_face_restriction(const orc_solver* s, double* lQc, double* lFc, const double* lQf, const double* lFf,
                          int level_coarse, int level_fine, const int* subface);
void orc_face_flux_restriction(const orc_solver* s, double* lFc, const double* lFf, const int* subface,
                               int level_delta);
void orc_volume_prolongation(const orc_solver* s, double* luhF, const double* luhC, int level_coarse,
                             int level_fine, const int* subcell);
void orc_volume_restriction(const orc_solver* s, double* luhC, const double* luhF, int level_coarse,
                            int level_fine, const int* subcell);
int orc_amr_nleaves(int n, const unsigned char* refined);
void orc_amr_leaves(int n, const unsigned char* refined, int* level, int* coords);
int orc_amr_run(const orc_solver* s, int n, const unsigned char* refined, double* luh, int steps, double* dt_io,
                int threads);

/* ---- boundary conditions (aderdg_bc_oracle.c) ---------------------------- */
enum { ORC_BC_PERIODIC = 0, ORC_BC_WALL = 1, ORC_BC_DIRICHLET_ENTROPY_WAVE = 2 };
void orc_entropy_wave(const orc_solver* s, const double* x, double t, double* Q);
void orc_boundary_values(const orc_solver* s, int kind, const double* x, double t, double dt, int faceIndex,
                         int direction, const double* FIn, const double* QIn, double* FOut, double* QOut);
void orc_boundary_conditions(const orc_solver* s, int kind, double* FIn, const double* QIn, const double* centre,
                             const double* dx, double t, double dt, int direction, int orientation);
int orc_run_bc(const orc_solver* s, double* luh, int n, double h, int steps, double* dt_io, double* t_io,
               const int* bc, int threads);
int orc_run_box(const orc_solver* s, double* luh, const int* ncell_d, double h, int steps, double* dt_io,
                double* t_io, const int* bc, int threads);

#ifdef __cplusplus
}
#endif
#endif

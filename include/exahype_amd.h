/*
 * exahype_amd -- MI355X-native ADER-DG cell-update engine, C ABI.
 *
 * Drop-in boundary for the reference's ADER-DG kernel plugin API
 * (exahype::solvers::ADERDGSolver pure virtuals, ExaHyPE/exahype/solvers/ADERDGSolver.h:2392-2660,
 * implemented per application by the generated Abstract<Solver>,
 * Toolkit/exahype/toolkit/templates/solvers/AbstractGenericADERDGSolverImplementation.template).
 *
 * Conventions (identical to the reference generic kernels):
 *   - all arrays are FP64, row-major, per cell:
 *       luh    [B^D][nVar+nPar]            (z,y,x, data fastest)
 *       lduh   [B^D][nVar]
 *       lQhbnd [2D][B^(D-1)][nVar+nPar]    face f = 2*dir + orientation,
 *       lFhbnd [2D][B^(D-1)][nVar]         x-faces (z,y), y-faces (z,x), z-faces (y,x)
 *   - B = order + 1 Gauss-Legendre nodes on [0,1] per dimension.
 *   - batched entry points take DEVICE pointers to arrays of cells stored back
 *     to back in that per-cell layout, and a hipStream_t passed as void*.
 *   - every function returns EXA_OK (0) or an error code; exa_last_error()
 *     returns a description of the last error of the calling thread.
 *     Nothing aborts the process (the reference calls std::abort /
 *     std::terminate instead, e.g. generic/c/3d/faceIntegralNonlinear.cpph:97-101).
 *   - a context is not thread-safe; use one context per host thread / stream.
 */
#ifndef EXAHYPE_AMD_H
#define EXAHYPE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum exa_status {
  EXA_OK = 0,
  EXA_ERR_ARG = 1,          /* invalid argument / size                         */
  EXA_ERR_HIP = 2,          /* HIP runtime error (launch, copy, allocation)    */
  EXA_ERR_UNSUPPORTED = 3,  /* combination not compiled into this library      */
  EXA_ERR_NOMEM = 4         /* device or host allocation failed                */
};

/* PDE hooks are compiled device functors (device code cannot call host
 * virtuals).  Reference hooks: flux / eigenvalues
 * (AbstractGenericADERDGSolverHeader.template:220,311-313). */
enum exa_pde {
  EXA_PDE_EULER = 0,      /* ApplicationExamples/Euler/Euler_ADERDG/EulerSolver_ADERDG.cpp:61-147 */
  EXA_PDE_ACOUSTIC = 1,   /* linear acoustics, 3D: p,u,v,w; parameters rho, c
                             (ApplicationExamples/Linear/LinearWave3D/MyLinearWaveSolver.cpp:108-209);
                             linear (Cauchy-Kowalevski) kernels, exa_acoustic_* below */
  EXA_PDE_ADVECTION = 2,  /* linear advection F_d = a_d Q, eigenvalues a_d (test PDE)          */
  EXA_PDE_USER = 3        /* the device hooks of struct ExaUserPde (include/exahype_amd_pde.h:
                             flux, eigenvalues, nonConservativeProduct, algebraicSource),
                             compiled into the library; 5 variables, no parameters.  With
                             a source or NCP: FP64, trivial guess, the four-kernel step,
                             periodic / exchange domain faces */
};

/* Arithmetic of the space-time predictor (Picard loop, time averages,
 * extrapolation, volume integral).  Riemann solve, face integrals, solution
 * update, Picard residual and CFL reduction always run in FP64. */
enum exa_precision {
  EXA_PREC_FP64 = 0,
  EXA_PREC_FP32 = 1,       /* FP32 contractions (MFMA where compiled), FP32 accumulation */
  EXA_PREC_BF16 = 2        /* experiment: the contractions' operands (operator entries,
                              flux and divergence values) rounded to BF16 on the BF16
                              matrix cores, FP32 accumulation, FP32 iterate; 3D Euler
                              order 5 with a fixed Picard count and the trivial guess */
};

typedef struct exa_ctx exa_ctx;

/* Create a solver context.
 *   dim          2 or 3
 *   order        N (B = N+1 nodes per dimension)
 *   nvar, npar   NumberOfVariables, NumberOfParameters
 *   pde          exa_pde
 *   precision    exa_precision
 *   picard_max   <= 0: adaptive Picard, tol 1e-7 on sqrt(sum of squared
 *                updates), at most 2(N+1) iterations (UseMaxPicardIterations
 *                = false, generic/c/3d/spaceTimePredictorNonlinear.cpph:84-86,407-423);
 *                > 0: exactly picard_max iterations (UseMaxPicardIterations
 *                = true, MaxPicardIterations = picard_max).
 *   cfl          CFL factor (reference default 0.9)
 * The PNPM factor follows the reference defaults per order
 * (Toolkit/exahype/toolkit/solverController.py:17-23). */
int exa_ctx_create(exa_ctx** out, int dim, int order, int nvar, int npar, int pde, int precision,
                   int picard_max, double cfl);
int exa_ctx_destroy(exa_ctx* ctx);
const char* exa_last_error(void);

/* Point source of the linear path's per-cell predictor (getPointSources /
 * deltaDistribution, generic/c/3d/deltaDistribution.cpph:37-146): pressure
 * source at x[3] with the Ricker time function amp (1 - 2a) exp(-a),
 * a = (pi f0 (t - t0))^2; x == NULL or amp == 0 removes it. */
int exa_ctx_set_point_source(exa_ctx* ctx, const double* x, double amp, double f0, double t0);

/* Advection velocity for EXA_PDE_ADVECTION (a[dim]). */
int exa_ctx_set_advection(exa_ctx* ctx, const double* a);
/* Parameters of the EXA_PDE_USER hooks (prm[0..n), n <= 16, the rest zero);
 * the hooks read them on the device.  Replaces the constants a user solver
 * keeps as members (AbstractGenericADERDGSolverHeader.template). */
int exa_ctx_set_pde_params(exa_ctx* ctx, const double* prm, int n);

/* Predictor kernel selection: 0 (default) picks the fastest kernel compiled
 * for the shape, 1 forces the generic node-column kernel (every shape); for
 * 3D Euler order 5 also 2 = x-line kernel (VALU), 3 = x-line kernel on 4x4x1
 * matrix-core steps (FP32), 4 = space-time tile kernel on 16x16x4 matrix-core
 * steps (FP32).  Meshes copy the context's settings when they are created. */
int exa_ctx_set_predictor_variant(exa_ctx* ctx, int variant);

/* Kernel dialect.  The reference generates a second, "optimised" set of
 * kernels (KernelGenerator/kernelgenerator/templates/aderdg/; selected by
 * optimisation=optimised in a specification file) whose nonlinear predictor
 * differs from the generic one in its Picard rule
 * (fusedSPTVI_nonlinear_cpp.template:177,307-309,607) and, optionally, its
 * initial guess (CERK, :121-173), and whose arrays pad the variable dimension
 * to a multiple of the vector width (controller.py:166-195 getSizeWithPadding).
 *   EXA_DIALECT_GENERIC        the generic kernels (default)
 *   EXA_DIALECT_OPTIMISED      MaxIterations = 2 nDof + 1, tolerance exit
 *                              always on, count = min(iter + 1, MaxIterations)
 *   EXA_DIALECT_OPTIMISED_CERK the same with the Continuous Extension
 *                              Runge-Kutta initial guess and MaxIterations =
 *                              max(nDof - 3, 1)
 *   EXA_DIALECT_SPLIT_CK       linear contexts only: the split Cauchy-Kowalevski
 *                              predictor + volume integral of the optimised
 *                              linear kernels (fusedSPTVI_linear_split_ck_cpp.template;
 *                              space_time_predictor split_ck in a specification
 *                              file, solverController.py:221-222) in
 *                              exa_fusedSpaceTimePredictorVolumeIntegral and the
 *                              acoustic mesh (x-line form): one derivative level
 *                              kept, B levels in the time average with the
 *                              template's factors (dt/2, then x dt/(t+2) from
 *                              t = 0: dt^k/(2 k!), not the Taylor dt^k/(k+1)!),
 *                              the material matrix on the flux part only, point
 *                              sources at the Gauss time nodes.  Riemann, face
 *                              integral and update stay generic; no adjoint.
 * vect_size (1, 2, 4 or 8 doubles: the generator's noarch, wsm, snb/hsw,
 * knl/skx widths) sets nVarPad and nDataPad of the exa_opt_* and
 * exa_converter_* entry points.  The optimised Picard dialects run on the
 * node-column predictor kernel.  Meshes copy the context's settings when
 * created. */
enum exa_dialect {
  EXA_DIALECT_GENERIC = 0,
  EXA_DIALECT_OPTIMISED = 1,
  EXA_DIALECT_OPTIMISED_CERK = 2,
  EXA_DIALECT_SPLIT_CK = 3
};
int exa_ctx_set_dialect(exa_ctx* ctx, int dialect, int vect_size);

/* Initial guess of the nonlinear Picard loop, independent of the Picard rule:
 * EXA_GUESS_TRIVIAL (luh at every time node, the generic kernels,
 * spaceTimePredictorNonlinear.cpph:71-82) or EXA_GUESS_CERK (the optimised
 * kernels' Continuous Extension Runge-Kutta guess,
 * fusedSPTVI_nonlinear_cpp.template:121-173).  exa_ctx_set_dialect sets it
 * too (CERK for EXA_DIALECT_OPTIMISED_CERK); a later call overrides.  With the
 * generic rule the CERK guess runs on the node-column kernel, and for 3D Euler
 * order 5 in FP32 with a fixed Picard count on the x-line matrix-core kernel. */
enum exa_guess { EXA_GUESS_TRIVIAL = 0, EXA_GUESS_CERK = 1 };
int exa_ctx_set_initial_guess(exa_ctx* ctx, int guess);

/* Picard guard of the FP32 one-iteration CERK form (FP32, fixed count 1,
 * EXA_GUESS_CERK; 3D Euler order 5).  That form computes no full residual: the
 * last iteration is collapsed into its time average.  Per cell it forms, with
 * FP64 accumulation, the relative proxy residual
 *   r = |lQhi(1) - lQhi(0)| / |lQhi(1)|   (2-norms over nodes and variables),
 * lQhi(k) = sum_t w_t lQi(k)(t) the time average of iterate k (lQi(0) the CERK
 * guess), and recomputes a cell with r >= tol with recompute_count fixed
 * iterations (0: max(nDof - 3, 1), the optimised kernels' CERK cap,
 * fusedSPTVI_nonlinear_cpp.template:177); that cell reports the count, the
 * others keep one.  The reference's rule is absolute, |lQi(1) - lQi(0)| < 1e-7
 * (spaceTimePredictorNonlinear.cpph:370-423); FP32 iterates cannot resolve it
 * (the flux divergences carry ~1e-8 relative rounding, ~1e-6 absolute per
 * cell), so the guard bounds the first iteration's relative change instead:
 * stopping there leaves an error of about the contraction factor times r.
 * Default: on, tol = 1e-6 (the FP32 form's accuracy target), recompute_count 0;
 * tol <= 0 switches it off.  No counterpart in the reference.  Meshes copy the
 * setting at creation. */
int exa_ctx_set_picard_guard(exa_ctx* ctx, double tol, int recompute_count);
/* *size = nVarPad (which = 0) or nDataPad (which = 1) of the context's dialect. */
int exa_ctx_padded_size(const exa_ctx* ctx, int which, int* size);

/* Operator tables of order N derived on the host (no device needed; each
 * output may be NULL).  Same layout as exa_ctx_tables plus
 * egp[B*B] = equidistantGridProjector. */
int exa_tables(int order, double* w, double* x, double* Kxi, double* iK1, double* dudx, double* F0, double* F1,
               double* fgp, double* egp);

/* Copy out the derived operator tables (each may be NULL):
 * w[B], x[B], Kxi[B*B], iK1[B*B], dudx[B*B], F0[B], F1[B], fgp[3*B*B]. */
int exa_ctx_tables(const exa_ctx* ctx, double* w, double* x, double* Kxi, double* iK1, double* dudx,
                   double* F0, double* F1, double* fgp);

/* ---------------------------------------------------------------------------
 * Per-cell compatibility entry points: host pointers, the reference argument
 * order and semantics.  They copy to the device, run the batched kernels on
 * one cell and copy back (synchronous).  For drop-in use and parity tests.
 * ------------------------------------------------------------------------- */

/* fusedSpaceTimePredictorVolumeIntegral
 * (AbstractGenericADERDGSolverImplementation.template:104-235; API ADERDGSolver.h:2542).
 * lGradQhbnd is unused (no viscous flux) and may be NULL.  centre is unused
 * by the compiled PDEs (no algebraic source).  Returns the Picard count in
 * *picard_iterations with the reference's per-dimension convention
 * (3D: 0-based index of the converged iteration, 2D: MaxIterations+2 on
 * convergence; MaxIterations when not converged). */
/* For an EXA_PDE_ACOUSTIC context the per-cell entry points run the linear
 * kernels (spaceTimePredictorLinear + volumeIntegralLinear with the context's
 * point source, riemannSolverLinear with FL/FR as outputs, faceIntegralLinear
 * for conforming faces added to the update, solutionUpdate, stableTimeStepSize
 * with eigenvalues (c, -c, 0, 0)); parameters may vary per node; the Picard
 * count is 0. */
int exa_fusedSpaceTimePredictorVolumeIntegral(exa_ctx* ctx, double* lduh, double* lQhbnd, double* lGradQhbnd,
                                              double* lFhbnd, double* luh, const double* centre,
                                              const double* dx, double t, double dt,
                                              int addVolumeIntegralResultToUpdate, int* picard_iterations);

/* fusedSpaceTimePredictorVolumeIntegral of the optimised kernels
 * (KernelGenerator/kernelgenerator/templates/aderdg/fusedSPTVI_nonlinear_cpp.template:27-47,
 * nonlinear, no source/NCP/viscous flux), context in an EXA_DIALECT_OPTIMISED*
 * dialect.  Padded layouts: lduh[B^D][nVarPad], lQhbnd[2D][B^(D-1)][nDataPad],
 * lFhbnd[2D][B^(D-1)][nVarPad] (padding entries written as 0, as the kernel
 * zeroes its outputs first); luh[B^D][nData] unpadded.  Cubic cells
 * (inverseDx = 1/dx for every direction).  The scratch arrays the generated
 * kernel takes (lQi, rhs, lFi, ...) live on the device and are not arguments.
 * *picard_iterations = min(iter + 1, MaxIterations) (:607). */
int exa_opt_fusedSpaceTimePredictorVolumeIntegral(exa_ctx* ctx, double* lduh, double* lQhbnd, double* lFhbnd,
                                                  const double* luh, double inverseDx, double dt,
                                                  int* picard_iterations);

/* Layout converters between the generic kernels' arrays and the optimised
 * kernels' padded arrays (KernelGenerator/kernelgenerator/templates/aderdg/
 * converter_cpp.template:16-173, same names; sizes from the context's shape
 * and exa_ctx_set_dialect's vect_size).  Padding entries of an optimised
 * output are left untouched, as in the reference.
 *   Qhbnd [2D][B^(D-1)][nData|nDataPad]   Fhbnd [2D][B^(D-1)][nVar|nVarPad]
 *   QFace [B^(D-1)][nData|nDataPad]       FFace [B^(D-1)][nVar|nVarPad]
 *   Qhi   [B^D][nData|nDataPad]           Fhi   [D+1][B^D][nVar|nVarPad]
 *   PSi   [B+1][B^D][nVar|nVarPad]        Lduh  [B^D][nVar|nVarPad] */
int exa_converter_Qhbnd_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_Qhbnd_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_Fhbnd_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_Fhbnd_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_QFace_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_QFace_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_FFace_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_FFace_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_Qhi_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_Qhi_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_Fhi_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_Fhi_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_PSi_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_PSi_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);
int exa_converter_Lduh_optimised2generic(const exa_ctx* ctx, const double* opt, double* gen);
int exa_converter_Lduh_generic2optimised(const exa_ctx* ctx, const double* gen, double* opt);

/* riemannSolver (…Implementation.template:287-305 -> generic/c/3d/riemannSolverNonlinear.cpph:37-128).
 * FL/FR are overwritten in place with the Rusanov flux. */
int exa_riemannSolver(exa_ctx* ctx, double* FL, double* FR, const double* QL, const double* QR, double t,
                      double dt, const double* dx, int direction, int isBoundaryFace, int faceIndex);

/* Boundary kinds of a domain face (Euler).  The two non-periodic kinds are the
 * boundaryValues hooks of the reference's Euler example
 * (ApplicationExamples/Euler/Euler/EulerSolver_ADERDG.cpp:525-554):
 * EXA_BC_WALL reflecting wall (outside state = inside state with the normal
 * momentum negated, outside flux = its Euler flux; the example's choice for
 * SodShockTube / SphericalExplosion / RarefactionWave);
 * EXA_BC_DIRICHLET_ENTROPY_WAVE outside state and flux = Gauss time average
 * over [t, t + dt] of the exact entropy wave (:266-286; rho = 1/2 + 3/10
 * exp(-|x - x0 - v0 t| / 0.3^D), x0 = (1/2, 1/2, 1/2), v0 = (1/2, 0, 0),
 * p = 1; the example's default). */
enum exa_boundary_kind {
  EXA_BC_PERIODIC = 0,
  EXA_BC_WALL = 1,
  EXA_BC_DIRICHLET_ENTROPY_WAVE = 2,
  EXA_BC_EXCHANGE = 3 /* mesh faces only: outside data from a halo message (exa_mesh_halo_unpack) */
};

/* The boundaryValues hook used by exa_boundaryConditions (default
 * EXA_BC_DIRICHLET_ENTROPY_WAVE, the example's default). */
int exa_ctx_set_boundary_values(exa_ctx* ctx, int kind);

/* boundaryConditions (…Implementation.template:307-350 -> generic/c/3d/boundaryConditions.cpph:29-122,
 * caller ADERDGSolver.cpp:1836-1900).  fluxIn [B^(D-1)][V] (in: the cell's
 * extrapolated flux on face 2*direction+orientation; out: the Riemann flux),
 * stateIn [B^(D-1)][nData] the extrapolated state; the outside data come from
 * the context's boundaryValues hook at the face nodes; the Riemann solve puts
 * them on the left of a left face (orientation 0) and on the right of a right
 * face.  gradStateIn and luh are unused (no viscous flux) and may be NULL. */
int exa_boundaryConditions(exa_ctx* ctx, double* fluxIn, const double* stateIn, const double* gradStateIn,
                           const double* luh, const double* cellCentre, const double* cellSize, double t, double dt,
                           int direction, int orientation);

/* faceIntegral (…Implementation.template:247-285 -> generic/c/3d/faceIntegralNonlinear.cpph:21-103).
 * levelDelta == 0: addToUpdate != 0: out is lduh and the face integral is
 * accumulated into it; addToUpdate == 0: out is luh and out += dt/w *
 * faceIntegral (as the reference glue).  levelDelta > 0 (a fine face seen from
 * a coarser cell, subfaceIndex[dim-1] in 0..3^levelDelta-1): the flux is
 * restricted over levelDelta levels (faceUnknownsRestriction) and integrated
 * with the coarse size 3^levelDelta dx into out (an update, as the reference
 * glue :265-270).  subfaceIndex may be NULL when levelDelta == 0. */
int exa_faceIntegral(exa_ctx* ctx, double* out, const double* lFhbnd, int direction, int orientation,
                     const int* subfaceIndex, int levelDelta, const double* dx, double dt, int addToUpdate);

/* surfaceIntegralNonlinear (generic/c/3d/surfaceIntegralNonlinear.cpph:21-90): all 2D faces. */
int exa_surfaceIntegral(exa_ctx* ctx, double* lduh, const double* lFhbnd, const double* dx);

/* addUpdateToSolution -> solutionUpdate (…Implementation.template:237-245,
 * generic/c/3d/solutionUpdate.cpph:28-67).  luh may alias luhOld. */
int exa_addUpdateToSolution(exa_ctx* ctx, double* luh, const double* luhOld, const double* lduh, double dt);

/* stableTimeStepSize (…Implementation.template:352-361, generic/c/3d/stableTimeStepSize.cpph:14-62). */
int exa_stableTimeStepSize(exa_ctx* ctx, const double* luh, const double* dx, double* dt_out);

/* ---------------------------------------------------------------------------
 * Batched device entry points (device pointers, cells back to back).
 * h: per-cell edge length (cubic cells), device array [ncells]; dt is read
 * from the device scalar d_dt so that a captured step needs no host value.
 * ------------------------------------------------------------------------- */

/* Cell extents of the batched entry points: per_direction = 0 (default): d_h
 * holds one edge length per cell [ncells]; 1: d_h holds the reference's dx
 * vector per cell [ncells][dim] (ADERDGSolver.h:2542-2553 passes dx[D]), used
 * per direction by the predictor (node-column kernel), the face integrals and
 * the stable dt.  The FP32 line and BF16 tile kernels take cubic cells: with 1
 * an FP32 context runs the FP32 node-column kernel and BF16 is refused. */
int exa_ctx_set_cell_dx(exa_ctx* ctx, int per_direction);

/* Predictor + volume integral for ncells cells (reference row: a4 + a5).
 * d_picard (nullable): per-cell Picard count.  With the Picard guard of the
 * FP32 one-iteration CERK form (exa_ctx_set_picard_guard) the cell list of
 * the recomputation lives in a context buffer that the first call of a given
 * size allocates. */
int exa_batch_predictor(exa_ctx* ctx, void* stream, long ncells, const double* d_luh, const double* d_h,
                        const double* d_dt, double* d_lduh, double* d_lQhbnd, double* d_lFhbnd, int* d_picard);

/* exa_batch_predictor without allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results as exa_batch_predictor; the Picard guard's cell list is
 * a static device array of 2 + 1048576 ints (4 MiB per GPU, held until the
 * process ends), so ncells > 1048576 is EXA_ERR_ARG.  Calls on one GPU share
 * that array: calls that overlap on one GPU (concurrent streams) race. */
int exa_batch_predictor_noalloc(exa_ctx* ctx, void* stream, long ncells, const double* d_luh, const double* d_h,
                                const double* d_dt, double* d_lduh, double* d_lQhbnd, double* d_lFhbnd,
                                int* d_picard);

/* Rusanov flux on nfaces faces.  d_faces: int[4] per face = {left cell,
 * right cell, direction, unused}; reads left face 2*dir+1 and right face
 * 2*dir of lQhbnd/lFhbnd and writes the flux into both lFhbnd faces. */
int exa_batch_riemann(exa_ctx* ctx, void* stream, long nfaces, const int* d_faces, const double* d_lQhbnd,
                      double* d_lFhbnd);

/* Face integrals of all 2D faces into lduh, then luh += dt/w * lduh, then the
 * cell's admissible dt is min-reduced into *d_dt_next (IEEE order-independent:
 * the result does not depend on the reduction order).  d_dt_next must hold a
 * value >= every cell's dt (e.g. DBL_MAX) before the call. */
int exa_batch_corrector(exa_ctx* ctx, void* stream, long ncells, double* d_luh, double* d_lduh,
                        const double* d_lFhbnd, const double* d_h, const double* d_dt, double* d_dt_next);

/* min over cells of stableTimeStepSize(luh) min-reduced into *d_dt_out. */
int exa_batch_stable_dt(exa_ctx* ctx, void* stream, long ncells, const double* d_luh, const double* d_h,
                        double* d_dt_out);

/* ---------------------------------------------------------------------------
 * Host mesh driver (replaces the Peano traversal for the path): a uniform
 * Cartesian grid of n^D cubic cells of edge h, periodic, one global time
 * step per step (ADERDGSolver.cpp step order: predictor, Riemann, surface
 * integral + update, global dt reduction).  Device-resident patch arrays.
 * ------------------------------------------------------------------------- */
typedef struct exa_mesh exa_mesh;

int exa_mesh_create_uniform(exa_mesh** out, exa_ctx* ctx, int n, double h);
int exa_mesh_destroy(exa_mesh* mesh);
long exa_mesh_ncells(const exa_mesh* mesh);
/* Host <-> device copies of the solution [ncells][B^D][nData], cells x fastest. */
int exa_mesh_set_solution(exa_mesh* mesh, const double* luh);
int exa_mesh_get_solution(exa_mesh* mesh, double* luh);
/* Device pointer of the resident solution (for zero-copy consumers). */
double* exa_mesh_device_solution(exa_mesh* mesh);

/* Initial conditions evaluated on the device at every cell's nodes, the
 * reference's adjustSolution at t = 0 (generic/c/3d/solutionAdjustment.cpph:18-63).
 * EXA_IC_DENSITY_WAVE (Euler): params = {a0, u0, v0, w0, p0, M, amp,
 *   M x (kx, ky, kz, phase) for rho, M x (kx, ky, kz, phase) for u};
 *   rho = 1 + a0 sin(2 pi (x+y[+z])) + amp sum cos(2 pi k.x + phase),
 *   u = u0 + amp sum cos(...), v = v0, w = w0 (0 in 2D), p = p0.
 * EXA_IC_ENTROPY_WAVE (Euler): params = {t}; the exact entropy wave of
 *   EXA_BC_DIRICHLET_ENTROPY_WAVE at time t (the example's initial condition
 *   at t = 0, EulerSolver_ADERDG.cpp:266-286). */
enum exa_initial_condition { EXA_IC_DENSITY_WAVE = 1, EXA_IC_ENTROPY_WAVE = 2 };
int exa_mesh_adjust_solution(exa_mesh* mesh, int ic, const double* params, int nparams);
/* dt <= 0: compute the first dt from the current solution (stableTimeStepSize). */
int exa_mesh_init_dt(exa_mesh* mesh, double dt);
/* Advance nsteps global time steps on the mesh's stream.  use_graph != 0
 * replays a captured hipGraph of one step. Asynchronous; see exa_mesh_sync. */
int exa_mesh_run(exa_mesh* mesh, int nsteps, int use_graph);
int exa_mesh_sync(exa_mesh* mesh);
/* Current time, dt of the next step, per-cell Picard counts of the last step. */
int exa_mesh_time(exa_mesh* mesh, double* t, double* dt);
int exa_mesh_picard(exa_mesh* mesh, int* counts);
/* Cells the Picard guard recomputed since the mesh was created (see
 * exa_ctx_set_picard_guard; no counterpart in the reference). */
int exa_mesh_picard_guard(exa_mesh* mesh, long* recomputed);
/* Step form: enabled != 0 (default where compiled) fuses the step into two
 * kernels -- the predictor adds the volume part of the update to luh, then one
 * kernel solves every cell's face Riemann problems from both sides' data and
 * applies the face part and the CFL reduction; 0 runs the reference's four
 * phases (predictor into lduh, Riemann per face, face integrals + update, dt)
 * as separate kernels.  Results agree to rounding. */
int exa_mesh_set_fused(exa_mesh* mesh, int enabled);
/* Slab-pipelined fused step (no counterpart in the reference: a schedule of
 * this library's two step kernels).  slabs >= 3 splits the mesh into slabs of
 * z planes; the predictors run slab by slab on the mesh stream and the fused
 * corrector of a slab runs on a second stream once the predictors of its
 * neighbour slabs are done, beside the predictors of the later slabs.  The
 * results equal the one-launch step's bit for bit.  Applies to 3D fused steps
 * with all faces periodic and the one-iteration FP32 CERK predictor (else
 * EXA_ERR_UNSUPPORTED), 3 <= slabs <= n; 0 switches it off (default).  Runs
 * with timing enabled (exa_mesh_set_timing) and the split-step entry points
 * use the one-launch step. */
int exa_mesh_set_overlap(exa_mesh* mesh, int slabs);
/* Boundary kind per domain face kinds[2*d + o] (o = 0 at x_d = 0, o = 1 at
 * x_d = n h), EXA_BC_*; default all periodic.  A direction is periodic on both
 * faces or on neither.  Non-periodic faces follow the reference's
 * applyBoundaryConditions (ADERDGSolver.cpp:1836-1900) at the step's t and dt:
 * after the predictor, the boundaryValues hook writes each boundary face's
 * outside data into a ghost face, and the Riemann solve uses it as the
 * neighbour.  Euler only. */
int exa_mesh_set_boundary(exa_mesh* mesh, const int* kinds);
/* Set the mesh time t (the Dirichlet data depend on it); exa_mesh_init_dt
 * resets it to 0. */
int exa_mesh_set_time(exa_mesh* mesh, double t);
/* Halo exchange of a partitioned domain (the reference's neighbour messages,
 * ADERDGSolver.cpp:1946-2096: per face of a cell the extrapolated predictor
 * lQhbnd [B^(D-1)][nData], then the fluctuation lFhbnd [B^(D-1)][V]).  A face of
 * kind EXA_BC_EXCHANGE takes its outside data from a message: per step call
 * exa_mesh_step_predict, pack the faces the neighbours need
 * (exa_mesh_halo_pack: the lQhbnd message [n^(D-1) cells][B^(D-1)][nData] and
 * the lFhbnd message [n^(D-1)][B^(D-1)][V] of domain face `face`, cells in
 * transverse order (iz,iy) / (iz,ix) / (iy,ix), 2D iy / ix), send them,
 * exa_mesh_halo_unpack the received ones (device pointers, same layout, the
 * neighbour's opposite face), exa_mesh_step_correct, reduce dt_next (device
 * state {t, dt, dt_next}) to the global minimum, exa_mesh_step_advance.  All
 * enqueue on the mesh stream; exa_mesh_run does the three parts per step. */
int exa_mesh_step_predict(exa_mesh* mesh);
int exa_mesh_step_correct(exa_mesh* mesh);
int exa_mesh_step_advance(exa_mesh* mesh);
int exa_mesh_halo_size(const exa_mesh* mesh, long* n_q, long* n_f);
int exa_mesh_halo_pack(exa_mesh* mesh, int face, double* d_msg_q, double* d_msg_f);
int exa_mesh_halo_unpack(exa_mesh* mesh, int face, const double* d_msg_q, const double* d_msg_f);
/* Host-staged forms (host messages, as the reference's MPI messages; synchronous). */
int exa_mesh_halo_pack_host(exa_mesh* mesh, int face, double* msg_q, double* msg_f);
int exa_mesh_halo_unpack_host(exa_mesh* mesh, int face, const double* msg_q, const double* msg_f);
double* exa_mesh_device_state(exa_mesh* mesh);
/* Host access to dt_next between correct and advance (synchronous). */
int exa_mesh_dt_next(exa_mesh* mesh, double* dt_next);
int exa_mesh_set_dt_next(exa_mesh* mesh, double dt_next);
/* Point probes (the reference's probe plotter, plotters/ADERDG2ProbeAscii.cpp:127-180 with
 * kernels::legendre::interpolate, GaussLegendreBasis.cpp:21-49): values[npoints][nData] of the
 * current solution at points[npoints][dim], each inside [0, n h)^dim, interpolated with the
 * cell's Lagrange basis in the cell with offset <= x < offset + h.  Synchronous. */
int exa_mesh_probe(exa_mesh* mesh, int npoints, const double* points, double* values);
/* Global observables of the current solution (the solver template's reset / map / merge cycle,
 * Toolkit/exahype/toolkit/templates/solvers/AbstractGenericADERDGSolverImplementation.template:412-453,
 * with the Euler example's dgcells / fvcells counts, ApplicationExamples/Euler/Euler/
 * EulerSolver_ADERDG.cpp:607-628): out[0] = dgcells, out[1] = fvcells (0: no FV limiter),
 * out[2 + k] = integral of quantity k over the domain (Gauss weights times h^dim),
 * out[2 + nData + k] = nodal minimum, out[2 + 2 nData + k] = nodal maximum; nout >= 2 + 3 nData.
 * Reduction order is fixed (bitwise reproducible).  Synchronous. */
int exa_mesh_global_observables(exa_mesh* mesh, double* out, int nout);
/* The mesh's HIP stream (hipStream_t as void*). */
void* exa_mesh_stream(exa_mesh* mesh);
/* Per-phase timing of the last exa_mesh_run with timing enabled: milliseconds
 * summed over the run's steps for predictor, riemann, corrector; *nsteps. */
int exa_mesh_set_timing(exa_mesh* mesh, int enabled);
int exa_mesh_phase_times(exa_mesh* mesh, double* ms_predictor, double* ms_riemann, double* ms_corrector,
                         int* nsteps);

/* ---------------------------------------------------------------------------
 * AMR projections between a cell of level l and its descendants at level
 * l + delta (3^D children per level), reference generic/c/{2d,3d}/amrRoutines.cpph.
 * Host arrays per cell, layouts as lQhbnd face / luh.  Prolongations overwrite
 * the fine array; restrictions accumulate into the coarse array (as the
 * reference).  subfaceIndex[dim-1] / subcellIndex[dim] in 0..3^delta-1.
 * ------------------------------------------------------------------------- */
/* faceUnknownsProlongation (3d/amrRoutines.cpph:43-143) */
int exa_faceUnknownsProlongation(exa_ctx* ctx, double* lQhbndFine, double* lFhbndFine, const double* lQhbndCoarse,
                                 const double* lFhbndCoarse, int coarseGridLevel, int fineGridLevel,
                                 const int* subfaceIndex);
/* faceUnknownsRestriction (3d/amrRoutines.cpph:248-314) */
int exa_faceUnknownsRestriction(exa_ctx* ctx, double* lQhbndCoarse, double* lFhbndCoarse, const double* lQhbndFine,
                                const double* lFhbndFine, int coarseGridLevel, int fineGridLevel,
                                const int* subfaceIndex);
/* volumeUnknownsProlongation (3d/amrRoutines.cpph:364-419) */
int exa_volumeUnknownsProlongation(exa_ctx* ctx, double* luhFine, const double* luhCoarse, int coarseGridLevel,
                                   int fineGridLevel, const int* subcellIndex);
/* volumeUnknownsRestriction (3d/amrRoutines.cpph:483-554) */
int exa_volumeUnknownsRestriction(exa_ctx* ctx, double* luhCoarse, const double* luhFine, int coarseGridLevel,
                                  int fineGridLevel, const int* subcellIndex);

/* ---------------------------------------------------------------------------
 * Static two-level mesh (3D, nonlinear path): n^3 coarse cells on the
 * periodic unit cube; every coarse cell with refined[c] != 0 (c = (z*n+y)*n+x)
 * is split into 3^3 children.  Leaves are ordered by coarse cell, a refined
 * cell contributing its 27 children (x fastest).  One global dt (min of the
 * leaves' stableTimeStepSize with their own cell size).  Coarse-fine faces
 * follow the reference's virtual-cell coupling: the coarse face data are
 * prolongated to the 9 subfaces (prolongateFaceDataToVirtualCell,
 * ADERDGSolver.cpp:1340-1399), the Riemann problems are solved on the fine
 * subfaces, and the restricted fluctuations are integrated over the coarse
 * face with the coarse cell size (restrictToTopMostParent, :1472-1563).
 * ------------------------------------------------------------------------- */
typedef struct exa_amr exa_amr;

int exa_amr_create(exa_amr** out, exa_ctx* ctx, int n, const unsigned char* refined);
int exa_amr_destroy(exa_amr* mesh);
long exa_amr_nleaves(const exa_amr* mesh);
/* level[nleaves] (0 coarse, 1 fine) and integer coordinates coords[nleaves][3]
 * on the leaf's level (either may be NULL). */
int exa_amr_leaves(exa_amr* mesh, int* level, int* coords);
int exa_amr_set_solution(exa_amr* mesh, const double* luh);  /* [nleaves][B^3][nData] */
int exa_amr_get_solution(exa_amr* mesh, double* luh);
int exa_amr_init_dt(exa_amr* mesh, double dt);  /* dt <= 0: from the state */
int exa_amr_run(exa_amr* mesh, int nsteps, int use_graph);
int exa_amr_sync(exa_amr* mesh);
int exa_amr_time(exa_amr* mesh, double* t, double* dt);
/* per-leaf Picard counts of the last step (conventions of exa_mesh_picard) */
int exa_amr_picard(exa_amr* mesh, int* counts);
/* Autotuner: times the candidate launch geometries on the mesh's own data and
 * keeps the fastest: cfg[3] = {predictor grid (0 = one workgroup per tile,
 * else persistent workgroups), Riemann faces per workgroup, corrector grid};
 * ms[3] = their times.  The state is restored afterwards. */
int exa_amr_autotune(exa_amr* mesh, int* cfg, double* ms);

/* ---------------------------------------------------------------------------
 * Ensemble: ninst independent uniform periodic meshes (n[i] cells per side on
 * the unit square / cube), each with its own dt, stepped together: every
 * kernel is one launch over the cells of all instances.
 * ------------------------------------------------------------------------- */
typedef struct exa_ensemble exa_ensemble;

int exa_ensemble_create(exa_ensemble** out, exa_ctx* ctx, int ninst, const int* n);
int exa_ensemble_destroy(exa_ensemble* ens);
/* cells of instance inst (inst < 0: all instances) */
long exa_ensemble_ncells(const exa_ensemble* ens, int inst);
int exa_ensemble_set_solution(exa_ensemble* ens, int inst, const double* luh);
int exa_ensemble_get_solution(exa_ensemble* ens, int inst, double* luh);
int exa_ensemble_init_dt(exa_ensemble* ens, double dt);  /* dt <= 0: per instance from its state */
int exa_ensemble_run(exa_ensemble* ens, int nsteps, int use_graph);
int exa_ensemble_sync(exa_ensemble* ens);
int exa_ensemble_time(exa_ensemble* ens, int inst, double* t, double* dt);
/* per-cell Picard counts of the last step, all instances */
int exa_ensemble_picard(exa_ensemble* ens, int* counts);
int exa_ensemble_autotune(exa_ensemble* ens, int* cfg, double* ms);

/* ---------------------------------------------------------------------------
 * Group: step two-level meshes and ensembles together.  use_graph != 0
 * captures one step of every member as one hipGraph (members on parallel
 * branches) and replays it once per step.  Synchronous.  Re-tuning a member
 * after the first captured run needs exa_group_reset.
 * ------------------------------------------------------------------------- */
typedef struct exa_group exa_group;

int exa_group_create(exa_group** out, int namr, exa_amr* const* amrs, int nens, exa_ensemble* const* ens);
int exa_group_destroy(exa_group* group);
int exa_group_run(exa_group* group, int nsteps, int use_graph);
int exa_group_reset(exa_group* group);

/* ---------------------------------------------------------------------------
 * Linear acoustics on a periodic n^3 mesh (context: dim 3, nvar 4, npar 2,
 * EXA_PDE_ACOUSTIC, EXA_PREC_FP64, orders 1..5), FP64.  One step is the
 * reference's linear ADER-DG step: spaceTimePredictorLinear + volumeIntegralLinear
 * (generic/c/3d/spaceTimePredictorLinear.cpph:30-647, volumeIntegralLinear.cpph:21-70),
 * riemannSolverLinear (riemannSolverLinear.cpph:23-167), faceIntegralLinear
 * (faceIntegralLinear.cpph:8-101) and solutionUpdate, with one point source
 * (getPointSources / deltaDistribution, deltaDistribution.cpph:37-146) whose
 * time function is a Ricker wavelet amp (1 - 2 a) exp(-a), a = (pi f0 (t - t0))^2,
 * acting on p.  dt is fixed; t = step * dt.
 * Material (rho, c) is constant per cell (the gradient is per cell).
 * The misfit gradient is this library's own discrete adjoint (the reference
 * has none).
 * ------------------------------------------------------------------------- */
typedef struct exa_acoustic exa_acoustic;

int exa_acoustic_create(exa_acoustic** out, exa_ctx* ctx, int n, double h);
int exa_acoustic_destroy(exa_acoustic* ac);
long exa_acoustic_ncells(const exa_acoustic* ac);
/* Solution in the reference layout [ncells][B^3][6] (p,u,v,w,rho,c), cells x
 * fastest.  set: rho and c must be equal at all nodes of a cell
 * (EXA_ERR_UNSUPPORTED otherwise); they become the cell's material. */
int exa_acoustic_set_solution(exa_acoustic* ac, const double* luh);
int exa_acoustic_get_solution(exa_acoustic* ac, double* luh);
/* Per-cell material rho[ncells], c[ncells] (either may be NULL to keep it). */
int exa_acoustic_set_material(exa_acoustic* ac, const double* rho, const double* c);
int exa_acoustic_get_material(exa_acoustic* ac, double* rho, double* c);
/* dt > 0: fixed step; dt <= 0: CFL * PNPM / sum_d(max c / h) (stableTimeStepSize
 * with eigenvalues (c, -c, 0, 0)) of the current material. */
int exa_acoustic_set_dt(exa_acoustic* ac, double dt);
int exa_acoustic_set_step(exa_acoustic* ac, int step);
int exa_acoustic_time(exa_acoustic* ac, int* step, double* t, double* dt);
/* One point source at x[3] (amp = 0 removes it). */
int exa_acoustic_set_source(exa_acoustic* ac, const double* x, double amp, double f0, double t0);
/* Receivers at xr[nrec][3]: p interpolated with the cell's basis after every step. */
int exa_acoustic_set_receivers(exa_acoustic* ac, int nrec, const double* xr);
/* Advance nsteps steps from the current state; rec (nullable) receives
 * p[nsteps][nrec] at the receivers.  use_graph != 0 replays a captured
 * hipGraph of one step.  Synchronous. */
int exa_acoustic_run(exa_acoustic* ac, int nsteps, int use_graph, double* rec);
/* Misfit J = 1/2 sum_k sum_r (p_r(t_{k+1}) - obs[k][r])^2 over nsteps steps
 * from the current state, and dJ/dc per cell (grad_c[ncells]) by the
 * discrete adjoint with recomputation from checkpoints stored every
 * checkpoint_every steps.  Leaves the state after the last step (as
 * exa_acoustic_run).  Synchronous. */
int exa_acoustic_misfit_gradient(exa_acoustic* ac, int nsteps, int checkpoint_every, const double* obs,
                                 double* J, double* grad_c);
/* Milliseconds of the last exa_acoustic_run / exa_acoustic_misfit_gradient:
 * forward stepping, forward recomputation, adjoint sweep. */
int exa_acoustic_last_times(exa_acoustic* ac, double* ms_forward, double* ms_recompute, double* ms_adjoint);
/* Per-kernel timing of exa_acoustic_run (eager path): HIP events around the
 * predictor and the corrector of every step; milliseconds summed over the
 * last run's steps. */
int exa_acoustic_set_timing(exa_acoustic* ac, int enabled);
int exa_acoustic_phase_times(exa_acoustic* ac, double* ms_predictor, double* ms_corrector, int* nsteps);
/* Forward predictor form: 0 = x-line kernel (one thread per x-line of a cell,
 * x-derivatives from registers; the default), 1 = node kernel (one thread per
 * node). Both compute the same Cauchy-Kowalevski series; results agree to
 * rounding (summation order differs). No reference counterpart: a tuning
 * switch of this engine. */
int exa_acoustic_set_predictor(exa_acoustic* ac, int variant);

#ifdef __cplusplus
}
#endif
#endif

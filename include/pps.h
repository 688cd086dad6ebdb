/*
 * pps.h -- C ABI of the MI355X pressure-Poisson solver (libpps.so).
 *
 * Drop-in for the PARIS linear-solver entry points:
 *   NewSolver(A,p,maxError,beta,maxit,it,ierr,norm)         newsolver.f90:34-57
 *   NewSolverMG(A,p,maxError,beta,maxit,it,ierr,norm,tres2) MGsolver.f90:346-405
 *   init_MG(nvar) / finalize_MG                             MGsolver.f90:34-82
 * The module state those routines read implicitly (module_grid bounds and
 * Ng, module_BC bdry_cond, the MultiGrid/nrelax/sym_MG flags,
 * modules.f90:57-73) is passed explicitly in pps_grid / pps_options.
 *
 * Array conventions are the reference's (Fortran order, x fastest):
 *   A : 8 planes of nx*ny*nz doubles; slots 1..6 = x-,x+,y-,y+,z-,z+
 *       couplings, 7 = diagonal, 8 = right-hand side.  Never written.
 *   p : (nx+2ng)*(ny+2ng)*(nz+2ng) doubles with ng ghost layers; initial
 *       guess in, solution out.
 *   equation: A7 p = A1 p(i-1) + A2 p(i+1) + A3 p(j-1) + A4 p(j+1)
 *                    + A5 p(k-1) + A6 p(k+1) + A8
 * Stopping rule (MGsolver.f90:385-396): tres2 = (sum |r|^norm / N)^(1/norm
 * if norm==2), exit when tres2 < maxError; *it = maxit+1 if not converged.
 *
 * Ghost semantics: periodic directions wrap (the solver fills the first
 * ghost layer of p with the wrapped solution); in wall directions the ghost
 * values of p are fixed boundary data and are returned unchanged.
 *
 * All entry points are synchronous on return, thread-compatible, and not
 * re-entrant per handle.  Device-pointer variants take a hipStream_t as
 * void* and run on it, after the work already queued there (so inputs
 * written on that stream are ready); NULL means the default stream.
 * Host-array variants run on the handle's own stream.
 */
#ifndef PPS_H
#define PPS_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes (0 = success) ------------------------------------ */
enum pps_status {
    PPS_OK = 0,
    PPS_ERR_NAN_OR_TINY_DIAG = 1,  /* check_poisson_setup, newsolver.f90:295-310 */
    PPS_ERR_VOLUME_SOURCE = 2,     /* mean(A8) > 1e-10, newsolver.f90:314-321 */
    PPS_ERR_NOT_IN_KERNEL = 3,     /* row sums != 0, newsolver.f90:325-343 */
    PPS_ERR_BAD_SIZE = 4,          /* init_MG power-of-two rule, MGsolver.f90:44-46 */
    PPS_ERR_DIVERGED = 5,          /* catch_divergence, newsolver.f90:138-173 */
    PPS_ERR_NOT_CONVERGED = 6,     /* it = maxit+1 (MGsolver.f90:402): warning */
    PPS_ERR_NONSYMMETRIC = 7,      /* fast multigrid needs A2(i)=A1(i+1) etc. */
    PPS_ERR_BAD_ARG = 8,
    PPS_ERR_HIP = 9,               /* HIP runtime failure (out of memory, ...) */
    PPS_ERR_NO_DEVICE = 10,
    PPS_ERR_UNSUPPORTED = 11,      /* e.g. a method or option an entry point lacks */
    PPS_ERR_SCRATCH = 12           /* the grid needs more than the fixed scratch array
                                      of pps_solve_device_noalloc */
};

/* ---- solver methods --------------------------------------------------- */
enum pps_method {
    /* NewSolverMG restated on the GPU: RB-SOR(beta), nrelax sweeps, 8-slot
     * child-average coarsening, piecewise-constant prolongation
     * (MGsolver.f90:203-320).  FP64.  Same iteration history as the
     * reference up to summation order. */
    PPS_MG_REFERENCE = 0,
    /* FP64 multigrid-preconditioned CG: red-black Gauss-Seidel (nu_pre,
     * nu_post sweeps), residual fused with the 8-child average
     * restriction, face-coefficient coarsening (a coarse face = 1/16 of the
     * sum of its 4 fine faces), piecewise-constant prolongation (the
     * transpose of the restriction, which keeps the V-cycle symmetric for
     * CG), dense coarsest solve in one workgroup.  pps_options.transfer
     * selects trilinear prolongation with its transpose (27-point full
     * weighting) instead. */
    PPS_MG_FP64 = 1,
    /* Same V-cycle in FP32 (FP32 coefficients and accumulation) inside the
     * FP64 outer iteration: residual, update and reductions stay FP64. */
    PPS_MG_MIXED = 2,
    /* NewSolver_std restated: plain RB-SOR to tolerance (MultiGrid = F,
     * newsolver.f90:78-174). */
    PPS_SOR = 3
};

enum pps_smoother {
    PPS_SMOOTH_FUSED = 0,          /* one kernel per red-black sweep (default) */
    PPS_SMOOTH_COLOUR_PASSES = 1   /* one kernel per colour */
};

enum pps_outer {
    PPS_OUTER_CG = 0,      /* preconditioned conjugate gradients */
    PPS_OUTER_REFINE = 1   /* plain iterative refinement p += V(b - A p) */
};

/* Grid transfers of the fast V-cycle.  Both pairs keep R = P^T / 8, so the
 * V-cycle stays a symmetric operator under CG. */
enum pps_transfer {
    /* 8-child average restriction, piecewise-constant prolongation
     * (default; the prolongation rides in the first post-smoothing sweep) */
    PPS_TRANSFER_CONSTANT = 0,
    /* trilinear prolongation (weights 3/4, 1/4 per axis towards the nearer
     * coarse neighbour; constant extrapolation at walls, wrap on periodic
     * axes) with its transpose as the restriction: full weighting over the
     * 4 x 4 x 4 fine cells around each coarse cell */
    PPS_TRANSFER_TRILINEAR = 1
};

typedef struct pps_grid {
    int nx, ny, nz;        /* interior cells (Nx, Ny, Nz) */
    int ng;                /* ghost layers of p and the fields (Ng) */
    int bdry_cond[6];      /* x-, y-, z-, x+, y+, z+ (modules.f90:870-877) */
    double xlength, ylength, zlength;  /* box size, used by assembly only */
} pps_grid;

typedef struct pps_options {
    int method;            /* enum pps_method */
    int outer;             /* enum pps_outer (PPS_MG_FP64 / PPS_MG_MIXED) */
    int nrelax;            /* reference nrelax (default 3) */
    int sym_mg;            /* reference sym_MG (default 0) */
    int nu_pre, nu_post;   /* fast V-cycle smoothing sweeps (default 2, 2); equal
                              counts keep the cycle symmetric, as the CG outer
                              iteration assumes */
    int check_setup;       /* run the check_poisson_setup tests first (default 0) */
    int multigrid;         /* reference MultiGrid flag for pps_new_solver:
                              1 -> method, 0 -> PPS_SOR (default 1) */
    int smoother;          /* enum pps_smoother (fast methods, default fused) */
    int sweep_rows;        /* fused sweep: thread rows per block, 0 = automatic */
    int sweep_planes;      /* fused sweep: planes per block, 0 = automatic */
    int tail_cells;        /* fast V-cycle: levels below the finest with at most
                              this many cells run in one workgroup;
                              0 = automatic (4096), < 0 = never */
    int transfer;          /* enum pps_transfer (fast methods, default constant) */
} pps_options;

typedef struct pps_handle pps_handle;

/* Fill *opt with the defaults named above (method PPS_MG_FP64). */
int pps_default_options(pps_options *opt);

/* init_MG (MGsolver.f90:34-65): validate the grid, pick the device's
 * current HIP device, allocate the level hierarchy lazily. */
int pps_init(const pps_grid *grid, const pps_options *opt, pps_handle **out);
/* finalize_MG (MGsolver.f90:67-82). */
int pps_finalize(pps_handle *h);
/* Change the method / options of an existing handle. */
int pps_set_options(pps_handle *h, const pps_options *opt);
/* Cell widths of a stretched grid (the reference's dx(is:ie), dy(js:je),
 * dz(ks:ke) of module_grid, set by xform or read_x: paris.f90:3056-3102),
 * host arrays of nx, ny, nz positive values; all NULL clears them.  The fast
 * methods then weight each row of A by its cell volume, which makes
 * SetupPoisson's system (modules.f90:2737-2746) symmetric again:
 * dx(i) A2(i) = dx(i+1) A1(i+1).  They still stop on the caller's
 * (unweighted) residual.  Not used by the reference / SOR methods, which
 * take any A; the adjoint refuses weighted rows. */
int pps_set_cell_widths(pps_handle *h, const double *dx, const double *dy,
                        const double *dz);

/* NewSolver (newsolver.f90:34-57), host arrays.  Dispatches on
 * opt.multigrid like the reference; *ierr = 0 on success (the reference's
 * last MPI code), otherwise the pps_status value. */
int pps_new_solver(pps_handle *h, const double *A, double *p, double maxError,
                   double beta, int maxit, int *it, int *ierr, int norm);
/* NewSolverMG (MGsolver.f90:346-405), host arrays, also returns tres2. */
int pps_new_solver_mg(pps_handle *h, const double *A, double *p,
                      double maxError, double beta, int maxit, int *it,
                      int *ierr, int norm, double *tres2);

/* Device-pointer solve: dA (8 planes) and dp (padded) already in HBM. */
int pps_solve_device(pps_handle *h, const double *dA, double *dp,
                     double maxError, double beta, int maxit, int norm,
                     int *it, double *tres2, void *stream);
/* NewSolver followed by project_velocity, staggered branch (the main
 * projection, paris.f90:499 then paris.f90:898-908): on return p holds the
 * solve's result and u_out, v_out, w_out (padded like p, not aliasing u, v,
 * w) the corrected velocities, ghosts copied from u, v, w.  The fast methods
 * with norm 2 apply the correction inside their last residual pass.  Uses
 * the grid's uniform spacings: PPS_ERR_UNSUPPORTED once cell widths are set
 * (stretched grids call pps_project_velocity_coords_device after the solve). */
int pps_solve_project_device(pps_handle *h, const double *dA, double *dp,
                             double maxError, double beta, int maxit, int norm,
                             int *it, double *tres2, const double *u,
                             const double *v, const double *w,
                             const double *umask, const double *vmask,
                             const double *wmask, const double *rho, double dt,
                             double *u_out, double *v_out, double *w_out,
                             void *stream);
/* The projection step from the fields, without forming A: SetupPoisson +
 * Poisson_BCs (modules.f90:2722-2892; codes 0-4) are evaluated inside the
 * fast solve's setup pass, the same rows bit for bit as
 * pps_setup_poisson_device (volume_source is S_vol, uniform spacings of
 * the grid), then the solve runs as pps_solve_device (norm 2), and when
 * u_out, v_out, w_out are given (all three, not aliasing u, v, w)
 * project_velocity (paris.f90:898-908) is applied in its last residual
 * pass, as pps_solve_project_device does.  Replaces the SetupPoisson ->
 * NewSolver -> project_velocity sequence of paris.f90:499, 898-908 on one
 * device.  Fast methods only, opt.check_setup 0, no cell widths (else
 * PPS_ERR_UNSUPPORTED).  Same p, it and tres2 as assembling A and calling
 * pps_solve_device; no initial-guess reuse beyond dp itself. */
int pps_solve_fields_device(pps_handle *h, const double *u, const double *v,
                            const double *w, const double *umask,
                            const double *vmask, const double *wmask,
                            const double *rho, double dt, double volume_source,
                            double *dp, double maxError, int maxit, int *it,
                            double *tres2, double *u_out, double *v_out,
                            double *w_out, void *stream);
/* pps_solve_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * solve (tight serving loops, harnesses that forbid hipMalloc).  The
 * solver's buffers (outer vectors, operator copy, reduction partials, the
 * V-cycle hierarchy) are taken from one fixed 128 MiB array that each GPU
 * holds from the library's first use until the process ends; pps_init
 * looks its address up.  Limits: methods PPS_MG_FP64 / PPS_MG_MIXED, norm 2,
 * opt.check_setup 0, maxit >= 1 (else PPS_ERR_UNSUPPORTED / PPS_ERR_BAD_ARG);
 * grids up to about 1.1 million cells (104^3 mixed, 108^3 FP64), larger
 * ones return PPS_ERR_SCRATCH before any work is queued.  Every call on
 * one GPU shares the array and nothing carries over between calls; calls
 * that overlap in time on one GPU (several threads or streams) race.
 * Results equal pps_solve_device's bit for bit. */
int pps_solve_device_noalloc(pps_handle *h, const double *dA, double *dp,
                             double maxError, double beta, int maxit, int norm,
                             int *it, double *tres2, void *stream);

/* Residual history of the last solve: tres2 after each outer iteration.
 * Returns the number of entries written (<= cap). */
/* FreeSolver (freesurface.f90:700-808; called at paris.f90:496 and 606):
 * RB-SOR (RedBlackRelax_fs, freesurface.f90:857-895) on the active cells of
 * the free-surface mask only: pcmask == 0 for solver_flag 1, pcmask in
 * {1, 2} for solver_flag 2.  First p = P_gas where pcmask != 0 (flag 1) or
 * p = 0 where pcmask == 3 (flag 2).  Stops when the max-norm residual over
 * the active cells is below maxError (the reference's use_L_inf); tres2 =
 * sqrt(sum r^2) / (active cells); it = maxit+1 when not converged.  The
 * reference takes pcmask, P_gas and solver_flag from module_freesurface
 * (modules.f90:411-420); here they are arguments: pcmask (int32) and
 * P_gas are padded like p (vofmodules.f90:446-447); P_gas may be NULL for
 * solver_flag 2.  Returns PPS_ERR_BAD_ARG for a solver_flag other than 1
 * or 2 (the reference aborts, freesurface.f90:722) and PPS_ERR_DIVERGED
 * where the reference aborts in catch_divergence_fs. */
int pps_free_solver(pps_handle *h, const double *A, double *p, const int *pcmask,
                    const double *P_gas, int solver_flag, double maxError, double beta,
                    int maxit, int *it, int *ierr, double *tres2);
/* The same on device arrays, on the given stream (NULL: the default). */
int pps_free_solver_device(pps_handle *h, const double *dA, double *dp, const int *dpcmask,
                           const double *dP_gas, int solver_flag, double maxError,
                           double beta, int maxit, int *it, double *tres2, void *stream);

/* poi_solve (modules.f90:2411-2593; called at paris.f90:490): the HYPRE
 * entry on the same 8-slot system (matrix A7 on the diagonal, -A1..-A6 off
 * it, right-hand side A8, modules.f90:2447-2469).  HYPRE is not linked
 * into this library: every solver type (1 SMG, 2 PFMG, 3 GMRES, 4 BiCGSTAB) runs
 * the handle's fast MG-PCG (PPS_MG_MIXED if the handle is mixed, else
 * PPS_MG_FP64) to ||A8 - M p||_2 / ||A8||_2 < maxError, the relative
 * tolerance HYPRE's SetTol gives.  p is the initial guess and the result;
 * *num_iterations = iterations performed (at most maxit).  Returns
 * PPS_ERR_BAD_ARG for another solver type, PPS_ERR_NOT_CONVERGED after
 * maxit iterations. */
int pps_poi_solve(pps_handle *h, const double *A, double *p, double maxError, int maxit,
                  int *num_iterations, int solver_type);
/* The same on device arrays, on the given stream (NULL: the default). */
int pps_poi_solve_device(pps_handle *h, const double *dA, double *dp, double maxError,
                         int maxit, int *num_iterations, int solver_type, void *stream);

int pps_last_history(pps_handle *h, double *hist, int cap);

/* Adjoint of the last-solved system (no reference counterpart).  With
 * M = diag(A7) - offdiag(A1..A6), solves M^T lam = g (g projected mean-free
 * when M is singular) to relative residual rtol, on the hierarchy of the
 * last solve when dA is that solve's A (same address and method; the
 * caller keeps its contents unchanged in between), else on one set up
 * from dA.
 * M^T = M when M is symmetric; with cell widths set (pps_set_cell_widths)
 * the volume-weighted rows W M are symmetric and lam = W (W M)^-1 g.
 * dg, dlam: padded like p. */
int pps_adjoint_device(pps_handle *h, const double *dA, const double *dg,
                       double *dlam, double rtol, int maxit, int *it,
                       void *stream);
/* Fused face kernel: grad_f = -w_f (lam_L - lam_R)(p_L - p_R) for the
 * +x, +y, +z face of every cell (3 planes of nx*ny*nz), w_f = the
 * geometric weight dt/(dx*dxh) of the face; the last face of a wall axis
 * (which couples to no cell) gets 0.  wx/wy/wz: per-plane weights (length
 * nx/ny/nz); dp, dlam padded like p (periodic neighbours are wrapped here,
 * ghosts are not read).  This is dJ/dc_f for the map of
 * pps_faces_to_A_device (A = w c), with c_f = 2/(rho_L+rho_R) = "1/rho at
 * the face" (modules.f90:2737-2742); a caller whose couplings also carry
 * SetupPoisson's masks multiplies the result by them. */
int pps_face_gradient_device(const pps_grid *grid, const double *dp,
                             const double *dlam, const double *wx,
                             const double *wy, const double *wz,
                             double *dgrad, void *stream);
/* The linear map from face values to the 8-slot system that the face
 * gradient differentiates: with c = 3 planes (+x, +y, +z face of every
 * cell) and per-plane weights w, A2(i) = A1(i+1) = w_x(i) c_x(i) (wrapping
 * in periodic directions, 0 through walls), likewise y and z,
 * A7 = sum(A1..A6), A8 = b (interior, nx*ny*nz).  It is SetupPoisson's
 * coefficient rule (modules.f90:2737-2743) with c_f = 2/(rho_L+rho_R) and
 * w = dt/(dx*dxh) on a uniform grid. */
int pps_faces_to_A_device(const pps_grid *grid, const double *cface,
                          const double *wx, const double *wy,
                          const double *wz, const double *b, double *A,
                          void *stream);

/* pps_faces_to_A_device followed by pps_solve_device (norm 2), without
 * forming A: the rows (the same products and sums, bit for bit) are taken
 * from the face values inside the fast solve's setup pass.  Same p, it and
 * tres2 as that pair of calls; b interior (nx*ny*nz).  Fast methods only,
 * opt.check_setup 0, no cell widths (else PPS_ERR_UNSUPPORTED).  The hierarchy stays known
 * by cface's address for pps_adjoint_faces_device.  The system is
 * SetupPoisson's (modules.f90:2737-2743) with face values c_f; the solve
 * replaces NewSolver (newsolver.f90:34-57) on it. */
int pps_solve_faces_device(pps_handle *h, const double *cface,
                           const double *wx, const double *wy,
                           const double *wz, const double *b, double *dp,
                           double maxError, int maxit, int *it,
                           double *tres2, void *stream);
/* pps_adjoint_device for the system of face values cface (weights w): the
 * hierarchy of the last pps_solve_faces_device on the same cface, wx, wy,
 * wz addresses and method is reused (the caller keeps their contents
 * unchanged in between), else set up from the face values.  Same lam and it
 * as pps_adjoint_device on the A that pps_faces_to_A_device forms.  No cell
 * widths (else PPS_ERR_UNSUPPORTED). */
int pps_adjoint_faces_device(pps_handle *h, const double *cface,
                             const double *wx, const double *wy,
                             const double *wz, const double *dg,
                             double *dlam, double rtol, int maxit, int *it,
                             void *stream);

/* ---- producer / consumer of the system (modules.f90, paris.f90) ------ */
/* linfunc's MeanFlag values (vofmodules.f90:109-110) */
enum pps_mean { PPS_ARITH_MEAN = 101, PPS_HARM_MEAN = 102 };
/* linfunc (vofmodules.f90:171-218) then do_all_ghost (modules.f90:2062-2072)
 * on one rank: field = cvof * (a2 - a1) + a1 (PPS_ARITH_MEAN) or
 * 1 / (cvof * (1/a2 - 1/a1) + 1/a1) (PPS_HARM_MEAN, each inverse of a + 1e-50),
 * with cvof 27-point filtered off the outermost ghost layer when nfilter is 1;
 * then the ghost layers of periodic axes are wrapped, axis by axis over the
 * full extent of the others.  cvof, field: padded like p; field must not be
 * cvof when nfilter is 1.  nfilter other than 0, 1: PPS_ERR_BAD_ARG. */
int pps_linfunc_device(const pps_grid *grid, const double *cvof, double a1, double a2,
                       int mean_flag, int nfilter, double *field, void *stream);
/* SetPressureBC on all-ones masks (modules.f90:916-946), padded arrays. */
int pps_pressure_masks_device(const pps_grid *grid, double *umask,
                              double *vmask, double *wmask, void *stream);
/* SetupPoisson + Poisson_BCs (codes 0..6) + the solid-cell rule of
 * check_and_debug_Poisson (modules.f90:2722-2953), padded fields in,
 * 8-slot A out; BoundaryPressure = 0. */
int pps_setup_poisson_device(const pps_grid *grid, const double *u,
                             const double *v, const double *w,
                             const double *umask, const double *vmask,
                             const double *wmask, const double *rho, double dt,
                             double volume_source, double *A, void *stream);
/* The same with the reference's BoundaryPressure(1:6) (module_BC,
 * modules.f90:878-883; -x, +x, -y, +y, -z, +z), a host array of 6 doubles,
 * which the pressure codes read: code 5 makes the face row
 * P_0 - P_inner / 3 = 2/3 P_b, code 6 (x+) sets p and du/dn at the face
 * (modules.f90:2768-2891).  Rows of codes 5 / 6 are not symmetric: only
 * PPS_MG_REFERENCE and PPS_SOR solve them (the reference switches
 * check_setup off for them, paris.f90:2992). */
int pps_setup_poisson_bp_device(const pps_grid *grid, const double *u,
                                const double *v, const double *w,
                                const double *umask, const double *vmask,
                                const double *wmask, const double *rho,
                                double dt, double volume_source,
                                const double *boundary_pressure, double *A,
                                void *stream);
/* The same on a stretched grid (xform, read_x: paris.f90:3056-3102): the
 * caller's dx, dxh, dy, dyh, dz, dzh (host arrays with ghosts, n + 2 ng
 * values each, the reference's dx(imin:imax) ...) replace the uniform
 * spacings of the box lengths. */
int pps_setup_poisson_coords_device(const pps_grid *grid, const double *u,
                                    const double *v, const double *w,
                                    const double *umask, const double *vmask,
                                    const double *wmask, const double *rho,
                                    double dt, double volume_source,
                                    const double *dx, const double *dxh,
                                    const double *dy, const double *dyh,
                                    const double *dz, const double *dzh,
                                    double *A, void *stream);
/* check_poisson_setup (newsolver.f90:275-344) with the face masks, as the
 * reference reads them: solid cells (A7 := 0 in the row-sum test) are those
 * whose six face masks are closed (< 1/2).  Returns PPS_OK or the first
 * failed test: PPS_ERR_NAN_OR_TINY_DIAG, PPS_ERR_VOLUME_SOURCE,
 * PPS_ERR_NOT_IN_KERNEL. */
int pps_check_poisson_setup_device(const pps_grid *grid, const double *A,
                                   const double *umask, const double *vmask,
                                   const double *wmask, void *stream);
/* The same without masks (what the solvers' check_setup option runs on A
 * alone): solid cells are those with six zero couplings, the same cells
 * when A comes from SetupPoisson. */
int pps_check_setup_device(const pps_grid *grid, const double *A, void *stream);
/* Subtract the mean of A8 (compatibility projection of the right-hand side). */
int pps_remove_rhs_mean_device(const pps_grid *grid, double *A, void *stream);
/* calcResiduals (modules.f90:3313-3384): out = {L1, L2, Linf}. */
int pps_calc_residuals_device(const pps_grid *grid, const double *A,
                              const double *p, double out[3], void *stream);
/* project_velocity, staggered branch (paris.f90:898-908). */
int pps_project_velocity_device(const pps_grid *grid, double *u, double *v,
                                double *w, const double *umask,
                                const double *vmask, const double *wmask,
                                const double *rho, const double *p, double dt,
                                void *stream);
/* The same on a stretched grid: dx, dxh, dy, dyh, dz, dzh as for
 * pps_setup_poisson_coords_device (the face corrections use dxh, dyh, dzh). */
int pps_project_velocity_coords_device(const pps_grid *grid, double *u,
                                       double *v, double *w,
                                       const double *umask, const double *vmask,
                                       const double *wmask, const double *rho,
                                       const double *p, double dt,
                                       const double *dx, const double *dxh,
                                       const double *dy, const double *dyh,
                                       const double *dz, const double *dzh,
                                       void *stream);

/* ---- synthetic inputs (benchmark / test fields, no reference code) ---- */
/* Density from spherical droplets (tanh interface of width `width` cells,
 * volume fraction capped at 1, rho = rho1 + (rho2-rho1) c) and iid
 * U(-1,1) face velocities from a counter-based splitmix64 stream.  Wall
 * faces get zero normal velocity; ghosts wrap (periodic) or copy the
 * nearest interior cell (walls).  drops: ndrop x {cx, cy, cz, R}. */
int pps_synth_fields_device(const pps_grid *grid, const double *drops,
                            int ndrop, double rho1, double rho2, double width,
                            unsigned long long seed, double *rho, double *u,
                            double *v, double *w, void *stream);
/* Droplet list: ndrop centres uniform in the box and radii uniform in
 * [rmin, rmax] from splitmix64(seed); host memory, 4 doubles per drop. */
int pps_synth_drops(const pps_grid *grid, int ndrop, double rmin, double rmax,
                    unsigned long long seed, double *drops);

/* Uniform-grid coordinates as paris.f90:2959-3102 builds them (xform=0):
 * arrays of N+2ng entries, entry q = Fortran index q+1. */
int pps_coords(int N, int ng, double length, double *x, double *xh,
               double *dx, double *dxh);

/* Human-readable name of a status code. */
const char *pps_status_string(int status);

/* Timing of the last solve on the handle's stream: milliseconds of the
 * whole solve and of the setup (pack + coarsening) part. */
int pps_last_timing(pps_handle *h, double *solve_ms, double *setup_ms);

/* Measurement hooks (no reference counterpart; the reference's analogue is
 * its wall-clock timer bucket "Poisson for pressure", modules.f90:252-296).
 * When enabled (on = 1), every timed finest-level smoother launch of the
 * fast methods (the plain sweep: not the first of a cycle, not the one that
 * adds the coarse correction or forms the dot products) carries a HIP event
 * pair, the kernel's own start and end (on = N > 1: every N-th such launch);
 * read returns the summed kernel milliseconds, the launch count and the
 * summed algorithmic bytes since the last read, and resets them. */
int pps_profile_enable(pps_handle *h, int on);
int pps_profile_read(pps_handle *h, double *kernel_ms, long long *launches,
                     double *bytes);

#ifdef __cplusplus
}
#endif
#endif

/*
 * pps_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the PARIS pressure-Poisson path in plain C (FP64).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline.  The product
 * library (libpps) never links or calls it.
 *
 * Conventions (identical to the reference, newsolver.f90:30-46):
 *   A  : 8 contiguous planes of nx*ny*nz doubles, x fastest (Fortran order);
 *        slot s (1..8) of cell (i,j,k) (0-based interior indices) is
 *        A[(s-1)*n + i + nx*(j + ny*k)].  Slots 1..6 = x-,x+,y-,y+,z-,z+
 *        couplings, 7 = diagonal, 8 = right-hand side.
 *   p  : (nx+2ng)*(ny+2ng)*(nz+2ng) doubles with ng ghost layers, x fastest.
 *   equation: A7*p = A1*p(i-1)+A2*p(i+1)+A3*p(j-1)+A4*p(j+1)+A5*p(k-1)+A6*p(k+1)+A8
 *   bc[6]: bdry_cond(1..6) = x-, y-, z-, x+, y+, z+ ; 0 wall, 1 periodic,
 *          2 shear, 3 velocity inflow, 4 velocity outflow (modules.f90:870-877).
 */
#ifndef PPS_ORACLE_H
#define PPS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int nx, ny, nz;       /* interior cells (one rank owns the whole box) */
    int ng;               /* ghost layers on p (reference Ng, default 2) */
    int bc[6];            /* bdry_cond */
    double xlength, ylength, zlength;
} orc_grid;

enum {
    ORC_OK = 0,
    ORC_ERR_NAN_OR_TINY_DIAG = 1,   /* check_poisson_setup: NaN or A7 < 1e-55 */
    ORC_ERR_VOLUME_SOURCE = 2,      /* check_poisson_setup: mean(A8) > 1e-10 */
    ORC_ERR_NOT_IN_KERNEL = 3,      /* check_poisson_setup: row sum != 0 */
    ORC_ERR_BAD_SIZE = 4,           /* init_MG: not a power of two */
    ORC_ERR_DIVERGED = 5,           /* catch_divergence: residual > 1e16 or NaN */
    ORC_ERR_SOLID = 6,              /* check_and_debug_Poisson inconsistencies */
    ORC_ERR_UNSUPPORTED = 7,
    ORC_ERR_NOMEM = 8
};

/* Uniform-grid coordinates exactly as paris.f90:2959-3102 builds them
 * (xform = 0).  Arrays have N+2ng entries; entry q holds Fortran index q+1. */
void orc_coords(int N, int ng, double length, double *x, double *xh,
                double *dx, double *dxh);

/* SetPressureBC (modules.f90:916-946) applied to all-ones masks. */
void orc_pressure_masks(const orc_grid *g, double *umask, double *vmask,
                        double *wmask);

/* SetupPoisson + Poisson_BCs (codes 0..6) + check_and_debug_Poisson
 * (modules.f90:2722-2953) with BoundaryPressure = 0.  Fields carry ng
 * ghosts.  Returns ORC_* code. */
int orc_setup_poisson(const orc_grid *g, const double *u, const double *v,
                      const double *w, const double *umask,
                      const double *vmask, const double *wmask,
                      const double *rho, double dt, double volume_source,
                      double *A);
/* The same with the module's BoundaryPressure(1:6) (-x, +x, -y, +y, -z, +z,
 * modules.f90:878-883), which the pressure codes 5 / 6 read. */
int orc_setup_poisson_bp(const orc_grid *g, const double *u, const double *v,
                         const double *w, const double *umask,
                         const double *vmask, const double *wmask,
                         const double *rho, double dt, double volume_source,
                         const double bp[6], double *A);

/* check_poisson_setup (newsolver.f90:275-344); rowsum (optional, n cells)
 * receives sum(A1..6) - A7 with A7 := 0 in solid cells. */
int orc_check_poisson_setup(const orc_grid *g, const double *A,
                            const double *umask, const double *vmask,
                            const double *wmask, double *rowsum);

/* Residual r = A8 - A7 p + sum A_nb p_nb (MGsolver.f90:97-100) into r (n). */
void orc_residual_field(const orc_grid *g, const double *A, const double *p,
                        double *r);
/* compute_residual (MGsolver.f90:106-126): sum |r|^norm, reference order. */
double orc_compute_residual(const orc_grid *g, const double *A,
                            const double *p, int norm);
/* calcResiduals (modules.f90:3313-3384): out = {L1, L2, Linf}. */
void orc_calc_residuals(const orc_grid *g, const double *A, const double *p,
                        double out[3]);

/* Periodic one-layer ghost exchange of one level (ghost_MG_x/y/z with one
 * rank; wall directions untouched). */
void orc_apply_bc(const orc_grid *g, double *p);
/* RedBlackRelax (newsolver.f90:209-247) on the fine level, one sweep. */
void orc_redblack_relax(const orc_grid *g, const double *A, double *p,
                        double beta);

/* MG transfer operators on bare level arrays (MGsolver.f90:128-251).
 * nf = fine interior dims; Af/Ac: 8-slot level arrays without ghosts;
 * ef/ec: level vectors with ng ghosts. */
void orc_fill_coefficients(const int nf[3], const double *Af, double *Ac);
void orc_coarse_from_fine(const int nf[3], const double *Af, double *Ac);
void orc_coarse_fine_interp(const int nf[3], int ng, int sym_mg, double *ef,
                            const double *ec);

/* NewSolverMG (MGsolver.f90:346-405) including the module state of
 * init_MG (34-65).  hist (optional, maxit entries) receives tres2 after
 * each iteration.  *it follows the reference: maxit+1 when not converged.
 * Returns ORC_OK or ORC_ERR_BAD_SIZE. */
int orc_newsolver_mg(const orc_grid *g, const double *A, double *p,
                     double maxError, double beta, int maxit, int norm,
                     int nrelax, int sym_mg, int *it, double *tres2,
                     double *hist);
/* NewSolver_std (newsolver.f90:78-174), plain RB-SOR to tolerance. */
int orc_newsolver_std(const orc_grid *g, const double *A, double *p,
                      double maxError, double beta, int maxit, int norm,
                      int *it, double *tres2, double *hist);

/* FreeSolver (freesurface.f90:700-855) with RedBlackRelax_fs (857-895):
 * RB-SOR on the active cells of the free-surface mask, pcmask == 0 for
 * solver_flag 1, pcmask in {1, 2} for solver_flag 2.  pcmask (int) and
 * P_gas carry ng ghosts like p.  Stops when the max-norm residual over the
 * active cells is below maxError (use_L_inf); tres2 = sqrt(sum r^2) /
 * (active cells).  hist (optional, maxit) receives tres2 per iteration.
 * Returns ORC_ERR_UNSUPPORTED for a solver_flag other than 1 or 2 and
 * ORC_ERR_DIVERGED for NaN or sum r^2 > 1e18 (catch_divergence_fs). */
int orc_free_solver(const orc_grid *g, const double *A, double *p,
                    const int *pcmask, const double *pgas, int solver_flag,
                    double maxError, double beta, int maxit, int *it,
                    double *tres2, double *hist);

/* Setup_testMG (modules.f90:3160-3186) and get_MGtest_err
 * (MGsolver.f90:407-426); both use the reference's single-precision pi. */
void orc_setup_testmg(const orc_grid *g, const double *rho, double *A);
double orc_get_mgtest_err(const orc_grid *g, const double *p);

/* project_velocity, plain staggered branch (paris.f90:898-908). */
/* linfunc (vofmodules.f90:171-218) + do_all_ghost (modules.f90:2062-2072) on
 * padded arrays; mean_flag 101 ArithMean / 102 HarmMean, nfilter 0 or 1.
 * Returns 0, or -1 for other flags. */
int orc_linfunc(const orc_grid *g, const double *cvof, double a1, double a2, int mean_flag,
                int nfilter, double *field);

void orc_project_velocity(const orc_grid *g, double *u, double *v, double *w,
                          const double *umask, const double *vmask,
                          const double *wmask, const double *rho,
                          const double *p, double dt);

/* Independent exact solver (not a reference algorithm): Jacobi-preconditioned
 * CG on the same 8-slot system, ghosts handled as the reference does
 * (periodic wrap, wall ghosts fixed).  Used to pin converged solutions.
 * Runs until ||r||2/||b||2 <= rtol or maxit.  Returns iterations done. */
int orc_pcg_solve(const orc_grid *g, const double *A, double *p, double rtol,
                  int maxit, double *relres);

/* Threads used by the NewSolverMG restatement (default 1 = serial, the
 * order the tests pin); orc_max_threads() is OpenMP's maximum. */
void orc_set_threads(int n);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif

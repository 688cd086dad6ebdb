/*
 * pps_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * Plain-C FP64 restatement of the reference PARIS pressure-Poisson path
 * (MGsolver.f90, newsolver.f90, modules.f90, paris.f90).  Each function
 * cites the reference lines it follows.  Only tests/, smoke() and the CPU
 * baseline leg of bench.py use this file, as the checker; the product
 * library never links it.
 *
 * Built with -ffp-contract=off so every product of a sum is rounded
 * separately, as in the reference's plain Fortran expressions; sums run in
 * the reference's left-to-right order and its k,j,i loop order.
 *
 * Single-rank semantics: MPI halo exchanges become a periodic wrap of one
 * ghost layer in periodic directions, and no-ops at walls (MPI_PROC_NULL);
 * MPI_ALLREDUCE over one rank is the identity.
 */
#include "pps_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* level geometry                                                      */
/* ------------------------------------------------------------------ */
typedef struct {
    int n[3];     /* interior cells */
    int ng;       /* ghost layers */
    int X, Y, Z;  /* padded extents */
    int per[3];   /* periodic flags */
} lvl_t;

static lvl_t lvl_make(const int n[3], int ng, const int per[3]) {
    lvl_t L;
    for (int d = 0; d < 3; ++d) { L.n[d] = n[d]; L.per[d] = per[d]; }
    L.ng = ng;
    L.X = n[0] + 2 * ng; L.Y = n[1] + 2 * ng; L.Z = n[2] + 2 * ng;
    return L;
}
static inline long ic(const lvl_t *L, int i, int j, int k) {
    return (long)i + (long)L->n[0] * ((long)j + (long)L->n[1] * k);
}
static inline long ip(const lvl_t *L, int i, int j, int k) {
    return (long)(i + L->ng) + (long)L->X * ((long)(j + L->ng) + (long)L->Y * (k + L->ng));
}
/* Threads of the MG restatement's loops (1: serial, the order every test
 * pins).  Colour passes, residuals and transfers are independent per cell,
 * so more threads change only the summation order of the residual norms.
 * Used to time the reference algorithm on all host cores. */
static int orc_nthreads = 1;

void orc_set_threads(int n) { orc_nthreads = n < 1 ? 1 : n; }
int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline long ncell(const lvl_t *L) { return (long)L->n[0] * L->n[1] * L->n[2]; }
static inline long npad(const lvl_t *L) { return (long)L->X * L->Y * L->Z; }

static lvl_t grid_lvl(const orc_grid *g) {
    int n[3] = {g->nx, g->ny, g->nz};
    int per[3] = {g->bc[0] == 1, g->bc[1] == 1, g->bc[2] == 1};
    return lvl_make(n, g->ng, per);
}

/* ------------------------------------------------------------------ */
/* coordinates: paris.f90:3045-3102 with xform = 0 and read_x = F       */
/* ------------------------------------------------------------------ */
void orc_coords(int N, int ng, double length, double *x, double *xh,
                double *dx, double *dxh) {
    int nt = N + 2 * ng;
    /* Fortran index q = 1..nt maps to array entry q-1 */
    for (int q = 1; q <= nt; ++q) {
        double s = (double)(q - ng) / (double)N;
        double form = 0.0 * s * (0.5 - s) * (1.0 - s);  /* xform = 0 */
        xh[q - 1] = length * (form + s);
    }
    for (int q = 2; q <= nt; ++q) x[q - 1] = 0.5 * (xh[q - 1] + xh[q - 2]);
    x[0] = 2.0 * xh[0] - x[1];
    for (int q = 1; q <= nt - 1; ++q) dxh[q - 1] = x[q] - x[q - 1];
    dxh[nt - 1] = dxh[nt - 2];
    for (int q = 2; q <= nt; ++q) dx[q - 1] = xh[q - 1] - xh[q - 2];
    dx[0] = dx[1];
}

typedef struct {
    double *x, *xh, *dx, *dxh;
} axis_t;

static int axis_alloc(axis_t *a, int N, int ng, double length) {
    int nt = N + 2 * ng;
    a->x = malloc(sizeof(double) * nt * 4);
    if (!a->x) return -1;
    a->xh = a->x + nt; a->dx = a->xh + nt; a->dxh = a->dx + nt;
    orc_coords(N, ng, length, a->x, a->xh, a->dx, a->dxh);
    return 0;
}
static void axis_free(axis_t *a) { free(a->x); }

/* ------------------------------------------------------------------ */
/* SetPressureBC: modules.f90:916-946                                  */
/* ------------------------------------------------------------------ */
void orc_pressure_masks(const orc_grid *g, double *umask, double *vmask,
                        double *wmask) {
    lvl_t L = grid_lvl(g);
    long np = npad(&L);
    for (long q = 0; q < np; ++q) umask[q] = vmask[q] = wmask[q] = 1.0;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    /* wall (0) or shear (2): zero the wall-normal face mask, over the
     * interior +-1 cell band of the tangential directions */
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j) {
            if (g->bc[0] == 0 || g->bc[0] == 2) umask[ip(&L, -1, j, k)] = 0.0;
            if (g->bc[3] == 0 || g->bc[3] == 2) umask[ip(&L, nx - 1, j, k)] = 0.0;
        }
    for (int k = -1; k <= nz; ++k)
        for (int i = -1; i <= nx; ++i) {
            if (g->bc[1] == 0 || g->bc[1] == 2) vmask[ip(&L, i, -1, k)] = 0.0;
            if (g->bc[4] == 0 || g->bc[4] == 2) vmask[ip(&L, i, ny - 1, k)] = 0.0;
        }
    for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i) {
            if (g->bc[2] == 0 || g->bc[2] == 2) wmask[ip(&L, i, j, -1)] = 0.0;
            if (g->bc[5] == 0 || g->bc[5] == 2) wmask[ip(&L, i, j, nz - 1)] = 0.0;
        }
}

/* ------------------------------------------------------------------ */
/* SetupPoisson (2722-2753) + Poisson_BCs codes 3/4 (2755-2892)         */
/* + check_and_debug_Poisson (2894-2953)                                */
/* ------------------------------------------------------------------ */
int orc_setup_poisson(const orc_grid *g, const double *u, const double *v,
                      const double *w, const double *umask,
                      const double *vmask, const double *wmask,
                      const double *rho, double dt, double volume_source,
                      double *A) {
    const double bp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    return orc_setup_poisson_bp(g, u, v, w, umask, vmask, wmask, rho, dt, volume_source, bp, A);
}

int orc_setup_poisson_bp(const orc_grid *g, const double *u, const double *v,
                         const double *w, const double *umask,
                         const double *vmask, const double *wmask,
                         const double *rho, double dt, double volume_source,
                         const double bp[6], double *A) {
    for (int f = 0; f < 6; ++f)
        if (g->bc[f] < 0 || g->bc[f] > 6) return ORC_ERR_UNSUPPORTED;
    lvl_t L = grid_lvl(g);
    long n = ncell(&L);
    axis_t ax, ay, az;
    if (axis_alloc(&ax, g->nx, g->ng, g->xlength)) return ORC_ERR_NOMEM;
    if (axis_alloc(&ay, g->ny, g->ng, g->ylength)) { axis_free(&ax); return ORC_ERR_NOMEM; }
    if (axis_alloc(&az, g->nz, g->ng, g->zlength)) { axis_free(&ax); axis_free(&ay); return ORC_ERR_NOMEM; }
    const int G = g->ng;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long c = ic(&L, i, j, k);
                long q = ip(&L, i, j, k);
                /* coordinate arrays are indexed by padded index i+G */
                int I = i + G, J = j + G, K = k + G;
                double r0 = rho[q];
                double a1 = 2.0 * dt * umask[ip(&L, i - 1, j, k)] /
                            (ax.dx[I] * ax.dxh[I - 1] * (rho[ip(&L, i - 1, j, k)] + r0));
                double a2 = 2.0 * dt * umask[q] /
                            (ax.dx[I] * ax.dxh[I] * (rho[ip(&L, i + 1, j, k)] + r0));
                double a3 = 2.0 * dt * vmask[ip(&L, i, j - 1, k)] /
                            (ay.dx[J] * ay.dxh[J - 1] * (rho[ip(&L, i, j - 1, k)] + r0));
                double a4 = 2.0 * dt * vmask[q] /
                            (ay.dx[J] * ay.dxh[J] * (rho[ip(&L, i, j + 1, k)] + r0));
                double a5 = 2.0 * dt * wmask[ip(&L, i, j, k - 1)] /
                            (az.dx[K] * az.dxh[K - 1] * (rho[ip(&L, i, j, k - 1)] + r0));
                double a6 = 2.0 * dt * wmask[q] /
                            (az.dx[K] * az.dxh[K] * (rho[ip(&L, i, j, k + 1)] + r0));
                A[0 * n + c] = a1; A[1 * n + c] = a2; A[2 * n + c] = a3;
                A[3 * n + c] = a4; A[4 * n + c] = a5; A[5 * n + c] = a6;
                A[6 * n + c] = ((((a1 + a2) + a3) + a4) + a5) + a6;
                A[7 * n + c] = -(((volume_source + (u[q] - u[ip(&L, i - 1, j, k)]) / ax.dx[I]) +
                                  (v[q] - v[ip(&L, i, j - 1, k)]) / ay.dx[J]) +
                                 (w[q] - w[ip(&L, i, j, k - 1)]) / az.dx[K]);
            }
    axis_free(&ax); axis_free(&ay); axis_free(&az);

    /* Poisson_BCs on a single rank (every face is a domain face), block by
     * block in the reference's statement order; later blocks overwrite the
     * corner entries earlier ones set.  Codes 3 / 4: drop the outward
     * coupling (flux condition).  Code 5: the face row becomes
     * P_0 - P_inner / 3 = 2/3 P_b (linear extrapolation to the face value
     * BoundaryPressure), corners adding the conditions of both faces
     * (P_bc flags).  Code 6 (x+ only): p_b and du/dn at the face.  The z+
     * block of code 5 writes its y-coupling corner entries into the ks plane
     * (modules.f90:2884-2887), as the reference does. */
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int ie = nx - 1, je = ny - 1, ke = nz - 1;
    double P_bc[4] = {0.0, 0.0, 0.0, 0.0};
    const double c23 = 2.0 / 3.0, c13 = 1.0 / 3.0;
#define AA(s, i, j, k) A[(long)(s) * n + ic(&L, (i), (j), (k))]
    /* x- (2768-2781) */
    if (g->bc[0] == 3) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) { AA(6, 0, j, k) = AA(6, 0, j, k) - AA(0, 0, j, k); AA(0, 0, j, k) = 0.0; }
    } else if (g->bc[0] == 5) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                AA(7, 0, j, k) = c23 * bp[0];
                AA(6, 0, j, k) = 1.0;
                for (int s = 0; s < 6; ++s) AA(s, 0, j, k) = 0.0;
                AA(1, 0, j, k) = c13;
            }
        P_bc[0] = 1.0;
    }
    /* x+ (2784-2801) */
    if (g->bc[3] == 4 || g->bc[3] == 3) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) { AA(6, ie, j, k) = AA(6, ie, j, k) - AA(1, ie, j, k); AA(1, ie, j, k) = 0.0; }
    } else if (g->bc[3] == 5) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                AA(7, ie, j, k) = c23 * bp[1];
                AA(6, ie, j, k) = 1.0;
                for (int s = 1; s < 6; ++s) AA(s, ie, j, k) = 0.0;
                AA(0, ie, j, k) = c13;
            }
        P_bc[1] = 1.0;
    } else if (g->bc[3] == 6) {
        axis_t bx;
        if (axis_alloc(&bx, g->nx, g->ng, g->xlength)) return ORC_ERR_NOMEM;
        const double dxe = bx.dx[ie + G];
        axis_free(&bx);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                AA(6, ie, j, k) = AA(6, ie, j, k) - AA(0, ie, j, k) * 2.0 / 3.0 + AA(1, ie, j, k) * 5.0 / 3.0;
                AA(7, ie, j, k) = AA(7, ie, j, k) + 8.0 / 3.0 * AA(1, ie, j, k) * bp[1] +
                                  (u[ip(&L, ie, j, k)] - u[ip(&L, ie - 1, j, k)]) / dxe;
                AA(0, ie, j, k) = AA(0, ie, j, k) * 1.0 / 3.0;
                AA(1, ie, j, k) = 0.0;
            }
    }
    /* y- (2804-2822) */
    if (g->bc[1] == 3) {
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) { AA(6, i, 0, k) = AA(6, i, 0, k) - AA(2, i, 0, k); AA(2, i, 0, k) = 0.0; }
    } else if (g->bc[1] == 5) {
        for (int k = 0; k < nz; ++k) {
            for (int i = 0; i < nx; ++i) AA(7, i, 0, k) = c23 * bp[2];
            AA(7, 0, 0, k) = c23 * (bp[2] + bp[0] * P_bc[0]);
            AA(7, ie, 0, k) = c23 * (bp[2] + bp[1] * P_bc[1]);
            for (int i = 0; i < nx; ++i) AA(6, i, 0, k) = 1.0;
            AA(6, 0, 0, k) = 1.0 + P_bc[0];
            AA(6, ie, 0, k) = 1.0 + P_bc[1];
            for (int i = 0; i < nx; ++i)
                for (int s = 0; s < 6; ++s) AA(s, i, 0, k) = 0.0;
            AA(1, 0, 0, k) = c13 * P_bc[0];
            AA(0, ie, 0, k) = c13 * P_bc[1];
            for (int i = 0; i < nx; ++i) AA(3, i, 0, k) = c13;
        }
        P_bc[2] = 1.0;
    }
    /* y+ (2824-2843) */
    if (g->bc[4] == 3) {
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) { AA(6, i, je, k) = AA(6, i, je, k) - AA(3, i, je, k); AA(3, i, je, k) = 0.0; }
    } else if (g->bc[4] == 5) {
        for (int k = 0; k < nz; ++k) {
            for (int i = 0; i < nx; ++i) AA(7, i, je, k) = c23 * bp[3];
            AA(7, 0, je, k) = c23 * (bp[3] + bp[0] * P_bc[0]);
            AA(7, ie, je, k) = c23 * (bp[3] + bp[1] * P_bc[1]);
            for (int i = 0; i < nx; ++i) AA(6, i, je, k) = 1.0;
            AA(6, 0, je, k) = 1.0 + P_bc[0];
            AA(6, ie, je, k) = 1.0 + P_bc[1];
            for (int i = 0; i < nx; ++i)
                for (int s = 0; s < 6; ++s) AA(s, i, je, k) = 0.0;
            AA(1, 0, je, k) = c13 * P_bc[0];
            AA(0, ie, je, k) = c13 * P_bc[1];
            for (int i = 0; i < nx; ++i) AA(2, i, je, k) = c13;
        }
        P_bc[3] = 1.0;
    }
    /* z- (2845-2867) and z+ (2869-2891) */
    for (int side = 0; side < 2; ++side) {
        const int code = g->bc[side ? 5 : 2];
        const int kk = side ? ke : 0;
        const double P = bp[side ? 5 : 4];
        if (code == 3) {
            const int s0 = side ? 5 : 4;
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) { AA(6, i, j, kk) = AA(6, i, j, kk) - AA(s0, i, j, kk); AA(s0, i, j, kk) = 0.0; }
        } else if (code == 5) {
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) AA(7, i, j, kk) = c23 * P;
            AA(7, 0, 0, kk) = c23 * ((P + bp[0] * P_bc[0]) + bp[2] * P_bc[2]);
            AA(7, ie, 0, kk) = c23 * ((P + bp[1] * P_bc[1]) + bp[2] * P_bc[2]);
            AA(7, 0, je, kk) = c23 * ((P + bp[0] * P_bc[0]) + bp[3] * P_bc[3]);
            AA(7, ie, je, kk) = c23 * ((P + bp[1] * P_bc[1]) + bp[3] * P_bc[3]);
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) AA(6, i, j, kk) = 1.0;
            AA(6, 0, 0, kk) = (1.0 + P_bc[0]) + P_bc[2];
            AA(6, ie, 0, kk) = (1.0 + P_bc[1]) + P_bc[2];
            AA(6, 0, je, kk) = (1.0 + P_bc[0]) + P_bc[3];
            AA(6, ie, je, kk) = (1.0 + P_bc[1]) + P_bc[3];
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    for (int s = 0; s < 6; ++s) AA(s, i, j, kk) = 0.0;
            /* the y-coupling corner entries: the ks plane in both blocks */
            AA(1, 0, 0, kk) = c13 * P_bc[0]; AA(3, 0, 0, 0) = c13 * P_bc[2];
            AA(0, ie, 0, kk) = c13 * P_bc[1]; AA(3, ie, 0, 0) = c13 * P_bc[2];
            AA(1, 0, je, kk) = c13 * P_bc[0]; AA(2, 0, je, 0) = c13 * P_bc[3];
            AA(0, ie, je, kk) = c13 * P_bc[1]; AA(2, ie, je, 0) = c13 * P_bc[3];
            const int s1 = side ? 4 : 5;
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) AA(s1, i, j, kk) = c13;
        }
    }
#undef AA

    /* check_and_debug_Poisson: a tiny diagonal is only legal in a solid
     * (all six face masks closed), where A7 becomes 1 */
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long c = ic(&L, i, j, k);
                if (!(A[6 * n + c] < 1e-50)) continue;
                int solid = umask[ip(&L, i - 1, j, k)] < 0.5 && umask[ip(&L, i, j, k)] < 0.5 &&
                            vmask[ip(&L, i, j - 1, k)] < 0.5 && vmask[ip(&L, i, j, k)] < 0.5 &&
                            wmask[ip(&L, i, j, k - 1)] < 0.5 && wmask[ip(&L, i, j, k)] < 0.5;
                if (!solid) return ORC_ERR_SOLID;
                if (A[7 * n + c] > 1e-50) return ORC_ERR_SOLID;
                double mx = A[c], mn = A[c];
                for (int s = 1; s < 6; ++s) {
                    if (A[s * n + c] > mx) mx = A[s * n + c];
                    if (A[s * n + c] < mn) mn = A[s * n + c];
                }
                if (mx > 1e-50 || mn < 0.0) return ORC_ERR_SOLID;
                A[6 * n + c] = 1.0;
            }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* check_poisson_setup: newsolver.f90:275-344                          */
/* ------------------------------------------------------------------ */
int orc_check_poisson_setup(const orc_grid *g, const double *A,
                            const double *umask, const double *vmask,
                            const double *wmask, double *rowsum) {
    lvl_t L = grid_lvl(g);
    long n = ncell(&L);
    for (long c = 0; c < n; ++c) {
        for (int s = 0; s < 8; ++s)
            if (A[s * n + c] != A[s * n + c]) return ORC_ERR_NAN_OR_TINY_DIAG;
        if (A[6 * n + c] < 1e-55) return ORC_ERR_NAN_OR_TINY_DIAG;
    }
    /* Fortran sum() over the whole plane, then divided by the cell count;
     * the test is one-sided as in the reference */
    double src = 0.0;
    for (long c = 0; c < n; ++c) src += A[7 * n + c];
    src = src / (double)n;
    if (src > 1e-10) return ORC_ERR_VOLUME_SOURCE;
    double mx = -1e50, mn = 1e50;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long c = ic(&L, i, j, k);
                double a7 = A[6 * n + c];
                if (umask[ip(&L, i - 1, j, k)] < 0.5 && umask[ip(&L, i, j, k)] < 0.5 &&
                    vmask[ip(&L, i, j - 1, k)] < 0.5 && vmask[ip(&L, i, j, k)] < 0.5 &&
                    wmask[ip(&L, i, j, k - 1)] < 0.5 && wmask[ip(&L, i, j, k)] < 0.5)
                    a7 = 0.0;
                double t = (((((A[c] + A[n + c]) + A[2 * n + c]) + A[3 * n + c]) + A[4 * n + c]) +
                            A[5 * n + c]) - a7;
                if (rowsum) rowsum[c] = t;
                if (t > mx) mx = t;
                if (t < mn) mn = t;
            }
    double big = mx > -mn ? mx : -mn;
    if (big > 1e-10) return ORC_ERR_NOT_IN_KERNEL;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* residuals                                                            */
/* ------------------------------------------------------------------ */
/* get_residual / compute_residual expression: MGsolver.f90:97-100, 119-123 */
static inline double resid_mg(const double *A, long n, long c, const double *p,
                              long q, long sx, long sy, long sz) {
    return A[7 * n + c] - A[6 * n + c] * p[q] + A[0 * n + c] * p[q - sx] +
           A[1 * n + c] * p[q + sx] + A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] +
           A[4 * n + c] * p[q - sz] + A[5 * n + c] * p[q + sz];
}

void orc_residual_field(const orc_grid *g, const double *A, const double *p,
                        double *r) {
    lvl_t L = grid_lvl(g);
    long n = ncell(&L), sx = 1, sy = L.X, sz = (long)L.X * L.Y;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long c = ic(&L, i, j, k);
                r[c] = resid_mg(A, n, c, p, ip(&L, i, j, k), sx, sy, sz);
            }
}

static double powabs(double r, int norm) {
    double a = fabs(r);
    if (norm == 1) return a;
    if (norm == 2) return a * a;
    return pow(a, (double)norm);
}

static double compute_residual_lvl(const lvl_t *L, const double *A,
                                   const double *p, int norm) {
    long n = ncell(L), sx = 1, sy = L->X, sz = (long)L->X * L->Y;
    double acc = 0.0;
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) reduction(+ : acc) schedule(static)
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) {
                long c = ic(L, i, j, k);
                acc = acc + powabs(resid_mg(A, n, c, p, ip(L, i, j, k), sx, sy, sz), norm);
            }
    return acc;
}

double orc_compute_residual(const orc_grid *g, const double *A,
                            const double *p, int norm) {
    lvl_t L = grid_lvl(g);
    return compute_residual_lvl(&L, A, p, norm);
}

/* calcResidual (modules.f90:3339-3384) without ExcludeBoundCellCalcRes */
void orc_calc_residuals(const orc_grid *g, const double *A, const double *p,
                        double out[3]) {
    lvl_t L = grid_lvl(g);
    long n = ncell(&L), sx = 1, sy = L.X, sz = (long)L.X * L.Y;
    double s1 = 0.0, s2 = 0.0, smax = 0.0;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long c = ic(&L, i, j, k), q = ip(&L, i, j, k);
                double loc = fabs(-p[q] * A[6 * n + c] + A[0 * n + c] * p[q - sx] +
                                  A[1 * n + c] * p[q + sx] + A[2 * n + c] * p[q - sy] +
                                  A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                                  A[5 * n + c] * p[q + sz] + A[7 * n + c]);
                s1 = s1 + loc;
                s2 = s2 + loc * loc;
                if (loc > smax) smax = loc;
            }
    out[0] = s1 / (double)g->nx / (double)g->ny / (double)g->nz;
    out[1] = sqrt(s2 / (double)g->nx / (double)g->ny / (double)g->nz);
    out[2] = smax;
}

/* ------------------------------------------------------------------ */
/* ghost exchange, single rank: ghost_MG_x/y/z (modules.f90:1846-1971)  */
/* ------------------------------------------------------------------ */
static void apply_bc_lvl(const lvl_t *L, double *p) {
    const int G = L->ng;
    if (L->per[0])
        for (int k = -G; k < L->n[2] + G; ++k)
            for (int j = -G; j < L->n[1] + G; ++j) {
                p[ip(L, -1, j, k)] = p[ip(L, L->n[0] - 1, j, k)];
                p[ip(L, L->n[0], j, k)] = p[ip(L, 0, j, k)];
            }
    if (L->per[1])
        for (int k = -G; k < L->n[2] + G; ++k)
            for (int i = -G; i < L->n[0] + G; ++i) {
                p[ip(L, i, -1, k)] = p[ip(L, i, L->n[1] - 1, k)];
                p[ip(L, i, L->n[1], k)] = p[ip(L, i, 0, k)];
            }
    if (L->per[2])
        for (int j = -G; j < L->n[1] + G; ++j)
            for (int i = -G; i < L->n[0] + G; ++i) {
                p[ip(L, i, j, -1)] = p[ip(L, i, j, L->n[2] - 1)];
                p[ip(L, i, j, L->n[2])] = p[ip(L, i, j, 0)];
            }
}

void orc_apply_bc(const orc_grid *g, double *p) {
    lvl_t L = grid_lvl(g);
    apply_bc_lvl(&L, p);
}

/* ------------------------------------------------------------------ */
/* RedBlackRelax: newsolver.f90:209-247.  Pass 1 updates cells whose    */
/* local index parity (i+j+k, 0-based) is even, then the halo is         */
/* exchanged, then pass 2 updates the odd cells.                         */
/* ------------------------------------------------------------------ */
static void rb_relax_lvl(const lvl_t *L, const double *A, double *p, double beta) {
    long n = ncell(L), sx = 1, sy = L->X, sz = (long)L->X * L->Y;
    for (int pass = 0; pass < 2; ++pass) {
        /* the cells of one colour are independent */
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
        for (int k = 0; k < L->n[2]; ++k)
            for (int j = 0; j < L->n[1]; ++j)
                for (int i = (j + k + pass) & 1; i < L->n[0]; i += 2) {
                    long c = ic(L, i, j, k), q = ip(L, i, j, k);
                    p[q] = (1.0 - beta) * p[q] +
                           beta / A[6 * n + c] *
                               (A[0 * n + c] * p[q - sx] + A[1 * n + c] * p[q + sx] +
                                A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] +
                                A[4 * n + c] * p[q - sz] + A[5 * n + c] * p[q + sz] +
                                A[7 * n + c]);
                }
        if (pass == 0) apply_bc_lvl(L, p);
    }
}

void orc_redblack_relax(const orc_grid *g, const double *A, double *p,
                        double beta) {
    lvl_t L = grid_lvl(g);
    rb_relax_lvl(&L, A, p, beta);
}

/* get_residual, in place on slot 8: MGsolver.f90:84-104 */
static void get_residual_lvl(const lvl_t *L, double *A, const double *p) {
    long n = ncell(L), sx = 1, sy = L->X, sz = (long)L->X * L->Y;
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
    for (int k = 0; k < L->n[2]; ++k)
        for (int j = 0; j < L->n[1]; ++j)
            for (int i = 0; i < L->n[0]; ++i) {
                long c = ic(L, i, j, k);
                A[7 * n + c] = resid_mg(A, n, c, p, ip(L, i, j, k), sx, sy, sz);
            }
}

/* ------------------------------------------------------------------ */
/* transfers: MGsolver.f90:128-251                                      */
/* ------------------------------------------------------------------ */
/* fill_coefficients (203-220): every slot of a coarse cell is the mean of
 * its 8 children, summed in the reference's child order */
void orc_fill_coefficients(const int nf[3], const double *Af, double *Ac) {
    int nc[3] = {nf[0] / 2, nf[1] / 2, nf[2] / 2};
    long n_f = (long)nf[0] * nf[1] * nf[2], n_c = (long)nc[0] * nc[1] * nc[2];
#define FI(i, j, k) ((long)(i) + (long)nf[0] * ((long)(j) + (long)nf[1] * (k)))
    for (int s = 0; s < 8; ++s)
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
        for (int k = 0; k < nc[2]; ++k)
            for (int j = 0; j < nc[1]; ++j)
                for (int i = 0; i < nc[0]; ++i) {
                    const double *a = Af + s * n_f;
                    int I = 2 * i, J = 2 * j, K = 2 * k;
                    double t = a[FI(I, J, K)] + a[FI(I, J + 1, K)] + a[FI(I, J, K + 1)] +
                               a[FI(I, J + 1, K + 1)] + a[FI(I + 1, J, K)] + a[FI(I + 1, J + 1, K)] +
                               a[FI(I + 1, J, K + 1)] + a[FI(I + 1, J + 1, K + 1)];
                    Ac[s * n_c + i + (long)nc[0] * (j + (long)nc[1] * k)] = 0.125 * t;
                }
}

/* coarse_from_fine (185-201): coarse slot 8 = sum over children, i1 outer,
 * k1 inner, of 0.125 * child slot 8 */
void orc_coarse_from_fine(const int nf[3], const double *Af, double *Ac) {
    int nc[3] = {nf[0] / 2, nf[1] / 2, nf[2] / 2};
    long n_f = (long)nf[0] * nf[1] * nf[2], n_c = (long)nc[0] * nc[1] * nc[2];
    const double *a = Af + 7 * n_f;
    double *o = Ac + 7 * n_c;
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
    for (int i = 0; i < nc[0]; ++i)
        for (int j = 0; j < nc[1]; ++j)
            for (int k = 0; k < nc[2]; ++k) {
                double t = 0.0;
                for (int i1 = 0; i1 < 2; ++i1)
                    for (int j1 = 0; j1 < 2; ++j1)
                        for (int k1 = 0; k1 < 2; ++k1)
                            t = t + a[FI(2 * i + i1, 2 * j + j1, 2 * k + k1)] * 0.125;
                o[i + (long)nc[0] * (j + (long)nc[1] * k)] = t;
            }
}
#undef FI

/* coarse_fine_interp (128-183) */
void orc_coarse_fine_interp(const int nf[3], int ng, int sym_mg, double *ef,
                            const double *ec) {
    int nc[3] = {nf[0] / 2, nf[1] / 2, nf[2] / 2};
    int per0[3] = {0, 0, 0};
    lvl_t F = lvl_make(nf, ng, per0), C = lvl_make(nc, ng, per0);
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
    for (int i = 0; i < nc[0]; ++i)
        for (int j = 0; j < nc[1]; ++j)
            for (int k = 0; k < nc[2]; ++k) {
                double pc = ec[ip(&C, i, j, k)];
                if (sym_mg) {
                    /* the children are weighted by |ef|; the reference sums in
                     * the order (000)(100)(010)(110)(001)(101)(011)(111) */
                    static const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                                {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
                    double tot = 0.0;
                    for (int m = 0; m < 8; ++m)
                        tot = tot + fabs(ef[ip(&F, 2 * i + o[m][0], 2 * j + o[m][1], 2 * k + o[m][2])]);
                    tot = tot + 1e-15;
                    for (int m = 0; m < 8; ++m) {
                        long q = ip(&F, 2 * i + o[m][0], 2 * j + o[m][1], 2 * k + o[m][2]);
                        ef[q] = fabs(ef[q]) / tot * pc;
                    }
                } else {
                    for (int di = 0; di < 2; ++di)
                        for (int dj = 0; dj < 2; ++dj)
                            for (int dk = 0; dk < 2; ++dk)
                                ef[ip(&F, 2 * i + di, 2 * j + dj, 2 * k + dk)] = pc;
                }
            }
}

/* ------------------------------------------------------------------ */
/* NewSolverMG + MG_iteration + init_MG: MGsolver.f90:34-65, 253-405      */
/* ------------------------------------------------------------------ */
typedef struct {
    int lmg;
    lvl_t *lv;       /* lv[1..lmg] */
    double **data;   /* 8-slot level operators */
    double **pmg, **emg;
} mg_t;

static void mg_free(mg_t *m) {
    if (!m->lv) return;
    for (int l = 1; l <= m->lmg; ++l) { free(m->data[l]); free(m->pmg[l]); free(m->emg[l]); }
    free(m->lv); free(m->data); free(m->pmg); free(m->emg);
    m->lv = NULL;
}

static int mg_init(mg_t *m, const orc_grid *g) {
    int nd[3] = {g->nx, g->ny, g->nz};
    for (int d = 0; d < 3; ++d)
        if (nd[d] <= 0 || (nd[d] & (nd[d] - 1))) return ORC_ERR_BAD_SIZE;
    int lmg = 1 << 30;
    for (int d = 0; d < 3; ++d) {
        int l = 0;
        while ((1 << (l + 1)) <= nd[d]) ++l;
        if (l < lmg) lmg = l;
    }
    if (lmg < 2) return ORC_ERR_BAD_SIZE;  /* the V-cycle touches level 2 */
    m->lmg = lmg;
    m->lv = calloc(lmg + 1, sizeof(lvl_t));
    m->data = calloc(lmg + 1, sizeof(double *));
    m->pmg = calloc(lmg + 1, sizeof(double *));
    m->emg = calloc(lmg + 1, sizeof(double *));
    if (!m->lv || !m->data || !m->pmg || !m->emg) return ORC_ERR_NOMEM;
    int per[3] = {g->bc[0] == 1, g->bc[1] == 1, g->bc[2] == 1};
    for (int l = lmg; l >= 1; --l) {
        int nl[3] = {nd[0] >> (lmg - l), nd[1] >> (lmg - l), nd[2] >> (lmg - l)};
        m->lv[l] = lvl_make(nl, g->ng, per);
        m->data[l] = calloc((size_t)ncell(&m->lv[l]) * 8, sizeof(double));
        m->pmg[l] = calloc((size_t)npad(&m->lv[l]), sizeof(double));
        m->emg[l] = calloc((size_t)npad(&m->lv[l]), sizeof(double));
        if (!m->data[l] || !m->pmg[l] || !m->emg[l]) return ORC_ERR_NOMEM;
    }
    return ORC_OK;
}

static void vzero(double *a, long n) {
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
    for (long q = 0; q < n; ++q) a[q] = 0.0;
}
static void vadd(double *a, const double *b, long n) {
#pragma omp parallel for if (orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)
    for (long q = 0; q < n; ++q) a[q] = a[q] + b[q];
}

/* MG_iteration (253-320) */
static void mg_iteration(mg_t *m, double beta, int nrelax, int sym_mg) {
    int lmg = m->lmg;
    for (int l = lmg; l >= 2; --l) {
        lvl_t *L = &m->lv[l];
        vzero(m->pmg[l], npad(L));
        for (int r = 0; r < nrelax; ++r) {
            rb_relax_lvl(L, m->data[l], m->pmg[l], beta);
            apply_bc_lvl(L, m->pmg[l]);
        }
        get_residual_lvl(L, m->data[l], m->pmg[l]);
        orc_coarse_from_fine(L->n, m->data[l], m->data[l - 1]);
    }
    rb_relax_lvl(&m->lv[2], m->data[2], m->pmg[2], beta);
    apply_bc_lvl(&m->lv[2], m->pmg[2]);
    for (int l = 3; l <= lmg; ++l) {
        lvl_t *L = &m->lv[l];
        long np = npad(L);
        vzero(m->emg[l], np);
        orc_coarse_fine_interp(L->n, L->ng, sym_mg, m->emg[l], m->pmg[l - 1]);
        apply_bc_lvl(L, m->emg[l]);
        get_residual_lvl(L, m->data[l], m->emg[l]);
        vadd(m->pmg[l], m->emg[l], np);
        vzero(m->emg[l], np);
        for (int r = 0; r < nrelax; ++r) {
            rb_relax_lvl(L, m->data[l], m->emg[l], beta);
            apply_bc_lvl(L, m->emg[l]);
        }
        vadd(m->pmg[l], m->emg[l], np);
        get_residual_lvl(L, m->data[l], m->emg[l]);
    }
}

int orc_newsolver_mg(const orc_grid *g, const double *A, double *p,
                     double maxError, double beta, int maxit, int norm,
                     int nrelax, int sym_mg, int *it, double *tres2,
                     double *hist) {
    mg_t m;
    memset(&m, 0, sizeof m);
    int rc = mg_init(&m, g);
    if (rc) { mg_free(&m); return rc; }
    lvl_t *F = &m.lv[m.lmg];
    long n = ncell(F), np = npad(F);
    double N = (double)g->nx * (double)g->ny * (double)g->nz;
    /* get_MGmatrix_coef (322-344); level 1 is filled but never used */
    memcpy(m.data[m.lmg], A, sizeof(double) * 8 * n);
    for (int l = m.lmg - 1; l >= 1; --l) orc_fill_coefficients(m.lv[l + 1].n, m.data[l + 1], m.data[l]);
    /* initial residual into slot 8 (370-377) */
    memcpy(m.pmg[m.lmg], p, sizeof(double) * np);
    double res0 = compute_residual_lvl(F, m.data[m.lmg], m.pmg[m.lmg], norm);
    get_residual_lvl(F, m.data[m.lmg], m.pmg[m.lmg]);
    double t2 = res0 / N;
    vzero(m.pmg[m.lmg], np);
    int k;
    for (k = 1; k <= maxit; ++k) {
        mg_iteration(&m, beta, nrelax, sym_mg);
        vadd(p, m.pmg[m.lmg], np);
        double res = compute_residual_lvl(F, A, p, norm);
        t2 = res / N;
        if (norm == 2) t2 = sqrt(t2);
        if (hist) hist[k - 1] = t2;
        if (t2 < maxError) break;
    }
    *it = k;
    *tres2 = t2;
    mg_free(&m);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* NewSolver_std: newsolver.f90:78-174                                  */
/* ------------------------------------------------------------------ */
int orc_newsolver_std(const orc_grid *g, const double *A, double *p,
                      double maxError, double beta, int maxit, int norm,
                      int *it, double *tres2, double *hist) {
    lvl_t L = grid_lvl(g);
    long n = ncell(&L), sx = 1, sy = L.X, sz = (long)L.X * L.Y;
    double N = (double)g->nx * (double)g->ny * (double)g->nz;
    double t2 = 0.0;
    int k;
    for (k = 1; k <= maxit; ++k) {
        rb_relax_lvl(&L, A, p, beta);
        /* the halo exchange overlaps the interior sum; the boundary shell is
         * summed after it, so the total runs interior first, shell second */
        apply_bc_lvl(&L, p);
        double res2 = 0.0;
        for (int pass = 0; pass < 2; ++pass)
            for (int kk = 0; kk < g->nz; ++kk)
                for (int j = 0; j < g->ny; ++j)
                    for (int i = 0; i < g->nx; ++i) {
                        int inner = i > 0 && i < g->nx - 1 && j > 0 && j < g->ny - 1 &&
                                    kk > 0 && kk < g->nz - 1;
                        if (inner != (pass == 0)) continue;
                        long c = ic(&L, i, j, kk), q = ip(&L, i, j, kk);
                        double r = -p[q] * A[6 * n + c] + A[0 * n + c] * p[q - sx] +
                                   A[1 * n + c] * p[q + sx] + A[2 * n + c] * p[q - sy] +
                                   A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                                   A[5 * n + c] * p[q + sz] + A[7 * n + c];
                        res2 = res2 + powabs(r, norm);
                    }
        res2 = res2 / N;
        /* catch_divergence (138-173) */
        for (int kk = 0; kk < g->nz; ++kk)
            for (int j = 0; j < g->ny; ++j)
                for (int i = 0; i < g->nx; ++i) {
                    long c = ic(&L, i, j, kk);
                    if (p[ip(&L, i, j, kk)] != p[ip(&L, i, j, kk)]) return ORC_ERR_DIVERGED;
                    for (int s = 0; s < 8; ++s)
                        if (A[s * n + c] != A[s * n + c]) return ORC_ERR_DIVERGED;
                }
        if (res2 > 1e16 || res2 != res2) { *it = k; *tres2 = res2; return ORC_ERR_DIVERGED; }
        t2 = res2;
        if (norm == 2) t2 = sqrt(t2);
        if (hist) hist[k - 1] = t2;
        if (t2 < maxError) break;
    }
    *it = k;
    *tres2 = t2;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* FreeSolver: freesurface.f90:700-855, RedBlackRelax_fs: 857-895        */
/* ------------------------------------------------------------------ */
static int fs_active(int m, int flag) {
    return flag == 1 ? m == 0 : (m == 1 || m == 2);   /* 880-881 */
}

int orc_free_solver(const orc_grid *g, const double *A, double *p,
                    const int *pcmask, const double *pgas, int solver_flag,
                    double maxError, double beta, int maxit, int *it,
                    double *tres2, double *hist) {
    if (solver_flag != 1 && solver_flag != 2) return ORC_ERR_UNSUPPORTED;   /* 722 */
    lvl_t L = grid_lvl(g);
    long n = ncell(&L), sx = 1, sy = L.X, sz = (long)L.X * L.Y;
    /* 723-726: fixed values of the cells the solve does not update */
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long q = ip(&L, i, j, k);
                if (solver_flag == 1 && pcmask[q] != 0) p[q] = pgas[q];
                if (solver_flag == 2 && pcmask[q] == 3) p[q] = 0.0;
            }
    apply_bc_lvl(&L, p);
    double t2 = 0.0;
    int it_ = 1;
    for (it_ = 1; it_ <= maxit; ++it_) {
        /* RedBlackRelax_fs: even local parity first, halo, then odd */
        for (int pass = 0; pass < 2; ++pass) {
            for (int k = 0; k < g->nz; ++k)
                for (int j = 0; j < g->ny; ++j)
                    for (int i = (j + k + pass) & 1; i < g->nx; i += 2) {
                        long c = ic(&L, i, j, k), q = ip(&L, i, j, k);
                        if (!fs_active(pcmask[q], solver_flag)) continue;
                        p[q] = (1.0 - beta) * p[q] +
                               beta / A[6 * n + c] *
                                   (A[0 * n + c] * p[q - sx] + A[1 * n + c] * p[q + sx] +
                                    A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] +
                                    A[4 * n + c] * p[q - sz] + A[5 * n + c] * p[q + sz] +
                                    A[7 * n + c]);
                    }
            if (pass == 0) apply_bc_lvl(&L, p);
        }
        apply_bc_lvl(&L, p);
        /* 738-785: interior cells first, then the boundary shell */
        double res2 = 0.0, resinf = 0.0, cells = 0.0;
        int nan = 0;
        for (int pass = 0; pass < 2; ++pass)
            for (int k = 0; k < g->nz; ++k)
                for (int j = 0; j < g->ny; ++j)
                    for (int i = 0; i < g->nx; ++i) {
                        int inner = i > 0 && i < g->nx - 1 && j > 0 && j < g->ny - 1 &&
                                    k > 0 && k < g->nz - 1;
                        if (inner != (pass == 0)) continue;
                        long c = ic(&L, i, j, k), q = ip(&L, i, j, k);
                        if (!fs_active(pcmask[q], solver_flag)) continue;
                        double r = fabs(-p[q] * A[6 * n + c] + A[0 * n + c] * p[q - sx] +
                                        A[1 * n + c] * p[q + sx] + A[2 * n + c] * p[q - sy] +
                                        A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                                        A[5 * n + c] * p[q + sz] + A[7 * n + c]);
                        res2 = res2 + r * r;
                        if (r > resinf) resinf = r;
                        cells = cells + 1.0;
                        if (r != r) nan = 1;
                        for (int s = 0; s < 8; ++s)
                            if (A[s * n + c] != A[s * n + c]) nan = 1;
                    }
        /* catch_divergence_fs (811-855) with one rank */
        if (nan || res2 > 1e18 || res2 != res2) { *it = it_; *tres2 = res2; return ORC_ERR_DIVERGED; }
        t2 = sqrt(res2) / cells;   /* 786-791, norm = 2 */
        if (hist) hist[it_ - 1] = t2;
        if (resinf < maxError) break;
    }
    *it = it_;
    *tres2 = t2;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Setup_testMG (modules.f90:3160-3186), get_MGtest_err (407-426)        */
/* `real(8) :: pi=3.14159265359` is a default-real literal: the value   */
/* stored is the single-precision rounding of pi.                        */
/* ------------------------------------------------------------------ */
static const float ORC_PI_SP = 3.14159265359f;

void orc_setup_testmg(const orc_grid *g, const double *rho, double *A) {
    lvl_t L = grid_lvl(g);
    long n = ncell(&L);
    double pi = (double)ORC_PI_SP;
    axis_t ax, ay, az;
    axis_alloc(&ax, g->nx, g->ng, g->xlength);
    axis_alloc(&ay, g->ny, g->ng, g->ylength);
    axis_alloc(&az, g->nz, g->ng, g->zlength);
    const int G = g->ng;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long c = ic(&L, i, j, k), q = ip(&L, i, j, k);
                int I = i + G, J = j + G, K = k + G;
                double r0 = rho[q];
                double a1 = 2.0 / (ax.dx[I] * ax.dxh[I - 1] * (rho[ip(&L, i - 1, j, k)] + r0));
                double a2 = 2.0 / (ax.dx[I] * ax.dxh[I] * (rho[ip(&L, i + 1, j, k)] + r0));
                double a3 = 2.0 / (ay.dx[J] * ay.dxh[J - 1] * (rho[ip(&L, i, j - 1, k)] + r0));
                double a4 = 2.0 / (ay.dx[J] * ay.dxh[J] * (rho[ip(&L, i, j + 1, k)] + r0));
                double a5 = 2.0 / (az.dx[K] * az.dxh[K - 1] * (rho[ip(&L, i, j, k - 1)] + r0));
                double a6 = 2.0 / (az.dx[K] * az.dxh[K] * (rho[ip(&L, i, j, k + 1)] + r0));
                A[c] = a1; A[n + c] = a2; A[2 * n + c] = a3;
                A[3 * n + c] = a4; A[4 * n + c] = a5; A[5 * n + c] = a6;
                A[6 * n + c] = ((((a1 + a2) + a3) + a4) + a5) + a6;
                A[7 * n + c] = 3.0 * (2.0 * pi) * (2.0 * pi) * sin(2.0 * pi * ax.x[I]) *
                               sin(2.0 * pi * ay.x[J]) * sin(2.0 * pi * az.x[K]);
            }
    axis_free(&ax); axis_free(&ay); axis_free(&az);
}

double orc_get_mgtest_err(const orc_grid *g, const double *p) {
    lvl_t L = grid_lvl(g);
    double pi = (double)ORC_PI_SP;
    axis_t ax, ay, az;
    axis_alloc(&ax, g->nx, g->ng, g->xlength);
    axis_alloc(&ay, g->ny, g->ng, g->ylength);
    axis_alloc(&az, g->nz, g->ng, g->zlength);
    const int G = g->ng;
    double res = 0.0, ref = 0.0;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                double s = sin(2.0 * pi * ax.x[i + G]) * sin(2.0 * pi * ay.x[j + G]) *
                           sin(2.0 * pi * az.x[k + G]);
                res = res + fabs(p[ip(&L, i, j, k)] - s);
                ref = ref + fabs(s);
            }
    axis_free(&ax); axis_free(&ay); axis_free(&az);
    return res / ref;
}

/* ------------------------------------------------------------------ */
/* project_velocity, staggered branch: paris.f90:898-908                 */
/* ------------------------------------------------------------------ */
void orc_project_velocity(const orc_grid *g, double *u, double *v, double *w,
                          const double *umask, const double *vmask,
                          const double *wmask, const double *rho,
                          const double *p, double dt) {
    lvl_t L = grid_lvl(g);
    axis_t ax, ay, az;
    axis_alloc(&ax, g->nx, g->ng, g->xlength);
    axis_alloc(&ay, g->ny, g->ng, g->ylength);
    axis_alloc(&az, g->nz, g->ng, g->zlength);
    const int G = g->ng;
    /* ieu/jev/kew: the last face is skipped unless the direction is periodic */
    int ieu = g->nx - ((g->bc[0] != 1) ? 1 : 0);
    int jev = g->ny - ((g->bc[1] != 1) ? 1 : 0);
    int kew = g->nz - ((g->bc[2] != 1) ? 1 : 0);
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < ieu; ++i) {
                long q = ip(&L, i, j, k), e = ip(&L, i + 1, j, k);
                u[q] = u[q] - dt * (2.0 * umask[q] / ax.dxh[i + G]) * (p[e] - p[q]) / (rho[e] + rho[q]);
            }
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < jev; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long q = ip(&L, i, j, k), e = ip(&L, i, j + 1, k);
                v[q] = v[q] - dt * (2.0 * vmask[q] / ay.dxh[j + G]) * (p[e] - p[q]) / (rho[e] + rho[q]);
            }
    for (int k = 0; k < kew; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                long q = ip(&L, i, j, k), e = ip(&L, i, j, k + 1);
                w[q] = w[q] - dt * (2.0 * wmask[q] / az.dxh[k + G]) * (p[e] - p[q]) / (rho[e] + rho[q]);
            }
    axis_free(&ax); axis_free(&ay); axis_free(&az);
}

/* ------------------------------------------------------------------ */
/* Independent exact solver: Jacobi-PCG on the 8-slot operator            */
/* M p = A7 p - sum A_nb p_nb, with wall ghosts held at their input value */
/* (their contribution moves to the right-hand side) and periodic ghosts  */
/* wrapped.  Not a reference algorithm: it pins converged solutions.      */
/* ------------------------------------------------------------------ */
/* Threads of the Jacobi-PCG below (a checker of converged solutions, not
 * a restatement of a reference loop): all host threads by default.  The
 * thread count changes only the summation order of its dot products. */
static int orc_pcg_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* y = M x on the interior and the dot product x.y; x is a padded vector
 * whose wall ghosts are 0 */
static double op_apply_dot(const lvl_t *L, const double *A, double *x, double *y) {
    apply_bc_lvl(L, x);
    const long n = ncell(L), sx = 1, sy = L->X, sz = (long)L->X * L->Y;
    const int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    double dot = 0.0;
#pragma omp parallel for collapse(2) num_threads(orc_pcg_threads()) reduction(+ : dot) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long c = ic(L, i, j, k), q = ip(L, i, j, k);
                double v = A[6 * n + c] * x[q] - (A[0 * n + c] * x[q - sx] + A[1 * n + c] * x[q + sx] +
                                                  A[2 * n + c] * x[q - sy] + A[3 * n + c] * x[q + sy] +
                                                  A[4 * n + c] * x[q - sz] + A[5 * n + c] * x[q + sz]);
                y[c] = v;
                dot += x[q] * v;
            }
    return dot;
}

static double true_residual(const orc_grid *g, const lvl_t *L, const double *A, const double *p,
                            double *pc, double *r) {
    const long n = ncell(L), np = npad(L);
    memcpy(pc, p, sizeof(double) * np);
    apply_bc_lvl(L, pc);
    orc_residual_field(g, A, pc, r);
    double rn = 0.0;
#pragma omp parallel for num_threads(orc_pcg_threads()) reduction(+ : rn) schedule(static)
    for (long c = 0; c < n; ++c) rn += r[c] * r[c];
    return rn;
}

/* Jacobi-preconditioned CG on the 8-slot system (M = diag(A7) -
 * offdiag(A1..A6), the residual of MGsolver.f90:97-100), to ||r||/||A8|| <=
 * rtol.  Independent of every multigrid component: used only to pin
 * converged solutions. */
int orc_pcg_solve(const orc_grid *g, const double *A, double *p, double rtol,
                  int maxit, double *relres) {
    lvl_t L = grid_lvl(g);
    long n = ncell(&L), np = npad(&L);
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *q = malloc(sizeof(double) * n), *d = calloc(np, sizeof(double));
    double *pc = malloc(sizeof(double) * np);
    if (!r || !z || !q || !d || !pc) { free(r); free(z); free(q); free(d); free(pc); return -1; }
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    /* r = b - M p using the residual of the reference (ghosts included) */
    double rn = true_residual(g, &L, A, p, pc, r);
    double bn = 0.0;
    for (long c = 0; c < n; ++c) bn += A[7 * n + c] * A[7 * n + c];
    bn = sqrt(bn);
    if (bn == 0.0) bn = 1.0;
    double rz = 0.0;
#pragma omp parallel for num_threads(orc_pcg_threads()) reduction(+ : rz) schedule(static)
    for (long c = 0; c < n; ++c) { z[c] = r[c] / A[6 * n + c]; rz += r[c] * z[c]; }
    int it = 0;
#pragma omp parallel for collapse(2) num_threads(orc_pcg_threads()) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) d[ip(&L, i, j, k)] = z[ic(&L, i, j, k)];
    while (sqrt(rn) / bn > rtol && it < maxit) {
        const double dq = op_apply_dot(&L, A, d, q);
        const double alpha = rz / dq;
        double rn2 = 0.0, rz2 = 0.0;
#pragma omp parallel for collapse(2) num_threads(orc_pcg_threads()) reduction(+ : rn2, rz2) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    long c = ic(&L, i, j, k), qq = ip(&L, i, j, k);
                    p[qq] += alpha * d[qq];
                    r[c] -= alpha * q[c];
                    rn2 += r[c] * r[c];
                    z[c] = r[c] / A[6 * n + c];
                    rz2 += r[c] * z[c];
                }
        rn = rn2;
        const double beta = rz2 / rz;
        rz = rz2;
#pragma omp parallel for collapse(2) num_threads(orc_pcg_threads()) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    long qq = ip(&L, i, j, k);
                    d[qq] = z[ic(&L, i, j, k)] + beta * d[qq];
                }
        ++it;
        /* recompute the true residual now and then to avoid drift */
        if (it % 50 == 0) rn = true_residual(g, &L, A, p, pc, r);
    }
    apply_bc_lvl(&L, p);
    rn = true_residual(g, &L, A, p, pc, r);
    if (relres) *relres = sqrt(rn) / bn;
    free(r); free(z); free(q); free(d); free(pc);
    return it;
}

/* ------------------------------------------------------------------ */
/* linfunc: a material property from the volume fraction               */
/* (vofmodules.f90:133-147 weights, 171-218 linfunc), then do_all_ghost */
/* (modules.f90:2062-2072): ghost_x, ghost_y, ghost_z over the full     */
/* extent of the other two axes, i.e. a periodic wrap of the ng ghost   */
/* layers axis by axis, and nothing at walls (MPI_PROC_NULL).           */
/* ------------------------------------------------------------------ */
int orc_linfunc(const orc_grid *g, const double *cvof, double a1, double a2, int mean_flag,
                int nfilter, double *field) {
    if (mean_flag != 101 && mean_flag != 102) return -1;
    if (nfilter != 0 && nfilter != 1) return -1;
    lvl_t L = grid_lvl(g);
    const int X = L.X, Y = L.Y, Z = L.Z;
    const long sy = X, sz = (long)X * Y;
    const double b1 = 8.0 / 64, b2 = 4.0 / 64, b3 = 2.0 / 64, b4 = 1.0 / 64;
    const double inva1 = 1.0 / (a1 + 1e-50), inva2 = 1.0 / (a2 + 1e-50);
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < Y; ++j)
            for (int i = 0; i < X; ++i) {
                const long q = i + sy * j + sz * k;
                const double *c = cvof;
                double cf = c[q];
                if (nfilter == 1 && i > 0 && i < X - 1 && j > 0 && j < Y - 1 && k > 0 && k < Z - 1) {
                    const double s2 = c[q - 1] + c[q - sy] + c[q - sz] + c[q + 1] + c[q + sy] + c[q + sz];
                    const double s3 = c[q + 1 + sy] + c[q + 1 - sy] + c[q - 1 + sy] + c[q - 1 - sy] +
                                      c[q + 1 + sz] + c[q + 1 - sz] + c[q - 1 + sz] + c[q - 1 - sz] +
                                      c[q + sy + sz] + c[q + sy - sz] + c[q - sy + sz] + c[q - sy - sz];
                    const double s4 = c[q + 1 + sy + sz] + c[q + 1 + sy - sz] + c[q + 1 - sy + sz] +
                                      c[q + 1 - sy - sz] + c[q - 1 + sy + sz] + c[q - 1 + sy - sz] +
                                      c[q - 1 - sy + sz] + c[q - 1 - sy - sz];
                    cf = b1 * c[q] + b2 * s2 + b3 * s3 + b4 * s4;
                }
                field[q] = mean_flag == 101 ? cf * (a2 - a1) + a1 : 1.0 / (cf * (inva2 - inva1) + inva1);
            }
    const int ext[3] = {X, Y, Z}, n[3] = {L.n[0], L.n[1], L.n[2]};
    for (int d = 0; d < 3; ++d) {
        if (!L.per[d]) continue;
        for (int k = 0; k < Z; ++k)
            for (int j = 0; j < Y; ++j)
                for (int i = 0; i < X; ++i) {
                    int idx[3] = {i, j, k};
                    const int a = idx[d];
                    if (a >= L.ng && a < L.ng + n[d]) continue;
                    idx[d] = a < L.ng ? a + n[d] : a - n[d];
                    field[i + sy * j + sz * k] = field[idx[0] + sy * idx[1] + sz * idx[2]];
                }
        (void)ext;
    }
    return 0;
}

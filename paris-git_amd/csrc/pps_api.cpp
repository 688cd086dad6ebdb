// pps_api.cpp -- the C ABI of libpps: handle lifecycle (init_MG /
// finalize_MG, MGsolver.f90:34-82), NewSolver / NewSolverMG
// (newsolver.f90:34-57, MGsolver.f90:346-405) on host or device arrays,
// and the adjoint entry point.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pps_internal.h"

using namespace pps;

extern "C" const char *pps_status_string(int st) {
    switch (st) {
    case PPS_OK: return "ok";
    case PPS_ERR_NAN_OR_TINY_DIAG: return "A is NaN or A7 is too close to 0";
    case PPS_ERR_VOLUME_SOURCE: return "large volume source";
    case PPS_ERR_NOT_IN_KERNEL: return "constant p not in kernel";
    case PPS_ERR_BAD_SIZE: return "multigrid requires 2^n cells per direction";
    case PPS_ERR_DIVERGED: return "pressure solver diverged";
    case PPS_ERR_NOT_CONVERGED: return "solver reached maxit";
    case PPS_ERR_NONSYMMETRIC: return "operator is not symmetric";
    case PPS_ERR_BAD_ARG: return "bad argument";
    case PPS_ERR_HIP: return "HIP runtime error";
    case PPS_ERR_NO_DEVICE: return "no HIP device";
    case PPS_ERR_UNSUPPORTED: return "unsupported by this entry point (method, option or boundary condition)";
    case PPS_ERR_SCRATCH: return "grid too large for the no-allocation solve's fixed scratch array";
    default: return "unknown status";
    }
}

extern "C" int pps_default_options(pps_options *o) {
    if (!o) return PPS_ERR_BAD_ARG;
    o->method = PPS_MG_FP64;
    o->outer = PPS_OUTER_CG;
    o->nrelax = 3;
    o->sym_mg = 0;
    o->nu_pre = 2;
    o->nu_post = 2;
    o->check_setup = 0;
    o->multigrid = 1;
    o->smoother = PPS_SMOOTH_FUSED;
    o->sweep_rows = 0;
    o->sweep_planes = 0;
    o->tail_cells = 0;
    o->transfer = PPS_TRANSFER_CONSTANT;
    return PPS_OK;
}

static int valid_grid(const pps_grid *g) {
    if (!g || g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->ng < 1) return PPS_ERR_BAD_ARG;
    for (int d = 0; d < 3; ++d) {
        const bool p0 = g->bdry_cond[d] == 1, p1 = g->bdry_cond[d + 3] == 1;
        if (p0 != p1) return PPS_ERR_BAD_ARG;   // paris.f90:2980-2986
    }
    return PPS_OK;
}

static int valid_opt(const pps_options *o) {
    if (!o) return PPS_ERR_BAD_ARG;
    if (o->method < PPS_MG_REFERENCE || o->method > PPS_SOR) return PPS_ERR_BAD_ARG;
    if (o->outer != PPS_OUTER_CG && o->outer != PPS_OUTER_REFINE) return PPS_ERR_BAD_ARG;
    if (o->nrelax < 0 || o->nu_pre < 0 || o->nu_post < 0) return PPS_ERR_BAD_ARG;
    if (o->smoother != PPS_SMOOTH_FUSED && o->smoother != PPS_SMOOTH_COLOUR_PASSES) return PPS_ERR_BAD_ARG;
    if (o->sweep_rows < 0 || o->sweep_rows > 1024 || o->sweep_planes < 0) return PPS_ERR_BAD_ARG;
    if (o->transfer != PPS_TRANSFER_CONSTANT && o->transfer != PPS_TRANSFER_TRILINEAR) return PPS_ERR_BAD_ARG;
    return PPS_OK;
}

extern "C" int pps_init(const pps_grid *grid, const pps_options *opt, pps_handle **out) {
    if (!out) return PPS_ERR_BAD_ARG;
    *out = nullptr;
    PPS_TRY(valid_grid(grid));
    pps_options o;
    if (opt) o = *opt; else pps_default_options(&o);
    PPS_TRY(valid_opt(&o));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PPS_ERR_NO_DEVICE;
    pps_handle *h = new pps_handle();
    h->grid = *grid;
    h->opt = o;
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return PPS_ERR_NO_DEVICE; }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return PPS_ERR_HIP;
    }
    for (auto &e : h->ev) {
        if (hipEventCreate(&e) != hipSuccess) { pps_finalize(h); return PPS_ERR_HIP; }
    }
    if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess) {
        h->side = nullptr;
        pps_finalize(h);
        return PPS_ERR_HIP;
    }
    for (auto &e : h->sev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { pps_finalize(h); return PPS_ERR_HIP; }
    }
    // scalar read-back buffer, and the device's scratch array address (the
    // no-allocation solve then allocates nothing of its own)
    if (hipHostMalloc((void **)&h->pinned, 16 * sizeof(double), 0) != hipSuccess) {
        h->pinned = nullptr;
        pps_finalize(h);
        return PPS_ERR_HIP;
    }
    if (hipHostMalloc((void **)&h->mbox, sizeof(pps::Mailbox), hipHostMallocCoherent) != hipSuccess) {
        h->mbox = nullptr;
        pps_finalize(h);
        return PPS_ERR_HIP;
    }
    h->mbox->seq = 0;
    const int rc = fast_scratch_setup(h->device);
    if (rc != PPS_OK) { pps_finalize(h); return rc; }
    *out = h;
    return PPS_OK;
}

extern "C" int pps_finalize(pps_handle *h) {
    if (!h) return PPS_ERR_BAD_ARG;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->side) (void)hipStreamSynchronize(h->side);
    ref_destroy(h->ref);
    fast_destroy(h->fast);
    h->hA.release();
    h->hp.release();
    h->hmask.release();
    h->hpgas.release();
    h->spacing.release();
    h->widths.release();
    h->asplit.release();
    h->red.release();
    if (h->pinned) (void)hipHostFree(h->pinned);
    if (h->mbox) (void)hipHostFree(h->mbox);
    for (auto &e : h->prof.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->sev)
        if (e) (void)hipEventDestroy(e);
    if (h->side) {
        (void)hipStreamSynchronize(h->side);
        (void)hipStreamDestroy(h->side);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PPS_OK;
}

extern "C" int pps_set_options(pps_handle *h, const pps_options *opt) {
    if (!h) return PPS_ERR_BAD_ARG;
    PPS_TRY(valid_opt(opt));
    h->opt = *opt;
    // the next solve or adjoint builds its operator and hierarchy again
    // (method, transfers or smoothing may have changed)
    fast_invalidate(h->fast);
    return PPS_OK;
}

static int solve_with(pps_handle *h, int method, const double *dA, double *dp, double maxError,
                      double beta, int maxit, int norm, int *it, double *tres2, hipStream_t s,
                      const ProjK *proj = nullptr) {
    if (!dA || !dp || !it || !tres2 || norm < 1) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    const Dims D = dims_of(h->grid);
    if (h->opt.check_setup) PPS_TRY(check_setup(h, D, dA, s));
    h->setup_ms = 0.0;
    if (maxit < 1) {
        // the reference loop does not run: it = maxit+1, tres2 = the
        // initial sum |r|^norm / N (MGsolver.f90:372-375)
        double res = 0.0;
        PPS_TRY(compute_residual_norm(h, D, dA, dp, norm, &res, s));
        *tres2 = res / (double)D.cells();
        *it = maxit + 1;
        h->hist.clear();
        return PPS_ERR_NOT_CONVERGED;
    }
    int rc;
    if (method == PPS_MG_FP64 || method == PPS_MG_MIXED) {
        const int saved = h->opt.method;
        h->opt.method = method;
        rc = fast_solve(h, dA, dp, maxError, maxit, norm, it, tres2, s, false, proj);
        h->opt.method = saved;
    } else {
        PPS_HIP(hipEventRecord(h->ev[0], s));
        if (method == PPS_MG_REFERENCE) rc = ref_solve(h, dA, dp, maxError, beta, maxit, norm, it, tres2, s);
        else rc = sor_solve(h, dA, dp, maxError, beta, maxit, norm, it, tres2, s);
        PPS_HIP(hipEventRecord(h->ev[2], s));
        PPS_HIP(hipEventSynchronize(h->ev[2]));
        float ms = 0.f;
        PPS_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[2]));
        h->solve_ms = ms;
    }
    if (rc != PPS_OK) return rc;
    return *it > maxit ? PPS_ERR_NOT_CONVERGED : PPS_OK;
}

extern "C" int pps_solve_device(pps_handle *h, const double *dA, double *dp, double maxError,
                                double beta, int maxit, int norm, int *it, double *tres2,
                                void *stream) {
    if (!h) return PPS_ERR_BAD_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: the default stream
    return solve_with(h, h->opt.method, dA, dp, maxError, beta, maxit, norm, it, tres2, s);
}

// NewSolver followed by project_velocity (paris.f90:499, 898-908).  The fast
// methods with norm 2 apply the correction inside their last residual
// pass; the others solve, copy and project.  The outputs are separate from
// u, v, w so that a restart of the fast solve can rewrite them.
extern "C" int pps_solve_project_device(pps_handle *h, const double *dA, double *dp,
                                        double maxError, double beta, int maxit, int norm, int *it,
                                        double *tres2, const double *u, const double *v,
                                        const double *w, const double *umask, const double *vmask,
                                        const double *wmask, const double *rho, double dt,
                                        double *u_out, double *v_out, double *w_out, void *stream) {
    if (!h || !u || !v || !w || !umask || !vmask || !wmask || !rho || !u_out || !v_out || !w_out)
        return PPS_ERR_BAD_ARG;
    if (u_out == u || v_out == v || w_out == w) return PPS_ERR_BAD_ARG;
    // the correction uses the grid's uniform spacings; a stretched grid
    // (cell widths set) solves, then calls pps_project_velocity_coords_device
    if (h->have_widths) return PPS_ERR_UNSUPPORTED;
    (void)hipSetDevice(h->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Dims D = dims_of(h->grid);
    const int method = h->opt.method;
    if ((method == PPS_MG_FP64 || method == PPS_MG_MIXED) && norm == 2 && maxit >= 1) {
        ProjK P;
        const double *dh[3];
        PPS_TRY(handle_spacing(h, dh, s));
        P.u = u; P.v = v; P.w = w;
        P.um = umask; P.vm = vmask; P.wm = wmask; P.rho = rho;
        P.dxh = dh[0]; P.dyh = dh[1]; P.dzh = dh[2];
        P.uo = u_out; P.vo = v_out; P.wo = w_out;
        P.dt = dt;
        P.lx = h->grid.bdry_cond[0] != 1;
        P.ly = h->grid.bdry_cond[1] != 1;
        P.lz = h->grid.bdry_cond[2] != 1;
        const double *const src[3] = {u, v, w};
        double *const dst[3] = {u_out, v_out, w_out};
        PPS_TRY(copy_ghosts(D, src, dst, s));
        return solve_with(h, method, dA, dp, maxError, beta, maxit, norm, it, tres2, s, &P);
    }
    const int rc = solve_with(h, method, dA, dp, maxError, beta, maxit, norm, it, tres2, s);
    if (rc != PPS_OK && rc != PPS_ERR_NOT_CONVERGED) return rc;
    const size_t bytes = sizeof(double) * (size_t)D.padded();
    PPS_HIP(hipMemcpyAsync(u_out, u, bytes, hipMemcpyDeviceToDevice, s));
    PPS_HIP(hipMemcpyAsync(v_out, v, bytes, hipMemcpyDeviceToDevice, s));
    PPS_HIP(hipMemcpyAsync(w_out, w, bytes, hipMemcpyDeviceToDevice, s));
    PPS_TRY(pps_project_velocity_device(&h->grid, u_out, v_out, w_out, umask, vmask, wmask, rho,
                                        dp, dt, stream));
    return rc;
}

// Cell widths for the fast methods: the rows are weighted by the cell
// volume (normalised), which symmetrises SetupPoisson's system on a
// stretched grid (xform, read_x: paris.f90:3056-3102).  NULL clears them.
extern "C" int pps_set_cell_widths(pps_handle *h, const double *dx, const double *dy,
                                   const double *dz) {
    if (!h) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    if (!dx && !dy && !dz) {
        h->have_widths = 0;
        h->wmin = 1.0;
        fast_invalidate(h->fast);
        return PPS_OK;
    }
    if (!dx || !dy || !dz) return PPS_ERR_BAD_ARG;
    const int N[3] = {h->grid.nx, h->grid.ny, h->grid.nz};
    const double *src[3] = {dx, dy, dz};
    std::vector<double> w((size_t)N[0] + N[1] + N[2]);
    double wmin = 1.0;
    size_t o = 0;
    for (int a = 0; a < 3; ++a) {
        double mx = 0.0, mn = 0.0;
        for (int q = 0; q < N[a]; ++q) {
            const double v = src[a][q];
            if (!(v > 0.0) || !std::isfinite(v)) return PPS_ERR_BAD_ARG;
            mx = q == 0 ? v : std::max(mx, v);
            mn = q == 0 ? v : std::min(mn, v);
        }
        for (int q = 0; q < N[a]; ++q) w[o + q] = src[a][q] / mx;
        wmin *= mn / mx;
        o += (size_t)N[a];
    }
    PPS_TRY(h->widths.ensure(sizeof(double) * w.size()));
    PPS_HIP(hipMemcpy(h->widths.ptr, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    h->have_widths = 1;
    h->wmin = wmin;
    fast_invalidate(h->fast);   // rebuild the operator on the next solve
    return PPS_OK;
}

// host-array staging shared by NewSolver / NewSolverMG
static int host_solve(pps_handle *h, int method, const double *A, double *p, double maxError,
                      double beta, int maxit, int *it, int *ierr, int norm, double *tres2) {
    if (!h || !A || !p || !it) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    const Dims D = dims_of(h->grid);
    const size_t bA = sizeof(double) * 8 * (size_t)D.cells();
    const size_t bp = sizeof(double) * (size_t)D.padded();
    PPS_TRY(h->hA.ensure(bA));
    PPS_TRY(h->hp.ensure(bp));
    hipStream_t s = h->stream;
    PPS_HIP(hipMemcpyAsync(h->hA.ptr, A, bA, hipMemcpyHostToDevice, s));
    PPS_HIP(hipMemcpyAsync(h->hp.ptr, p, bp, hipMemcpyHostToDevice, s));
    double t2 = 0.0;
    int rc = solve_with(h, method, h->hA.as<double>(), h->hp.as<double>(), maxError, beta, maxit,
                        norm, it, &t2, s);
    if (rc == PPS_OK || rc == PPS_ERR_NOT_CONVERGED) {
        PPS_HIP(hipMemcpyAsync(p, h->hp.ptr, bp, hipMemcpyDeviceToHost, s));
        PPS_HIP(hipStreamSynchronize(s));
    }
    if (tres2) *tres2 = t2;
    // the reference reports the last MPI code in ierr; a failure aborts it
    if (ierr) *ierr = (rc == PPS_OK || rc == PPS_ERR_NOT_CONVERGED) ? 0 : rc;
    return rc;
}

extern "C" int pps_new_solver(pps_handle *h, const double *A, double *p, double maxError,
                              double beta, int maxit, int *it, int *ierr, int norm) {
    if (!h) return PPS_ERR_BAD_ARG;
    const int method = h->opt.multigrid ? h->opt.method : PPS_SOR;
    return host_solve(h, method, A, p, maxError, beta, maxit, it, ierr, norm, nullptr);
}

extern "C" int pps_new_solver_mg(pps_handle *h, const double *A, double *p, double maxError,
                                 double beta, int maxit, int *it, int *ierr, int norm,
                                 double *tres2) {
    if (!h) return PPS_ERR_BAD_ARG;
    const int method = h->opt.method == PPS_SOR ? PPS_MG_REFERENCE : h->opt.method;
    return host_solve(h, method, A, p, maxError, beta, maxit, it, ierr, norm, tres2);
}

// FreeSolver on device arrays (freesurface.f90:700-808)
static int free_with(pps_handle *h, const double *dA, double *dp, const int *dmask, const double *dpgas,
                     int solver_flag, double maxError, double beta, int maxit, int *it, double *tres2,
                     hipStream_t s) {
    if (!dA || !dp || !dmask || !it || !tres2) return PPS_ERR_BAD_ARG;
    if (solver_flag == 1 && !dpgas) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    PPS_HIP(hipEventRecord(h->ev[0], s));
    const int rc = free_solve(h, dA, dp, dmask, dpgas, solver_flag, maxError, beta, maxit, it, tres2, s);
    PPS_HIP(hipEventRecord(h->ev[2], s));
    PPS_HIP(hipEventSynchronize(h->ev[2]));
    float ms = 0.f;
    PPS_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[2]));
    h->solve_ms = ms;
    h->setup_ms = 0.0;
    if (rc != PPS_OK) return rc;
    return *it > maxit ? PPS_ERR_NOT_CONVERGED : PPS_OK;
}

extern "C" int pps_free_solver_device(pps_handle *h, const double *dA, double *dp, const int *dpcmask,
                                      const double *dP_gas, int solver_flag, double maxError,
                                      double beta, int maxit, int *it, double *tres2, void *stream) {
    if (!h) return PPS_ERR_BAD_ARG;
    return free_with(h, dA, dp, dpcmask, dP_gas, solver_flag, maxError, beta, maxit, it, tres2,
                     static_cast<hipStream_t>(stream));
}

extern "C" int pps_free_solver(pps_handle *h, const double *A, double *p, const int *pcmask,
                               const double *P_gas, int solver_flag, double maxError, double beta,
                               int maxit, int *it, int *ierr, double *tres2) {
    if (!h || !A || !p || !pcmask || !it) return PPS_ERR_BAD_ARG;
    if (solver_flag == 1 && !P_gas) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    const Dims D = dims_of(h->grid);
    const size_t bA = sizeof(double) * 8 * (size_t)D.cells();
    const size_t bp = sizeof(double) * (size_t)D.padded();
    const size_t bm = sizeof(int) * (size_t)D.padded();
    PPS_TRY(h->hA.ensure(bA));
    PPS_TRY(h->hp.ensure(bp));
    PPS_TRY(h->hmask.ensure(bm));
    hipStream_t s = h->stream;
    PPS_HIP(hipMemcpyAsync(h->hA.ptr, A, bA, hipMemcpyHostToDevice, s));
    PPS_HIP(hipMemcpyAsync(h->hp.ptr, p, bp, hipMemcpyHostToDevice, s));
    PPS_HIP(hipMemcpyAsync(h->hmask.ptr, pcmask, bm, hipMemcpyHostToDevice, s));
    const double *dpg = nullptr;
    if (P_gas) {
        PPS_TRY(h->hpgas.ensure(bp));
        PPS_HIP(hipMemcpyAsync(h->hpgas.ptr, P_gas, bp, hipMemcpyHostToDevice, s));
        dpg = h->hpgas.as<double>();
    }
    double t2 = 0.0;
    const int rc = free_with(h, h->hA.as<double>(), h->hp.as<double>(), h->hmask.as<int>(), dpg,
                             solver_flag, maxError, beta, maxit, it, &t2, s);
    if (rc == PPS_OK || rc == PPS_ERR_NOT_CONVERGED) {
        PPS_HIP(hipMemcpyAsync(p, h->hp.ptr, bp, hipMemcpyDeviceToHost, s));
        PPS_HIP(hipStreamSynchronize(s));
    }
    if (tres2) *tres2 = t2;
    if (ierr) *ierr = (rc == PPS_OK || rc == PPS_ERR_NOT_CONVERGED) ? 0 : rc;
    return rc;
}

// poi_solve (modules.f90:2411-2593) on device arrays: the handle's fast
// method (PPS_MG_FP64 unless it is PPS_MG_MIXED) to ||r||_2/||b||_2 <
// maxError; *num_iterations = iterations performed (at most maxit).
static int poi_with(pps_handle *h, const double *dA, double *dp, double maxError, int maxit,
                    int *num_iterations, int solver_type, hipStream_t s) {
    if (!dA || !dp || !num_iterations) return PPS_ERR_BAD_ARG;
    if (solver_type < 1 || solver_type > 4) return PPS_ERR_BAD_ARG;   // SMG, PFMG, GMRES, BiCGSTAB
    (void)hipSetDevice(h->device);
    *num_iterations = 0;
    if (maxit < 1) return PPS_ERR_NOT_CONVERGED;
    const int saved = h->opt.method;
    h->opt.method = saved == PPS_MG_MIXED ? PPS_MG_MIXED : PPS_MG_FP64;
    int it = 0;
    double t2 = 0.0;
    const int rc = fast_solve(h, dA, dp, maxError, maxit, 2, &it, &t2, s, true);
    h->opt.method = saved;
    if (rc != PPS_OK) return rc;
    *num_iterations = it > maxit ? maxit : it;
    return it > maxit ? PPS_ERR_NOT_CONVERGED : PPS_OK;
}

extern "C" int pps_poi_solve_device(pps_handle *h, const double *dA, double *dp, double maxError,
                                    int maxit, int *num_iterations, int solver_type, void *stream) {
    if (!h) return PPS_ERR_BAD_ARG;
    return poi_with(h, dA, dp, maxError, maxit, num_iterations, solver_type,
                    static_cast<hipStream_t>(stream));
}

extern "C" int pps_poi_solve(pps_handle *h, const double *A, double *p, double maxError, int maxit,
                             int *num_iterations, int solver_type) {
    if (!h || !A || !p || !num_iterations) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    const Dims D = dims_of(h->grid);
    const size_t bA = sizeof(double) * 8 * (size_t)D.cells();
    const size_t bp = sizeof(double) * (size_t)D.padded();
    PPS_TRY(h->hA.ensure(bA));
    PPS_TRY(h->hp.ensure(bp));
    hipStream_t s = h->stream;
    PPS_HIP(hipMemcpyAsync(h->hA.ptr, A, bA, hipMemcpyHostToDevice, s));
    PPS_HIP(hipMemcpyAsync(h->hp.ptr, p, bp, hipMemcpyHostToDevice, s));
    const int rc = poi_with(h, h->hA.as<double>(), h->hp.as<double>(), maxError, maxit, num_iterations,
                            solver_type, s);
    if (rc == PPS_OK || rc == PPS_ERR_NOT_CONVERGED) {
        PPS_HIP(hipMemcpyAsync(p, h->hp.ptr, bp, hipMemcpyDeviceToHost, s));
        PPS_HIP(hipStreamSynchronize(s));
    }
    return rc;
}

extern "C" int pps_last_history(pps_handle *h, double *hist, int cap) {
    if (!h || (!hist && cap > 0)) return -PPS_ERR_BAD_ARG;
    int n = (int)h->hist.size();
    if (n > cap) n = cap;
    for (int q = 0; q < n; ++q) hist[q] = h->hist[q];
    return n;
}

extern "C" int pps_last_timing(pps_handle *h, double *solve_ms, double *setup_ms) {
    if (!h) return PPS_ERR_BAD_ARG;
    if (solve_ms) *solve_ms = h->solve_ms;
    if (setup_ms) *setup_ms = h->setup_ms;
    return PPS_OK;
}

extern "C" int pps_profile_enable(pps_handle *h, int on) {
    if (!h) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    if (on && h->prof.ev.empty()) {
        h->prof.ev.resize(1024, nullptr);
        for (auto &e : h->prof.ev) PPS_HIP(hipEventCreate(&e));
    }
    h->prof.on = on ? 1 : 0;
    h->prof.every = on > 1 ? on : 1;
    h->prof.seen = 0;
    return PPS_OK;
}

extern "C" int pps_profile_read(pps_handle *h, double *kernel_ms, long long *launches,
                                double *bytes) {
    if (!h) return PPS_ERR_BAD_ARG;
    if (kernel_ms) *kernel_ms = h->prof.ms;
    if (launches) *launches = h->prof.launches;
    if (bytes) *bytes = h->prof.bytes;
    h->prof.ms = 0.0;
    h->prof.launches = 0;
    h->prof.bytes = 0.0;
    return PPS_OK;
}

// The solve of the system that pps_faces_to_A_device would form, with its
// rows taken from the face values inside the fast solve's setup pass.
extern "C" int pps_solve_faces_device(pps_handle *h, const double *cface, const double *wx,
                                      const double *wy, const double *wz, const double *b, double *dp,
                                      double maxError, int maxit, int *it, double *tres2,
                                      void *stream) {
    if (!h || !cface || !wx || !wy || !wz || !b || !dp || !it || !tres2 || maxit < 1)
        return PPS_ERR_BAD_ARG;
    const int method = h->opt.method;
    // the face-value system is symmetric as it stands: cell widths (the row
    // weights of SetupPoisson's stretched system) do not apply to it
    if ((method != PPS_MG_FP64 && method != PPS_MG_MIXED) || h->opt.check_setup || h->have_widths)
        return PPS_ERR_UNSUPPORTED;
    (void)hipSetDevice(h->device);
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: the default stream
    h->setup_ms = 0.0;
    const FacesK F{cface, wx, wy, wz, b};
    const int rc = fast_solve(h, nullptr, dp, maxError, maxit, 2, it, tres2, s, false, nullptr, nullptr, &F);
    if (rc != PPS_OK) return rc;
    return *it > maxit ? PPS_ERR_NOT_CONVERGED : PPS_OK;
}

// The adjoint of that system on the hierarchy of the last solve from the
// same face values (set up again from them otherwise).
extern "C" int pps_adjoint_faces_device(pps_handle *h, const double *cface, const double *wx,
                                        const double *wy, const double *wz, const double *dg,
                                        double *dlam, double rtol, int maxit, int *it, void *stream) {
    if (!h || !cface || !wx || !wy || !wz || !dg || !dlam || !it || !(rtol > 0.0) || maxit < 1)
        return PPS_ERR_BAD_ARG;
    if (h->have_widths) return PPS_ERR_UNSUPPORTED;
    (void)hipSetDevice(h->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int saved = h->opt.method;
    if (saved != PPS_MG_MIXED) h->opt.method = PPS_MG_FP64;
    const FacesK F{cface, wx, wy, wz, nullptr};
    const int rc = fast_adjoint(h, nullptr, dg, dlam, rtol, maxit, it, s, &F);
    h->opt.method = saved;
    return rc;
}

extern "C" int pps_adjoint_device(pps_handle *h, const double *dA, const double *dg, double *dlam,
                                  double rtol, int maxit, int *it, void *stream) {
    if (!h || !dA || !dg || !dlam || !it || !(rtol > 0.0) || maxit < 1) return PPS_ERR_BAD_ARG;
    (void)hipSetDevice(h->device);
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: the default stream
    // the adjoint runs on the fast hierarchy; the reference-faithful and
    // plain methods use its FP64 form
    const int saved = h->opt.method;
    if (saved != PPS_MG_MIXED) h->opt.method = PPS_MG_FP64;
    const int rc = fast_adjoint(h, dA, dg, dlam, rtol, maxit, it, s);
    h->opt.method = saved;
    return rc;
}

// pps_internal.h -- shared declarations of libpps (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/pps.h"

#define PPS_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            std::fprintf(stderr, "libpps: %s failed: %s (%s:%d)\n", #call,   \
                         hipGetErrorString(e_), __FILE__, __LINE__);         \
            return PPS_ERR_HIP;                                              \
        }                                                                    \
    } while (0)

#define PPS_TRY(expr)                 \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != PPS_OK) return rc_; \
    } while (0)

namespace pps {

// Geometry of one level: interior dims, ghost layers, padded extents,
// periodic flags.  Vectors "with ghosts" use the padded extents, x fastest.
struct Dims {
    int nx, ny, nz, ng;
    int per[3];
    long long cells() const { return (long long)nx * ny * nz; }
    int X() const { return nx + 2 * ng; }
    int Y() const { return ny + 2 * ng; }
    int Z() const { return nz + 2 * ng; }
    long long padded() const { return (long long)X() * Y() * Z(); }
};

inline Dims dims_of(const pps_grid &g) {
    Dims d;
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz; d.ng = g.ng;
    d.per[0] = g.bdry_cond[0] == 1;
    d.per[1] = g.bdry_cond[1] == 1;
    d.per[2] = g.bdry_cond[2] == 1;
    return d;
}

// Bump allocator over a fixed device region (the no-allocation solve's
// scratch array): hands out 256-byte aligned pieces, never frees.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t n) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at > cap || n > cap - at) return nullptr;
        used = at + n;
        return base + at;
    }
};

// Device buffer owned by the handle (freed with the handle), or a piece of
// an Arena (borrowed: never freed, never reallocated).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    int ensure(size_t n) {
        if (n <= bytes) return PPS_OK;
        if (borrowed) return PPS_ERR_BAD_SIZE;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr; bytes = 0;
        if (hipMalloc(&ptr, n) != hipSuccess) return PPS_ERR_HIP;
        bytes = n;
        return PPS_OK;
    }
    // at least n bytes: from the arena when one is given, else ensure(n)
    int ensure(size_t n, Arena *a) {
        if (!a) return ensure(n);
        if (n <= bytes) return PPS_OK;
        if (ptr && !borrowed) return PPS_ERR_BAD_ARG;   // never mix the two kinds
        void *p = a->take(n);
        if (!p) return PPS_ERR_SCRATCH;
        ptr = p; bytes = n; borrowed = true;
        return PPS_OK;
    }
    void release() {
        if (ptr && !borrowed) (void)hipFree(ptr);
        ptr = nullptr; bytes = 0; borrowed = false;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Reductions: every reducing kernel writes one partial per block; a tiny
// single-block kernel sums the partials in a fixed order so results do
// not depend on scheduling.
constexpr int kRedBlocks = 65536;  // partial slots per reduction (one per block)
constexpr int kRedThreads = 256;

struct RefState;                   // reference-faithful MG (pps_ref.hip)
struct FastState;                  // fast MG (pps_fast.hip)

// Event timing of the fine-level smoother launches of a solve: an event
// pair brackets each launch on the solver's stream; elapsed times are
// collected after the solve has synchronised.
struct Prof {
    int on = 0;
    int every = 1;          // time every `every`-th eligible launch
    long long seen = 0;     // eligible launches since enabled
    std::vector<hipEvent_t> ev;
    std::vector<double> pend_bytes;
    int used = 0;
    double ms = 0.0, bytes = 0.0;
    long long launches = 0;
    // returns the event to record before a launch of `b` algorithmic bytes,
    // or nullptr when not profiling / out of events
    hipEvent_t *begin(double b) {
        if (!on || used + 2 > (int)ev.size()) return nullptr;
        if (seen++ % every) return nullptr;
        pend_bytes.push_back(b);
        hipEvent_t *e = &ev[used];
        used += 2;
        return e;
    }
    int collect() {
        for (int q = 0; q + 1 < used; q += 2) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[q], ev[q + 1]) != hipSuccess) return PPS_ERR_HIP;
            ms += t;
            bytes += pend_bytes[q / 2];
            ++launches;
        }
        used = 0;
        pend_bytes.clear();
        return PPS_OK;
    }
};

// Scalars a reduction kernel writes straight into coherent host memory
// (rr = r.r, bb = b.b, the setup flags), then seq with a system-scope
// release, so the host can spin on seq instead of a stream synchronisation.
struct Mailbox {
    double rr, bb, flags, pad;
    unsigned long long seq;
};

}  // namespace pps

struct pps_handle {
    pps_grid grid;
    pps_options opt;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // second stream (default priority, non-blocking) on which the fast
    // solve builds its coarse hierarchy while the first fine sweeps run,
    // and two ordering events (no timing) between it and the solve stream
    hipStream_t side = nullptr;
    hipEvent_t sev[2] = {nullptr, nullptr};
    pps::RefState *ref = nullptr;
    pps::FastState *fast = nullptr;
    std::vector<double> hist;
    double solve_ms = 0.0, setup_ms = 0.0;
    // scratch device buffers shared by drivers (A/p staging for host calls)
    pps::DevBuf hA, hp;
    pps::DevBuf hmask, hpgas;      // pcmask / P_gas staging for host FreeSolver calls
    pps::DevBuf spacing;           // dx, dxh, x per axis (handle_spacing)
    pps::DevBuf asplit;            // colour-split A of the SOR sweeps (8n doubles)
    // row weights of the fast methods (pps_set_cell_widths): wx | wy | wz on
    // the device, normalised to a maximum of 1, and min wx * min wy * min wz
    pps::DevBuf widths;
    int have_widths = 0;
    double wmin = 1.0;
    pps::DevBuf red;               // reduction partials + results
    double *pinned = nullptr;      // 16 pinned host doubles for scalar readback
    pps::Mailbox *mbox = nullptr;  // coherent host memory (pps_init)
    unsigned long long mseq = 0;   // seq of the last mailbox write requested
    pps::Prof prof;                // fine-level smoother timing
};

namespace pps {

// drivers (each returns a pps_status)
int ref_solve(pps_handle *h, const double *dA, double *dp, double maxError,
              double beta, int maxit, int norm, int *it, double *tres2,
              hipStream_t s);
int sor_solve(pps_handle *h, const double *dA, double *dp, double maxError,
              double beta, int maxit, int norm, int *it, double *tres2,
              hipStream_t s);
void ref_destroy(RefState *r);

// project_velocity (paris.f90:898-908) applied inside the fast solve's last
// residual pass: out = in - dt (2 mask / dh) (p_e - p_c) / (rho_e + rho_c) on
// the faces between each cell and its +x / +y / +z neighbour, the last face
// of an axis left as it is unless the axis is periodic (l* = 1).  Padded
// arrays like p; dxh / dyh / dzh indexed like the reference (i + ng).
struct ProjK {
    const double *u, *v, *w, *um, *vm, *wm, *rho, *dxh, *dyh, *dzh;
    double *uo, *vo, *wo;
    double dt;
    int lx, ly, lz;
};

// Face values as the rows of the system (the map of pps_faces_to_A_device):
// cf = three planes of face values (+x, +y, +z face of every cell),
// wx / wy / wz the per-plane weights, b the interior right-hand side (null:
// zero, for an adjoint that only needs the operator).
struct FacesK {
    const double *cf, *wx, *wy, *wz, *b;
};

// rows: assemble the system from the fields instead of reading dA; faces:
// from face values (dA may then be null; norm 2 only)
struct RowK;
int fast_solve(pps_handle *h, const double *dA, double *dp, double maxError,
               int maxit, int norm, int *it, double *tres2, hipStream_t s,
               bool relative = false, const ProjK *proj = nullptr,
               const RowK *rows = nullptr, const FacesK *faces = nullptr);
int fast_adjoint(pps_handle *h, const double *dA, const double *dg,
                 double *dlam, double rtol, int maxit, int *it, hipStream_t s,
                 const FacesK *faces = nullptr);
void fast_destroy(FastState *f);
// forget the built operator (the next solve sets it up again)
void fast_invalidate(FastState *f);
// look up the scratch array of pps_solve_device_noalloc on `device` (the
// current device), once per device
int fast_scratch_setup(int device);

// shared helpers (pps_assemble.hip); results are read back through the
// handle's pinned scalars, so each call synchronises the stream.
int check_setup(pps_handle *h, const Dims &d, const double *A, hipStream_t s);
// sum |r|^norm of r = A8 - A7 p + sum A_nb p_nb over the interior.
int compute_residual_norm(pps_handle *h, const Dims &d, const double *A,
                          const double *p, int norm, double *out_sum,
                          hipStream_t s);
// dxh, dyh, dzh of the handle's grid on the device (built on first use).
int handle_spacing(pps_handle *h, const double *dh[3], hipStream_t s);
// dst[a] = src[a] on the ghost layers of three padded arrays (interiors
// untouched).
int copy_ghosts(const Dims &d, const double *const src[3], double *const dst[3], hipStream_t s);
// Block partials written by other kernels (comp c of block b at
// part[c * kRedBlocks + b]), then their fixed-order totals (comps in
// maxmask take the maximum) into out; synchronises.
double *stats_partials(pps_handle *h);
int stats_finish(pps_handle *h, int nblocks, int ncomp, int maxmask, double *out, hipStream_t s);
// {sum |r|, sum r^2, max |r|} (calcResiduals before normalisation).
int fs_residual_stats(pps_handle *h, const Dims &d, const double *A, const double *p,
                      const int *pcmask, int flag, double out[4], hipStream_t s);
int free_solve(pps_handle *h, const double *dA, double *dp, const int *dmask, const double *dpgas,
               int solver_flag, double maxError, double beta, int maxit, int *it, double *tres2,
               hipStream_t s);
int residual_stats(pps_handle *h, const Dims &d, const double *A,
                   const double *p, double out[3], hipStream_t s);

}  // namespace pps

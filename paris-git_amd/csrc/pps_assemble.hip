// pps_assemble.hip -- both sides of the solve on the GPU:
//   producer: SetPressureBC + SetupPoisson + Poisson_BCs + the solid-cell
//             rule of check_and_debug_Poisson (modules.f90:916-953,
//             2722-2953), check_poisson_setup (newsolver.f90:275-344),
//             right-hand-side mean removal;
//   consumer: calcResiduals (modules.f90:3313-3384), project_velocity
//             (paris.f90:898-908);
//   synthetic droplet / velocity fields for tests and benchmarks;
//   deterministic reductions shared by every driver.
// Assembly and projection keep the reference's operand order with fp
// contraction off, so A and u are bit-identical to the Fortran expressions.
#pragma clang fp contract(off)

#include <cmath>
#include <cstring>
#include <vector>

#include "pps_internal.h"
#include "pps_row.h"

namespace pps {

// ---------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}
// 256-thread block reductions; result valid in thread 0.  `sm` holds 4
// doubles; callers separate consecutive uses with __syncthreads().
__device__ __forceinline__ double block_sum256(double v, double *sm) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    return t;
}
__device__ __forceinline__ double block_max256(double v, double *sm) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    return t;
}

// Sum (or max, bit c of maxmask) of ncomp partial rows in a fixed order.
__global__ void __launch_bounds__(256)
k_final(const double *__restrict__ part, int nblocks, int ncomp, int maxmask,
        double *__restrict__ out) {
    __shared__ double sm[4];
    for (int c = 0; c < ncomp; ++c) {
        const bool mx = (maxmask >> c) & 1;
        double v = mx ? -INFINITY : 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) {
            const double x = part[c * kRedBlocks + b];
            v = mx ? fmax(v, x) : v + x;
        }
        const double t = mx ? block_max256(v, sm) : block_sum256(v, sm);
        if (threadIdx.x == 0) out[c] = t;
        __syncthreads();
    }
}

static int red_blocks(long long n) {
    long long b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (int)b;
}

static int ensure_red(pps_handle *h) {
    PPS_TRY(h->red.ensure(sizeof(double) * (8 * kRedBlocks + 16) + 64));
    if (!h->pinned) PPS_HIP(hipHostMalloc((void **)&h->pinned, 16 * sizeof(double), 0));
    return PPS_OK;
}
static double *red_part(pps_handle *h) { return h->red.as<double>(); }
static double *red_out(pps_handle *h) { return h->red.as<double>() + 8 * kRedBlocks; }
static int *red_flag(pps_handle *h) {
    return reinterpret_cast<int *>(h->red.as<double>() + 8 * kRedBlocks + 16);
}

// finalize + copy `ncomp` results into h->pinned; synchronises.
static int red_finish(pps_handle *h, int nblocks, int ncomp, int maxmask, hipStream_t s) {
    k_final<<<1, 256, 0, s>>>(red_part(h), nblocks, ncomp, maxmask, red_out(h));
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipMemcpyAsync(h->pinned, red_out(h), sizeof(double) * ncomp, hipMemcpyDeviceToHost, s));
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

struct DK {  // kernel-side dims
    int nx, ny, nz, ng, X, Y;
};
static DK dk(const Dims &d) {
    DK k;
    k.nx = d.nx; k.ny = d.ny; k.nz = d.nz; k.ng = d.ng; k.X = d.X(); k.Y = d.Y();
    return k;
}
__device__ __forceinline__ long long padi(const DK &d, int i, int j, int k) {
    return (long long)(i + d.ng) + (long long)d.X * ((long long)(j + d.ng) + (long long)d.Y * (k + d.ng));
}

// residual statistics: comp 0 = sum|r|, 1 = sum r^2, 2 = max|r|,
// 3 = sum |r|^norm.  form 0: A8 - A7 p + A1 pw + ... (MGsolver.f90:97-100);
// form 1: -p A7 + A1 pw + ... + A8 (newsolver.f90:111-114, modules.f90:3367-3370).
__global__ void __launch_bounds__(256)
k_resid_stats(DK d, const double *__restrict__ A, const double *__restrict__ p,
              int norm, int form, double *__restrict__ part) {
    __shared__ double sm[4];
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long sy = d.X, sz = (long long)d.X * d.Y;
    double s1 = 0.0, s2 = 0.0, sm_ = 0.0, sn = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const int i = (int)(c % d.nx);
        const long long t = c / d.nx;
        const int j = (int)(t % d.ny), k = (int)(t / d.ny);
        const long long q = padi(d, i, j, k);
        double r;
        if (form == 0)
            r = A[7 * n + c] - A[6 * n + c] * p[q] + A[c] * p[q - 1] + A[n + c] * p[q + 1] +
                A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                A[5 * n + c] * p[q + sz];
        else
            r = -p[q] * A[6 * n + c] + A[c] * p[q - 1] + A[n + c] * p[q + 1] +
                A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                A[5 * n + c] * p[q + sz] + A[7 * n + c];
        const double a = fabs(r);
        s1 += a;
        s2 += a * a;
        sm_ = fmax(sm_, a);
        if (norm == 1) sn += a;
        else if (norm == 2) sn += a * a;
        else sn += pow(a, (double)norm);
        if (r != r) sm_ = NAN;   // keep NaN visible through fmax
    }
    double t;
    t = block_sum256(s1, sm); if (threadIdx.x == 0) part[0 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_sum256(s2, sm); if (threadIdx.x == 0) part[1 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_max256(sm_, sm); if (threadIdx.x == 0) part[2 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_sum256(sn, sm); if (threadIdx.x == 0) part[3 * kRedBlocks + blockIdx.x] = t;
}

int compute_residual_norm(pps_handle *h, const Dims &d, const double *A,
                          const double *p, int norm, double *out_sum,
                          hipStream_t s) {
    PPS_TRY(ensure_red(h));
    const int nb = red_blocks(d.cells());
    k_resid_stats<<<nb, 256, 0, s>>>(dk(d), A, p, norm, 0, red_part(h));
    PPS_TRY(red_finish(h, nb, 4, 1 << 2, s));
    *out_sum = h->pinned[3];
    return PPS_OK;
}

// FreeSolver's residual statistics over the active cells (pcmask == 0 for
// solver_flag 1, pcmask in {1, 2} for solver_flag 2; freesurface.f90:
// 738-785, the expression of 742-745): comp 0 = sum r^2, 1 = max |r|,
// 2 = active cells, 3 = active cells whose residual or any A slot is NaN
// (catch_divergence_fs, 811-855).  pcmask is padded like p.
__global__ void __launch_bounds__(256)
k_resid_stats_fs(DK d, const double *__restrict__ A, const double *__restrict__ p,
                 const int *__restrict__ pcmask, int flag, double *__restrict__ part) {
    __shared__ double sm[4];
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long sy = d.X, sz = (long long)d.X * d.Y;
    double s2 = 0.0, mx = 0.0, cnt = 0.0, bad = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const int i = (int)(c % d.nx);
        const long long t = c / d.nx;
        const int j = (int)(t % d.ny), k = (int)(t / d.ny);
        const long long q = padi(d, i, j, k);
        const int m = pcmask[q];
        if (!(flag == 1 ? m == 0 : (m == 1 || m == 2))) continue;
        const double r = fabs(-p[q] * A[6 * n + c] + A[c] * p[q - 1] + A[n + c] * p[q + 1] +
                              A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] +
                              A[4 * n + c] * p[q - sz] + A[5 * n + c] * p[q + sz] + A[7 * n + c]);
        bool nan = r != r;
#pragma unroll
        for (int s = 0; s < 8; ++s) nan |= A[s * n + c] != A[s * n + c];
        s2 += r * r;
        mx = fmax(mx, r);
        cnt += 1.0;
        if (nan) bad += 1.0;
    }
    double t;
    t = block_sum256(s2, sm); if (threadIdx.x == 0) part[0 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_max256(mx, sm); if (threadIdx.x == 0) part[1 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_sum256(cnt, sm); if (threadIdx.x == 0) part[2 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_sum256(bad, sm); if (threadIdx.x == 0) part[3 * kRedBlocks + blockIdx.x] = t;
}

double *stats_partials(pps_handle *h) {
    if (ensure_red(h) != PPS_OK) return nullptr;
    return red_part(h);
}
int stats_finish(pps_handle *h, int nblocks, int ncomp, int maxmask, double *out, hipStream_t s) {
    PPS_TRY(red_finish(h, nblocks, ncomp, maxmask, s));
    for (int c = 0; c < ncomp; ++c) out[c] = h->pinned[c];
    return PPS_OK;
}
int fs_residual_stats(pps_handle *h, const Dims &d, const double *A, const double *p,
                      const int *pcmask, int flag, double out[4], hipStream_t s) {
    PPS_TRY(ensure_red(h));
    const int nb = red_blocks(d.cells());
    k_resid_stats_fs<<<nb, 256, 0, s>>>(dk(d), A, p, pcmask, flag, red_part(h));
    PPS_TRY(red_finish(h, nb, 4, 1 << 1, s));
    for (int c = 0; c < 4; ++c) out[c] = h->pinned[c];
    return PPS_OK;
}

int residual_stats(pps_handle *h, const Dims &d, const double *A,
                   const double *p, double out[3], hipStream_t s) {
    PPS_TRY(ensure_red(h));
    const int nb = red_blocks(d.cells());
    k_resid_stats<<<nb, 256, 0, s>>>(dk(d), A, p, 2, 1, red_part(h));
    PPS_TRY(red_finish(h, nb, 4, 1 << 2, s));
    out[0] = h->pinned[0];
    out[1] = h->pinned[1];
    out[2] = h->pinned[2];
    return PPS_OK;
}

// ---------------------------------------------------------------------
// check_poisson_setup (newsolver.f90:275-344).  Solid cells (A7 := 0 in the
// row-sum test): with the face masks, the cells whose six masks are all
// below 1/2, as the reference reads them; without masks (the solvers'
// check_setup option, which has only A), the cells whose six couplings are
// zero, the same cells when A comes from SetupPoisson (an open mask and a
// finite density give a positive coupling).
// comp 0: count of NaN / A7 < 1e-55 cells, 1: sum A8, 2: max |rowsum|
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_check(DK d, const double *__restrict__ A, const double *__restrict__ um,
        const double *__restrict__ vm, const double *__restrict__ wm, double *__restrict__ part) {
    __shared__ double sm[4];
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long sy = d.X, sz = (long long)d.X * d.Y;
    double bad = 0.0, src = 0.0, row = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        double a[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) a[s] = A[s * n + c];
        bool nan = false;
#pragma unroll
        for (int s = 0; s < 8; ++s) nan |= (a[s] != a[s]);
        if (nan || a[6] < 1e-55) bad += 1.0;
        src += a[7];
        bool solid;
        if (um) {
            const int i = (int)(c % d.nx);
            const long long t = c / d.nx;
            const int j = (int)(t % d.ny), k = (int)(t / d.ny);
            const long long q = padi(d, i, j, k);
            solid = um[q - 1] < 0.5 && um[q] < 0.5 && vm[q - sy] < 0.5 && vm[q] < 0.5 &&
                    wm[q - sz] < 0.5 && wm[q] < 0.5;
        } else {
            solid = a[0] == 0.0 && a[1] == 0.0 && a[2] == 0.0 && a[3] == 0.0 && a[4] == 0.0 &&
                    a[5] == 0.0;
        }
        const double a7 = solid ? 0.0 : a[6];
        const double t = (((((a[0] + a[1]) + a[2]) + a[3]) + a[4]) + a[5]) - a7;
        row = fmax(row, fabs(t));
    }
    double t;
    t = block_sum256(bad, sm); if (threadIdx.x == 0) part[0 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_sum256(src, sm); if (threadIdx.x == 0) part[1 * kRedBlocks + blockIdx.x] = t; __syncthreads();
    t = block_max256(row, sm); if (threadIdx.x == 0) part[2 * kRedBlocks + blockIdx.x] = t;
}

static int check_with(pps_handle *h, const Dims &d, const double *A, const double *um,
                      const double *vm, const double *wm, hipStream_t s) {
    PPS_TRY(ensure_red(h));
    const int nb = red_blocks(d.cells());
    k_check<<<nb, 256, 0, s>>>(dk(d), A, um, vm, wm, red_part(h));
    PPS_TRY(red_finish(h, nb, 3, 1 << 2, s));
    if (h->pinned[0] != 0.0) return PPS_ERR_NAN_OR_TINY_DIAG;
    if (h->pinned[1] / (double)d.cells() > 1e-10) return PPS_ERR_VOLUME_SOURCE;
    if (h->pinned[2] > 1e-10) return PPS_ERR_NOT_IN_KERNEL;
    return PPS_OK;
}

int check_setup(pps_handle *h, const Dims &d, const double *A, hipStream_t s) {
    return check_with(h, d, A, nullptr, nullptr, nullptr, s);
}

// ---------------------------------------------------------------------
// right-hand-side mean removal
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_sum_plane(const double *__restrict__ a, long long n, double *__restrict__ part) {
    __shared__ double sm[4];
    double v = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) v += a[c];
    const double t = block_sum256(v, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ void __launch_bounds__(256)
k_sub_mean(double *__restrict__ a, long long n, const double *__restrict__ sum) {
    const double m = sum[0] / (double)n;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) a[c] = a[c] - m;
}

// ---------------------------------------------------------------------
// SetPressureBC on all-ones masks (modules.f90:916-946)
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_masks(DK d, int b0, int b1, int b2, int b3, int b4, int b5, double *__restrict__ um,
        double *__restrict__ vm, double *__restrict__ wm) {
    const long long np = (long long)d.X * d.Y * (d.nz + 2 * d.ng);
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < np; q += (long long)gridDim.x * 256) {
        const int i = (int)(q % d.X) - d.ng;
        const long long t = q / d.X;
        const int j = (int)(t % d.Y) - d.ng, k = (int)(t / d.Y) - d.ng;
        double u = 1.0, v = 1.0, w = 1.0;
        const bool band_jk = j >= -1 && j <= d.ny && k >= -1 && k <= d.nz;
        const bool band_ik = i >= -1 && i <= d.nx && k >= -1 && k <= d.nz;
        const bool band_ij = i >= -1 && i <= d.nx && j >= -1 && j <= d.ny;
        if (band_jk && i == -1 && (b0 == 0 || b0 == 2)) u = 0.0;
        if (band_jk && i == d.nx - 1 && (b3 == 0 || b3 == 2)) u = 0.0;
        if (band_ik && j == -1 && (b1 == 0 || b1 == 2)) v = 0.0;
        if (band_ik && j == d.ny - 1 && (b4 == 0 || b4 == 2)) v = 0.0;
        if (band_ij && k == -1 && (b2 == 0 || b2 == 2)) w = 0.0;
        if (band_ij && k == d.nz - 1 && (b5 == 0 || b5 == 2)) w = 0.0;
        um[q] = u; vm[q] = v; wm[q] = w;
    }
}

// ---------------------------------------------------------------------
// SetupPoisson + Poisson_BCs (codes 3..6) + solid rule
// coordinate arrays are indexed by padded index (Fortran index - 1)
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_setup_poisson(DK d, RowK R, double *__restrict__ A, int *__restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    const long long sy = d.X, sz = (long long)d.X * d.Y;
    double a[8];
    if (setup_row(R, d.nx, d.ny, d.nz, d.ng, sy, sz, i, j, k, padi(d, i, j, k), a)) atomicOr(flag, 1);
#pragma unroll
    for (int s = 0; s < 8; ++s) A[s * n + c] = a[s];
}

// ---------------------------------------------------------------------
// calcResiduals + project_velocity
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_project(DK d, int comp, int last, double *__restrict__ vel, const double *__restrict__ mask,
          const double *__restrict__ rho, const double *__restrict__ p,
          const double *__restrict__ dh, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int ext[3] = {d.nx, d.ny, d.nz};
    int idx[3] = {i, j, k};
    if (i >= d.nx || j >= d.ny || k >= d.nz) return;
    if (idx[comp] >= ext[comp] - last) return;   // ieu / jev / kew
    const long long st[3] = {1, d.X, (long long)d.X * d.Y};
    const long long q = padi(d, i, j, k), e = q + st[comp];
    vel[q] = vel[q] - dt * (2.0 * mask[q] / dh[idx[comp] + d.ng]) * (p[e] - p[q]) / (rho[e] + rho[q]);
}

// ---------------------------------------------------------------------
// synthetic inputs
// ---------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// uniform [0,1) from the idx-th output of a splitmix64 stream seeded `seed`
__device__ __forceinline__ double u01(uint64_t seed, uint64_t idx) {
    const uint64_t z = splitmix64(seed + (idx + 1) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void __launch_bounds__(256)
k_synth_rho(DK d, const double *__restrict__ xc, const double *__restrict__ yc,
            const double *__restrict__ zc, const double *__restrict__ drops, int ndrop,
            double rho1, double rho2, double wh, double Lx, double Ly, double Lz, int px,
            int py, int pz, double *__restrict__ rho) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const double x = xc[i + d.ng], y = yc[j + d.ng], z = zc[k + d.ng];
    const double cut = 12.0 * wh;
    double c = 0.0;
    for (int m = 0; m < ndrop; ++m) {
        double dx = x - drops[4 * m], dy = y - drops[4 * m + 1], dz = z - drops[4 * m + 2];
        if (px) dx -= Lx * rint(dx / Lx);
        if (py) dy -= Ly * rint(dy / Ly);
        if (pz) dz -= Lz * rint(dz / Lz);
        const double R = drops[4 * m + 3];
        const double reach = R + cut;
        if (fabs(dx) > reach || fabs(dy) > reach || fabs(dz) > reach) continue;
        const double s = sqrt(dx * dx + dy * dy + dz * dz) - R;
        if (s > cut) continue;
        c += 0.5 * (1.0 - tanh(s / wh));
    }
    c = fmin(c, 1.0);
    rho[padi(d, i, j, k)] = c * (rho2 - rho1) + rho1;
}

__global__ void __launch_bounds__(256)
k_synth_vel(DK d, uint64_t seed, int comp, int wall, double *__restrict__ vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    const uint64_t n = (uint64_t)d.nx * d.ny * d.nz;
    double val = 2.0 * u01(seed, (uint64_t)comp * n + (uint64_t)c) - 1.0;
    const int idx[3] = {i, j, k}, ext[3] = {d.nx, d.ny, d.nz};
    if (wall && idx[comp] == ext[comp] - 1) val = 0.0;   // the + wall face
    vel[padi(d, i, j, k)] = val;
}

// Fill every ghost of a padded field: periodic directions wrap; wall
// directions copy the nearest interior value (mode 0) or write 0 (mode 1).
// Directions are processed one at a time (x, then y, then z) so edges and
// corners take the composed rule.
__global__ void __launch_bounds__(256)
k_fill_ghosts(DK d, int dir, int periodic, int mode, double *__restrict__ f) {
    const long long np = (long long)d.X * d.Y * (d.nz + 2 * d.ng);
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < np; q += (long long)gridDim.x * 256) {
        int idx[3];
        idx[0] = (int)(q % d.X) - d.ng;
        const long long t = q / d.X;
        idx[1] = (int)(t % d.Y) - d.ng;
        idx[2] = (int)(t / d.Y) - d.ng;
        const int ext[3] = {d.nx, d.ny, d.nz};
        const int a = idx[dir];
        if (a >= 0 && a < ext[dir]) continue;
        // the other directions: only rows already inside or processed
        bool ok = true;
        for (int e = dir + 1; e < 3; ++e) ok &= idx[e] >= 0 && idx[e] < ext[e];
        if (!ok) continue;
        int src[3] = {idx[0], idx[1], idx[2]};
        if (periodic) src[dir] = ((a % ext[dir]) + ext[dir]) % ext[dir];
        else src[dir] = a < 0 ? 0 : ext[dir] - 1;
        if (!periodic && mode == 1) { f[q] = 0.0; continue; }
        f[q] = f[padi(d, src[0], src[1], src[2])];
    }
}

// linfunc (vofmodules.f90:171-218) on every padded cell: the volume
// fraction, 27-point filtered (weights 8, 4, 2, 1 / 64 of
// vofmodules.f90:133-147) off the outermost ghost layer when nfilter is 1,
// mapped to a1..a2 arithmetically (harm 0) or harmonically (harm 1).  The
// operand order is the reference's, without contraction.
__global__ void __launch_bounds__(256)
k_linfunc(DK d, const double *__restrict__ c, double a1, double a2, int harm, int nfilter,
          double *__restrict__ field) {
#pragma clang fp contract(off)
    const int Z = d.nz + 2 * d.ng;
    const long long sy = d.X, sz = (long long)d.X * d.Y, np = sz * Z;
    const double b1 = 8.0 / 64, b2 = 4.0 / 64, b3 = 2.0 / 64, b4 = 1.0 / 64;
    const double inva1 = 1.0 / (a1 + 1e-50), inva2 = 1.0 / (a2 + 1e-50);
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < np; q += (long long)gridDim.x * 256) {
        const int i = (int)(q % d.X);
        const long long t = q / d.X;
        const int j = (int)(t % d.Y), k = (int)(t / d.Y);
        double cf = c[q];
        if (nfilter == 1 && i > 0 && i < d.X - 1 && j > 0 && j < d.Y - 1 && k > 0 && k < Z - 1) {
            const double s2 = c[q - 1] + c[q - sy] + c[q - sz] + c[q + 1] + c[q + sy] + c[q + sz];
            const double s3 = c[q + 1 + sy] + c[q + 1 - sy] + c[q - 1 + sy] + c[q - 1 - sy] +
                              c[q + 1 + sz] + c[q + 1 - sz] + c[q - 1 + sz] + c[q - 1 - sz] +
                              c[q + sy + sz] + c[q + sy - sz] + c[q - sy + sz] + c[q - sy - sz];
            const double s4 = c[q + 1 + sy + sz] + c[q + 1 + sy - sz] + c[q + 1 - sy + sz] +
                              c[q + 1 - sy - sz] + c[q - 1 + sy + sz] + c[q - 1 + sy - sz] +
                              c[q - 1 - sy + sz] + c[q - 1 - sy - sz];
            cf = b1 * c[q] + b2 * s2 + b3 * s3 + b4 * s4;
        }
        field[q] = harm ? 1.0 / (cf * (inva2 - inva1) + inva1) : cf * (a2 - a1) + a1;
    }
}

// One axis of do_all_ghost (modules.f90:2062-2072) on one rank: the ng
// ghost layers of a periodic axis take the wrapped interior values, over
// the full padded extent of the other two axes (ghost_x/y/z send whole
// planes), so the three launches compose edges and corners as the
// reference's sequence does.  Sources are interior along the axis and
// destinations ghosts, so one launch never reads what it writes.
__global__ void __launch_bounds__(256)
k_wrap_axis(DK d, int dir, double *__restrict__ f) {
    const int Z = d.nz + 2 * d.ng;
    const long long np = (long long)d.X * d.Y * Z;
    const int n[3] = {d.nx, d.ny, d.nz};
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < np; q += (long long)gridDim.x * 256) {
        int idx[3];
        idx[0] = (int)(q % d.X);
        const long long t = q / d.X;
        idx[1] = (int)(t % d.Y);
        idx[2] = (int)(t / d.Y);
        const int a = idx[dir];
        if (a >= d.ng && a < d.ng + n[dir]) continue;
        idx[dir] = a < d.ng ? a + n[dir] : a - n[dir];
        f[q] = f[idx[0] + (long long)d.X * (idx[1] + (long long)d.Y * idx[2])];
    }
}

static unsigned lin_blocks(long long n) {
    long long b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return (unsigned)b;
}
static dim3 cgrid(const Dims &d) {
    return dim3((unsigned)((d.nx + 63) / 64), (unsigned)((d.ny + 3) / 4), (unsigned)d.nz);
}
static const dim3 kBlk(64, 4, 1);

// host-side coordinate tables uploaded for a call
struct Coords {
    DevBuf buf;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double *xc = nullptr, *yc = nullptr, *zc = nullptr;
    int make(const pps_grid &g, hipStream_t s) {
        const int nt[3] = {g.nx + 2 * g.ng, g.ny + 2 * g.ng, g.nz + 2 * g.ng};
        const int N[3] = {g.nx, g.ny, g.nz};
        const double Ls[3] = {g.xlength, g.ylength, g.zlength};
        size_t tot = 0;
        for (int a = 0; a < 3; ++a) tot += 3 * (size_t)nt[a];
        std::vector<double> host(tot);
        size_t off[3];
        size_t o = 0;
        for (int a = 0; a < 3; ++a) {
            off[a] = o;
            std::vector<double> x(nt[a]), xh(nt[a]), dx(nt[a]), dxh(nt[a]);
            pps_coords(N[a], g.ng, Ls[a], x.data(), xh.data(), dx.data(), dxh.data());
            for (int q = 0; q < nt[a]; ++q) {
                host[o + q] = dx[q];
                host[o + nt[a] + q] = dxh[q];
                host[o + 2 * nt[a] + q] = x[q];
            }
            o += 3 * (size_t)nt[a];
        }
        PPS_TRY(buf.ensure(sizeof(double) * tot));
        PPS_HIP(hipMemcpyAsync(buf.ptr, host.data(), sizeof(double) * tot, hipMemcpyHostToDevice, s));
        PPS_HIP(hipStreamSynchronize(s));
        double *b = buf.as<double>();
        ax = b + off[0]; ay = b + off[1]; az = b + off[2];
        xc = ax + 2 * nt[0]; yc = ay + 2 * nt[1]; zc = az + 2 * nt[2];
        return PPS_OK;
    }
    ~Coords() { buf.release(); }
};

}  // namespace pps

using namespace pps;

static hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

static int grid_ok(const pps_grid *g) {
    if (!g || g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->ng < 1) return PPS_ERR_BAD_ARG;
    for (int d = 0; d < 3; ++d)
        if ((g->bdry_cond[d] == 1) != (g->bdry_cond[d + 3] == 1)) return PPS_ERR_BAD_ARG;
    return PPS_OK;
}

extern "C" int pps_coords(int N, int ng, double length, double *x, double *xh,
                          double *dx, double *dxh) {
    if (N <= 0 || ng < 0 || !x || !xh || !dx || !dxh) return PPS_ERR_BAD_ARG;
    const int nt = N + 2 * ng;
    // paris.f90:3046-3102 with xform = 0; entry q holds Fortran index q+1
    for (int q = 1; q <= nt; ++q) {
        const double s = (double)(q - ng) / (double)N;
        const double form = 0.0 * s * (0.5 - s) * (1.0 - s);
        xh[q - 1] = length * (form + s);
    }
    for (int q = 2; q <= nt; ++q) x[q - 1] = 0.5 * (xh[q - 1] + xh[q - 2]);
    x[0] = 2.0 * xh[0] - x[1];
    for (int q = 1; q <= nt - 1; ++q) dxh[q - 1] = x[q] - x[q - 1];
    dxh[nt - 1] = dxh[nt - 2];
    for (int q = 2; q <= nt; ++q) dx[q - 1] = xh[q - 1] - xh[q - 2];
    dx[0] = dx[1];
    return PPS_OK;
}

// A tiny stand-alone handle for the free functions that need reduction
// scratch; it lives until the process ends.
static pps_handle *scratch_handle() {
    static pps_handle *h = nullptr;
    if (!h) {
        h = new pps_handle();
        (void)hipGetDevice(&h->device);
    }
    return h;
}

extern "C" int pps_pressure_masks_device(const pps_grid *g, double *um, double *vm,
                                         double *wm, void *stream) {
    PPS_TRY(grid_ok(g));
    const Dims d = dims_of(*g);
    hipStream_t s = as_stream(stream);
    k_masks<<<lin_blocks(d.padded()), 256, 0, s>>>(dk(d), g->bdry_cond[0], g->bdry_cond[1],
                                                  g->bdry_cond[2], g->bdry_cond[3],
                                                  g->bdry_cond[4], g->bdry_cond[5], um, vm, wm);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

// linfunc + do_all_ghost: field (padded) from cvof (padded)
extern "C" int pps_linfunc_device(const pps_grid *g, const double *cvof, double a1, double a2,
                                  int mean_flag, int nfilter, double *field, void *stream) {
    PPS_TRY(grid_ok(g));
    if (!cvof || !field) return PPS_ERR_BAD_ARG;
    if (mean_flag != PPS_ARITH_MEAN && mean_flag != PPS_HARM_MEAN) return PPS_ERR_BAD_ARG;
    if (nfilter != 0 && nfilter != 1) return PPS_ERR_BAD_ARG;
    if (nfilter == 1 && field == cvof) return PPS_ERR_BAD_ARG;   // the filter reads neighbours
    const Dims d = dims_of(*g);
    hipStream_t s = as_stream(stream);
    k_linfunc<<<lin_blocks(d.padded()), 256, 0, s>>>(dk(d), cvof, a1, a2, mean_flag == PPS_HARM_MEAN,
                                                     nfilter, field);
    for (int a = 0; a < 3; ++a)
        if (g->bdry_cond[a] == 1) k_wrap_axis<<<lin_blocks(d.padded()), 256, 0, s>>>(dk(d), a, field);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

// {dx, dxh} per axis from the caller's padded host arrays into one device
// buffer; ax / ay / az point into it
static int upload_spacings(const pps_grid *g, const double *dx, const double *dxh,
                           const double *dy, const double *dyh, const double *dz,
                           const double *dzh, DevBuf &buf, const double *ax[3], hipStream_t s) {
    const int nt[3] = {g->nx + 2 * g->ng, g->ny + 2 * g->ng, g->nz + 2 * g->ng};
    const double *src[3][2] = {{dx, dxh}, {dy, dyh}, {dz, dzh}};
    std::vector<double> host(2 * ((size_t)nt[0] + nt[1] + nt[2]));
    size_t off[3], o = 0;
    for (int a = 0; a < 3; ++a) {
        off[a] = o;
        for (int h = 0; h < 2; ++h)
            for (int q = 0; q < nt[a]; ++q) host[o + (size_t)h * nt[a] + q] = src[a][h][q];
        o += 2 * (size_t)nt[a];
    }
    PPS_TRY(buf.ensure(sizeof(double) * host.size()));
    PPS_HIP(hipMemcpyAsync(buf.ptr, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice, s));
    PPS_HIP(hipStreamSynchronize(s));
    for (int a = 0; a < 3; ++a) ax[a] = buf.as<double>() + off[a];
    return PPS_OK;
}

// ax / ay / az: {dx, dxh} of each axis on the device, padded (n + 2 ng each)
static int setup_poisson_with(const pps_grid *g, const double *u, const double *v, const double *w,
                              const double *um, const double *vm, const double *wm,
                              const double *rho, double dt, double volume_source, const double *ax,
                              const double *ay, const double *az, const double *bp, double *A,
                              hipStream_t s) {
    if (!u || !v || !w || !um || !vm || !wm || !rho || !A) return PPS_ERR_BAD_ARG;
    const Dims d = dims_of(*g);
    pps_handle *h = scratch_handle();
    PPS_TRY(ensure_red(h));
    BcK bc;
    for (int f = 0; f < 6; ++f) {
        bc.b[f] = g->bdry_cond[f];
        bc.bp[f] = bp ? bp[f] : 0.0;
    }
    PPS_HIP(hipMemsetAsync(red_flag(h), 0, sizeof(int), s));
    const RowK R{u, v, w, um, vm, wm, rho, ax, ay, az, dt, volume_source, bc};
    k_setup_poisson<<<cgrid(d), kBlk, 0, s>>>(dk(d), R, A, red_flag(h));
    PPS_HIP(hipGetLastError());
    int flag = 0;
    PPS_HIP(hipMemcpyAsync(&flag, red_flag(h), sizeof(int), hipMemcpyDeviceToHost, s));
    PPS_HIP(hipStreamSynchronize(s));
    return flag ? PPS_ERR_NAN_OR_TINY_DIAG : PPS_OK;
}

// codes 0..6 (modules.f90:877): wall, periodic, shear, vel-in, vel-out,
// pressure, pres-out
static int assembly_bc_ok(const pps_grid *g) {
    for (int f = 0; f < 6; ++f)
        if (g->bdry_cond[f] > 6 || g->bdry_cond[f] < 0) return PPS_ERR_UNSUPPORTED;
    return PPS_OK;
}

extern "C" int pps_setup_poisson_bp_device(const pps_grid *g, const double *u, const double *v,
                                           const double *w, const double *um, const double *vm,
                                           const double *wm, const double *rho, double dt,
                                           double volume_source, const double *boundary_pressure,
                                           double *A, void *stream) {
    PPS_TRY(grid_ok(g));
    PPS_TRY(assembly_bc_ok(g));
    hipStream_t s = as_stream(stream);
    Coords co;
    PPS_TRY(co.make(*g, s));
    return setup_poisson_with(g, u, v, w, um, vm, wm, rho, dt, volume_source, co.ax, co.ay, co.az,
                              boundary_pressure, A, s);
}

extern "C" int pps_setup_poisson_device(const pps_grid *g, const double *u, const double *v,
                                        const double *w, const double *um, const double *vm,
                                        const double *wm, const double *rho, double dt,
                                        double volume_source, double *A, void *stream) {
    return pps_setup_poisson_bp_device(g, u, v, w, um, vm, wm, rho, dt, volume_source, nullptr, A,
                                       stream);
}

extern "C" int pps_setup_poisson_coords_device(const pps_grid *g, const double *u, const double *v,
                                               const double *w, const double *um, const double *vm,
                                               const double *wm, const double *rho, double dt,
                                               double volume_source, const double *dx,
                                               const double *dxh, const double *dy,
                                               const double *dyh, const double *dz,
                                               const double *dzh, double *A, void *stream) {
    PPS_TRY(grid_ok(g));
    PPS_TRY(assembly_bc_ok(g));
    if (!dx || !dxh || !dy || !dyh || !dz || !dzh) return PPS_ERR_BAD_ARG;
    hipStream_t s = as_stream(stream);
    DevBuf buf;
    const double *ax[3];
    int rc = upload_spacings(g, dx, dxh, dy, dyh, dz, dzh, buf, ax, s);
    if (rc == PPS_OK)
        rc = setup_poisson_with(g, u, v, w, um, vm, wm, rho, dt, volume_source, ax[0], ax[1], ax[2],
                                nullptr, A, s);
    buf.release();
    return rc;
}

extern "C" int pps_check_setup_device(const pps_grid *g, const double *A, void *stream) {
    PPS_TRY(grid_ok(g));
    if (!A) return PPS_ERR_BAD_ARG;
    return check_setup(scratch_handle(), dims_of(*g), A, as_stream(stream));
}

extern "C" int pps_check_poisson_setup_device(const pps_grid *g, const double *A,
                                              const double *umask, const double *vmask,
                                              const double *wmask, void *stream) {
    PPS_TRY(grid_ok(g));
    if (!A || !umask || !vmask || !wmask) return PPS_ERR_BAD_ARG;
    return check_with(scratch_handle(), dims_of(*g), A, umask, vmask, wmask, as_stream(stream));
}

extern "C" int pps_remove_rhs_mean_device(const pps_grid *g, double *A, void *stream) {
    PPS_TRY(grid_ok(g));
    const Dims d = dims_of(*g);
    hipStream_t s = as_stream(stream);
    pps_handle *h = scratch_handle();
    PPS_TRY(ensure_red(h));
    const long long n = d.cells();
    const int nb = red_blocks(n);
    k_sum_plane<<<nb, 256, 0, s>>>(A + 7 * n, n, red_part(h));
    k_final<<<1, 256, 0, s>>>(red_part(h), nb, 1, 0, red_out(h));
    k_sub_mean<<<lin_blocks(n), 256, 0, s>>>(A + 7 * n, n, red_out(h));
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

extern "C" int pps_calc_residuals_device(const pps_grid *g, const double *A, const double *p,
                                         double out[3], void *stream) {
    PPS_TRY(grid_ok(g));
    const Dims d = dims_of(*g);
    double st[3];
    PPS_TRY(residual_stats(scratch_handle(), d, A, p, st, as_stream(stream)));
    out[0] = st[0] / (double)g->nx / (double)g->ny / (double)g->nz;
    out[1] = std::sqrt(st[1] / (double)g->nx / (double)g->ny / (double)g->nz);
    out[2] = st[2];
    return PPS_OK;
}

namespace pps {

// ghost layers of three padded arrays: dst = src (one thread a padded cell)
__global__ void __launch_bounds__(256)
k_copy_ghosts(DK d, int Z, const double *__restrict__ s0, const double *__restrict__ s1,
              const double *__restrict__ s2, double *__restrict__ d0, double *__restrict__ d1,
              double *__restrict__ d2) {
    const long long n = (long long)d.X * d.Y * Z;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % d.X), j = (int)((t / d.X) % d.Y), k = (int)(t / ((long long)d.X * d.Y));
    const bool ghost = i < d.ng || i >= d.nx + d.ng || j < d.ng || j >= d.ny + d.ng || k < d.ng ||
                       k >= d.nz + d.ng;
    if (ghost) {
        d0[t] = s0[t];
        d1[t] = s1[t];
        d2[t] = s2[t];
    }
}

int copy_ghosts(const Dims &d, const double *const src[3], double *const dst[3], hipStream_t s) {
    const long long n = d.padded();
    k_copy_ghosts<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dk(d), d.Z(), src[0], src[1], src[2],
                                                              dst[0], dst[1], dst[2]);
    PPS_HIP(hipGetLastError());
    return PPS_OK;
}

int handle_spacing(pps_handle *h, const double *dh[3], hipStream_t s) {
    const pps_grid &g = h->grid;
    const int nt[3] = {g.nx + 2 * g.ng, g.ny + 2 * g.ng, g.nz + 2 * g.ng};
    if (!h->spacing.ptr) {
        Coords co;
        PPS_TRY(co.make(g, s));
        PPS_TRY(h->spacing.ensure(co.buf.bytes));
        PPS_HIP(hipMemcpyAsync(h->spacing.ptr, co.buf.ptr, co.buf.bytes, hipMemcpyDeviceToDevice, s));
        PPS_HIP(hipStreamSynchronize(s));
    }
    // per axis: dx, dxh, x, each nt long (Coords layout)
    const double *b = h->spacing.as<double>();
    dh[0] = b + nt[0];
    dh[1] = b + 3 * nt[0] + nt[1];
    dh[2] = b + 3 * nt[0] + 3 * nt[1] + nt[2];
    return PPS_OK;
}

// {dx, dxh} of each axis of the handle's grid on the device (padded), the
// layout RowK reads
static int handle_axes(pps_handle *h, const double *ax[3], hipStream_t s) {
    const double *dh[3];
    PPS_TRY(handle_spacing(h, dh, s));
    const pps_grid &g = h->grid;
    const int nt[3] = {g.nx + 2 * g.ng, g.ny + 2 * g.ng, g.nz + 2 * g.ng};
    for (int a = 0; a < 3; ++a) ax[a] = dh[a] - nt[a];
    return PPS_OK;
}

}  // namespace pps

// ax / ay / az: {dx, dxh} of each axis on the device, padded
static int project_with(const pps_grid *g, double *u, double *v, double *w, const double *um,
                        const double *vm, const double *wm, const double *rho, const double *p,
                        double dt, const double *ax, const double *ay, const double *az,
                        hipStream_t s) {
    const Dims d = dims_of(*g);
    const int nt[3] = {g->nx + 2 * g->ng, g->ny + 2 * g->ng, g->nz + 2 * g->ng};
    const double *dh[3] = {ax + nt[0], ay + nt[1], az + nt[2]};
    double *vel[3] = {u, v, w};
    const double *mk[3] = {um, vm, wm};
    for (int c = 0; c < 3; ++c) {
        const int last = g->bdry_cond[c] != 1 ? 1 : 0;
        k_project<<<cgrid(d), kBlk, 0, s>>>(dk(d), c, last, vel[c], mk[c], rho, p, dh[c], dt);
    }
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

extern "C" int pps_project_velocity_device(const pps_grid *g, double *u, double *v, double *w,
                                           const double *um, const double *vm, const double *wm,
                                           const double *rho, const double *p, double dt,
                                           void *stream) {
    PPS_TRY(grid_ok(g));
    hipStream_t s = as_stream(stream);
    Coords co;
    PPS_TRY(co.make(*g, s));
    return project_with(g, u, v, w, um, vm, wm, rho, p, dt, co.ax, co.ay, co.az, s);
}


extern "C" int pps_project_velocity_coords_device(const pps_grid *g, double *u, double *v, double *w,
                                                  const double *um, const double *vm,
                                                  const double *wm, const double *rho,
                                                  const double *p, double dt, const double *dx,
                                                  const double *dxh, const double *dy,
                                                  const double *dyh, const double *dz,
                                                  const double *dzh, void *stream) {
    PPS_TRY(grid_ok(g));
    if (!dx || !dxh || !dy || !dyh || !dz || !dzh) return PPS_ERR_BAD_ARG;
    hipStream_t s = as_stream(stream);
    DevBuf buf;
    const double *ax[3];
    int rc = upload_spacings(g, dx, dxh, dy, dyh, dz, dzh, buf, ax, s);
    if (rc == PPS_OK) rc = project_with(g, u, v, w, um, vm, wm, rho, p, dt, ax[0], ax[1], ax[2], s);
    buf.release();
    return rc;
}

extern "C" int pps_synth_drops(const pps_grid *g, int ndrop, double rmin, double rmax,
                               unsigned long long seed, double *drops) {
    PPS_TRY(grid_ok(g));
    if (ndrop < 0 || !drops) return PPS_ERR_BAD_ARG;
    // splitmix64 as a sequential generator: state += golden; output mix
    uint64_t st = seed;
    auto next = [&st]() {
        st += 0x9E3779B97F4A7C15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int m = 0; m < ndrop; ++m) {
        drops[4 * m + 0] = g->xlength * next();
        drops[4 * m + 1] = g->ylength * next();
        drops[4 * m + 2] = g->zlength * next();
        drops[4 * m + 3] = rmin + (rmax - rmin) * next();
    }
    return PPS_OK;
}

extern "C" int pps_synth_fields_device(const pps_grid *g, const double *drops, int ndrop,
                                       double rho1, double rho2, double width,
                                       unsigned long long seed, double *rho, double *u,
                                       double *v, double *w, void *stream) {
    PPS_TRY(grid_ok(g));
    if (ndrop < 0 || (ndrop > 0 && !drops)) return PPS_ERR_BAD_ARG;
    const Dims d = dims_of(*g);
    hipStream_t s = as_stream(stream);
    Coords co;
    PPS_TRY(co.make(*g, s));
    DevBuf dd;
    if (ndrop > 0) {
        PPS_TRY(dd.ensure(sizeof(double) * 4 * ndrop));
        PPS_HIP(hipMemcpyAsync(dd.ptr, drops, sizeof(double) * 4 * ndrop, hipMemcpyHostToDevice, s));
    }
    double h = g->xlength / g->nx;
    if (g->ylength / g->ny < h) h = g->ylength / g->ny;
    if (g->zlength / g->nz < h) h = g->zlength / g->nz;
    k_synth_rho<<<cgrid(d), kBlk, 0, s>>>(dk(d), co.xc, co.yc, co.zc, dd.as<double>(), ndrop, rho1,
                                          rho2, width * h, g->xlength, g->ylength, g->zlength,
                                          d.per[0], d.per[1], d.per[2], rho);
    double *vel[3] = {u, v, w};
    for (int c = 0; c < 3; ++c)
        k_synth_vel<<<cgrid(d), kBlk, 0, s>>>(dk(d), (uint64_t)seed, c, d.per[c] ? 0 : 1, vel[c]);
    for (int dir = 0; dir < 3; ++dir) {
        k_fill_ghosts<<<lin_blocks(d.padded()), 256, 0, s>>>(dk(d), dir, d.per[dir], 0, rho);
        for (int c = 0; c < 3; ++c)
            k_fill_ghosts<<<lin_blocks(d.padded()), 256, 0, s>>>(dk(d), dir, d.per[dir], 1, vel[c]);
    }
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    dd.release();
    return PPS_OK;
}

// SetupPoisson + Poisson_BCs (codes 0-4) fused into the fast solve's setup
// pass, and optionally project_velocity fused into its last residual pass:
// the 8-slot A is never formed.
extern "C" int pps_solve_fields_device(pps_handle *h, const double *u, const double *v,
                                       const double *w, const double *umask, const double *vmask,
                                       const double *wmask, const double *rho, double dt,
                                       double volume_source, double *dp, double maxError, int maxit,
                                       int *it, double *tres2, double *u_out, double *v_out,
                                       double *w_out, void *stream) {
    if (!h || !u || !v || !w || !umask || !vmask || !wmask || !rho || !dp || !it || !tres2 || maxit < 1)
        return PPS_ERR_BAD_ARG;
    const bool proj = u_out || v_out || w_out;
    if (proj && (!u_out || !v_out || !w_out || u_out == u || v_out == v || w_out == w))
        return PPS_ERR_BAD_ARG;
    const int method = h->opt.method;
    if ((method != PPS_MG_FP64 && method != PPS_MG_MIXED) || h->opt.check_setup || h->have_widths)
        return PPS_ERR_UNSUPPORTED;
    // codes 5 / 6 give non-symmetric rows, which the fast methods refuse
    for (int f = 0; f < 6; ++f)
        if (h->grid.bdry_cond[f] < 0 || h->grid.bdry_cond[f] > 4) return PPS_ERR_UNSUPPORTED;
    (void)hipSetDevice(h->device);
    hipStream_t s = as_stream(stream);
    const Dims D = dims_of(h->grid);
    const double *ax[3];
    PPS_TRY(handle_axes(h, ax, s));
    BcK bc;
    for (int f = 0; f < 6; ++f) {
        bc.b[f] = h->grid.bdry_cond[f];
        bc.bp[f] = 0.0;
    }
    const RowK R{u, v, w, umask, vmask, wmask, rho, ax[0], ax[1], ax[2], dt, volume_source, bc};
    ProjK P{};
    if (proj) {
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
    }
    const int rc = fast_solve(h, nullptr, dp, maxError, maxit, 2, it, tres2, s, false, proj ? &P : nullptr, &R);
    if (rc != PPS_OK) return rc;
    return *it > maxit ? PPS_ERR_NOT_CONVERGED : PPS_OK;
}

// pps_ref.hip -- the reference multigrid (NewSolverMG, MGsolver.f90:34-405)
// and the plain RB-SOR solver (NewSolver_std, newsolver.f90:78-174)
// restated as gfx950 kernels on the reference's own 8-slot layout.
//
// Every arithmetic expression keeps the reference's operand order and no
// product is fused into an add (fp contract off), so a colour pass, a
// residual or a transfer gives the same bits as the Fortran loops; only
// the residual-norm reductions sum in a different order.
//
// Single-GPU halo semantics: periodic directions wrap one ghost layer, wall
// directions leave the ghosts alone (MPI_PROC_NULL in ghost_MG_x/y/z,
// modules.f90:1846-1971).  Only face ghosts are maintained: edge and corner
// ghosts are never read by the 7-point stencil.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>

#include "pps_internal.h"

namespace pps {

struct LvlK {
    int nx, ny, nz, ng, X, Y;
    int per0, per1, per2;
};

static LvlK lvlk(int nx, int ny, int nz, int ng, const int per[3]) {
    LvlK L;
    L.nx = nx; L.ny = ny; L.nz = nz; L.ng = ng;
    L.X = nx + 2 * ng; L.Y = ny + 2 * ng;
    L.per0 = per[0]; L.per1 = per[1]; L.per2 = per[2];
    return L;
}
__host__ __device__ static long long lk_cells(const LvlK &L) { return (long long)L.nx * L.ny * L.nz; }
static long long lk_padded(const LvlK &L) {
    return (long long)L.X * L.Y * (L.nz + 2 * L.ng);
}

__device__ __forceinline__ long long ipad(const LvlK &L, int i, int j, int k) {
    return (long long)(i + L.ng) + (long long)L.X * ((long long)(j + L.ng) + (long long)L.Y * (k + L.ng));
}
__device__ __forceinline__ long long icel(const LvlK &L, int i, int j, int k) {
    return (long long)i + (long long)L.nx * ((long long)j + (long long)L.ny * k);
}

// ---------------------------------------------------------------------
// RedBlackRelax colour pass (newsolver.f90:222-238): cells with
// (i+j+k) % 2 == colour (0-based local indices), in place.
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_rb_pass(LvlK L, const double *__restrict__ A, double *__restrict__ p,
          double beta, int colour) {
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int i = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x) + ((j + k + colour) & 1);
    if (j >= L.ny || i >= L.nx) return;
    const long long n = lk_cells(L);
    const long long c = icel(L, i, j, k), q = ipad(L, i, j, k);
    const long long sy = L.X, sz = (long long)L.X * L.Y;
    const double s = A[c] * p[q - 1] + A[n + c] * p[q + 1] + A[2 * n + c] * p[q - sy] +
                     A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                     A[5 * n + c] * p[q + sz] + A[7 * n + c];
    p[q] = (1.0 - beta) * p[q] + beta / A[6 * n + c] * s;
}

// Colour-split copy of the 8-slot A for the repeated sweeps of NewSolver_std
// and FreeSolver (nx even): per colour (i+j+k)&1 and slot, the cells of that
// colour in pair order h = (k ny + j) nx/2 + i/2.  A colour pass then
// streams only its own half of A, where the interleaved layout fetches
// every line of all 8 planes in both passes.  Values are copied, so the
// updates stay bit for bit those of k_rb_pass.
__global__ void __launch_bounds__(256)
k_split_colours(LvlK L, const double *__restrict__ A, double *__restrict__ As) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long long n = lk_cells(L), hn = n / 2;
    const long long c = icel(L, i, j, k);
    const int colour = (i + j + k) & 1;
    const long long h = ((long long)k * L.ny + j) * (L.nx / 2) + (i >> 1);
    double *a = As + (long long)colour * 8 * hn;
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s * hn + h] = A[s * n + c];
}

// k_rb_pass on the colour-split A (the same expression)
__global__ void __launch_bounds__(256)
k_rb_pass_split(LvlK L, const double *__restrict__ As, double *__restrict__ p,
                const int *__restrict__ pcmask, int flag, double beta, int colour) {
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = 2 * m + ((j + k + colour) & 1);
    if (j >= L.ny || i >= L.nx) return;
    const long long q = ipad(L, i, j, k);
    if (pcmask && !(flag == 1 ? pcmask[q] == 0 : (pcmask[q] == 1 || pcmask[q] == 2))) return;
    const long long hn = lk_cells(L) / 2;
    const long long h = ((long long)k * L.ny + j) * (L.nx / 2) + m;
    const double *a = As + (long long)colour * 8 * hn;
    const long long sy = L.X, sz = (long long)L.X * L.Y;
    const double s = a[h] * p[q - 1] + a[hn + h] * p[q + 1] + a[2 * hn + h] * p[q - sy] +
                     a[3 * hn + h] * p[q + sy] + a[4 * hn + h] * p[q - sz] +
                     a[5 * hn + h] * p[q + sz] + a[7 * hn + h];
    p[q] = (1.0 - beta) * p[q] + beta / a[6 * hn + h] * s;
}

// 256-thread block sum / max (t: linear thread id, wavefronts of 64
// consecutive ids), result in thread 0 (sm: 4 doubles)
__device__ __forceinline__ double blk_red(double v, double *sm, bool mx, int t) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_down(v, o, 64);
        v = mx ? fmax(v, w) : v + w;
    }
    if ((t & 63) == 0) sm[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) {
        r = sm[0];
        for (int w = 1; w < 4; ++w) r = mx ? fmax(r, sm[w]) : r + sm[w];
    }
    __syncthreads();
    return r;
}

// FreeSolver on the colour-split A, the residual statistics fused into the
// sweep: the cells of one colour (64 x 4 pair tiles, the planes strided by
// gridDim.z), with
// UPDATE the RB-SOR update of k_rb_pass_split first.  Every neighbour of a
// cell has the other colour, so after the second pass both the cells it
// updated (UPDATE) and, in a pass of its own, the first colour's cells see
// final values: their statistics equal k_resid_stats_fs's (same residual
// expression, sums in another order).  Block partials {sum r^2, max |r|,
// active cells, NaN cells} at part[c * kRedBlocks + boff + block].
// !FS (NewSolver_std): every cell, the residual of MGsolver.f90:97-100 as
// compute_residual_norm forms it, partial 0 = sum |r|^norm.
template <bool UPDATE, bool FS = true>
__global__ void __launch_bounds__(256)
k_fs_pass_stats(LvlK L, const double *__restrict__ As, double *__restrict__ p,
                const int *__restrict__ pcmask, int flag, double beta, int colour,
                double *__restrict__ part, int boff, int norm = 2) {
    __shared__ double sm[4];
    const long long hn = lk_cells(L) / 2;
    const int hx = L.nx / 2;
    const long long sy = L.X, sz = (long long)L.X * L.Y;
    const double *a = As + (long long)colour * 8 * hn;
    double s2 = 0.0, mx = 0.0, cnt = 0.0, bad = 0.0;
    const int m = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    for (int k = blockIdx.z; k < L.nz && m < hx && j < L.ny; k += gridDim.z) {
        const long long h = ((long long)k * L.ny + j) * hx + m;
        const int i = 2 * m + ((j + k + colour) & 1);
        const long long q = ipad(L, i, j, k);
        if (FS) {
            const int mk = pcmask[q];
            if (!(flag == 1 ? mk == 0 : (mk == 1 || mk == 2))) continue;
        }
        double pc = p[q];
        if (UPDATE) {
            const double s = a[h] * p[q - 1] + a[hn + h] * p[q + 1] + a[2 * hn + h] * p[q - sy] +
                             a[3 * hn + h] * p[q + sy] + a[4 * hn + h] * p[q - sz] +
                             a[5 * hn + h] * p[q + sz] + a[7 * hn + h];
            pc = (1.0 - beta) * pc + beta / a[6 * hn + h] * s;
            p[q] = pc;
        }
        if (!FS) {
            const double r = fabs(a[7 * hn + h] - a[6 * hn + h] * pc + a[h] * p[q - 1] +
                                  a[hn + h] * p[q + 1] + a[2 * hn + h] * p[q - sy] +
                                  a[3 * hn + h] * p[q + sy] + a[4 * hn + h] * p[q - sz] +
                                  a[5 * hn + h] * p[q + sz]);
            s2 += norm == 1 ? r : (norm == 2 ? r * r : pow(r, (double)norm));
            continue;
        }
        const double r = fabs(-pc * a[6 * hn + h] + a[h] * p[q - 1] + a[hn + h] * p[q + 1] +
                              a[2 * hn + h] * p[q - sy] + a[3 * hn + h] * p[q + sy] +
                              a[4 * hn + h] * p[q - sz] + a[5 * hn + h] * p[q + sz] + a[7 * hn + h]);
        bool nan = r != r;
#pragma unroll
        for (int s = 0; s < 8; ++s) nan |= a[s * hn + h] != a[s * hn + h];
        s2 += r * r;
        mx = fmax(mx, r);
        cnt += 1.0;
        if (nan) bad += 1.0;
    }
    const int t = threadIdx.x + 64 * threadIdx.y;
    const long long b = boff + blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    double v = blk_red(s2, sm, false, t);
    if (t == 0) part[b] = v;
    v = blk_red(mx, sm, true, t);
    if (t == 0) part[kRedBlocks + b] = v;
    v = blk_red(cnt, sm, false, t);
    if (t == 0) part[2 * kRedBlocks + b] = v;
    v = blk_red(bad, sm, false, t);
    if (t == 0) part[3 * kRedBlocks + b] = v;
}

// One-layer periodic wrap of the face ghosts of a level.
__global__ void __launch_bounds__(256)
k_halo(LvlK L, double *__restrict__ p) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (L.per0 && a < L.ny && b < L.nz) {
        p[ipad(L, -1, a, b)] = p[ipad(L, L.nx - 1, a, b)];
        p[ipad(L, L.nx, a, b)] = p[ipad(L, 0, a, b)];
    }
    if (L.per1 && a < L.nx && b < L.nz) {
        p[ipad(L, a, -1, b)] = p[ipad(L, a, L.ny - 1, b)];
        p[ipad(L, a, L.ny, b)] = p[ipad(L, a, 0, b)];
    }
    if (L.per2 && a < L.nx && b < L.ny) {
        p[ipad(L, a, b, -1)] = p[ipad(L, a, b, L.nz - 1)];
        p[ipad(L, a, b, L.nz)] = p[ipad(L, a, b, 0)];
    }
}

// get_residual in place (MGsolver.f90:96-102): slot 8 <- r.
__global__ void __launch_bounds__(256)
k_residual_inplace(LvlK L, double *__restrict__ A, const double *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long long n = lk_cells(L);
    const long long c = icel(L, i, j, k), q = ipad(L, i, j, k);
    const long long sy = L.X, sz = (long long)L.X * L.Y;
    A[7 * n + c] = A[7 * n + c] - A[6 * n + c] * p[q] + A[c] * p[q - 1] + A[n + c] * p[q + 1] +
                   A[2 * n + c] * p[q - sy] + A[3 * n + c] * p[q + sy] +
                   A[4 * n + c] * p[q - sz] + A[5 * n + c] * p[q + sz];
}

// fill_coefficients (MGsolver.f90:210-220): all 8 slots = child mean.
__global__ void __launch_bounds__(256)
k_fill_coefficients(int fx, int fy, int fz, const double *__restrict__ Af,
                    double *__restrict__ Ac) {
    const int cx = fx / 2, cy = fy / 2, cz = fz / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= cx || j >= cy) return;
    const long long nf = (long long)fx * fy * fz, nc = (long long)cx * cy * cz;
    const long long sy = fx, sz = (long long)fx * fy;
    const long long f0 = 2LL * i + sy * (2LL * j) + sz * (2LL * k);
    const long long co = i + (long long)cx * (j + (long long)cy * k);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const double *a = Af + s * nf;
        const double t = a[f0] + a[f0 + sy] + a[f0 + sz] + a[f0 + sy + sz] + a[f0 + 1] +
                         a[f0 + 1 + sy] + a[f0 + 1 + sz] + a[f0 + 1 + sy + sz];
        Ac[s * nc + co] = 0.125 * t;
    }
}

// coarse_from_fine (MGsolver.f90:193-199): coarse slot 8 = sum of
// 0.125 * child residual, children in (i1, j1, k1) order, k1 fastest.
__global__ void __launch_bounds__(256)
k_coarse_from_fine(int fx, int fy, int fz, const double *__restrict__ Af,
                   double *__restrict__ Ac) {
    const int cx = fx / 2, cy = fy / 2, cz = fz / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= cx || j >= cy) return;
    const long long nf = (long long)fx * fy * fz, nc = (long long)cx * cy * cz;
    const long long sy = fx, sz = (long long)fx * fy;
    const double *a = Af + 7 * nf;
    const long long f0 = 2LL * i + sy * (2LL * j) + sz * (2LL * k);
    double t = 0.0;
    for (int i1 = 0; i1 < 2; ++i1)
        for (int j1 = 0; j1 < 2; ++j1)
            for (int k1 = 0; k1 < 2; ++k1) t = t + a[f0 + i1 + sy * j1 + sz * k1] * 0.125;
    Ac[7 * nc + i + (long long)cx * (j + (long long)cy * k)] = t;
}

// coarse_fine_interp (MGsolver.f90:128-183) into the children of each
// coarse cell; sym = sym_MG.
__global__ void __launch_bounds__(256)
k_coarse_fine_interp(LvlK F, LvlK C, double *__restrict__ ef,
                     const double *__restrict__ ec, int sym) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= C.nx || j >= C.ny) return;
    const double pc = ec[ipad(C, i, j, k)];
    if (sym) {
        const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                             {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        double tot = 0.0;
        for (int m = 0; m < 8; ++m)
            tot = tot + fabs(ef[ipad(F, 2 * i + o[m][0], 2 * j + o[m][1], 2 * k + o[m][2])]);
        tot = tot + 1e-15;
        for (int m = 0; m < 8; ++m) {
            const long long q = ipad(F, 2 * i + o[m][0], 2 * j + o[m][1], 2 * k + o[m][2]);
            ef[q] = fabs(ef[q]) / tot * pc;
        }
    } else {
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) ef[ipad(F, 2 * i + di, 2 * j + dj, 2 * k + dk)] = pc;
    }
}

__global__ void __launch_bounds__(256)
k_zero(double *__restrict__ x, long long n) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x)
        x[q] = 0.0;
}

__global__ void __launch_bounds__(256)
k_add(double *__restrict__ y, const double *__restrict__ x, long long n) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x)
        y[q] = y[q] + x[q];
}

// ---------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------
struct RefState {
    int lmg = 0;
    std::vector<LvlK> lv;          // lv[1..lmg]
    std::vector<DevBuf> data, pmg, emg;
};

void ref_destroy(RefState *r) {
    if (!r) return;
    for (auto &b : r->data) b.release();
    for (auto &b : r->pmg) b.release();
    for (auto &b : r->emg) b.release();
    delete r;
}

static dim3 cell_grid(int nx, int ny, int nz) {
    return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz);
}
static const dim3 kCellBlock(64, 4, 1);

static unsigned lin_blocks(long long n) {
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

static int ref_ensure(pps_handle *h) {
    if (h->ref) return PPS_OK;
    const pps_grid &g = h->grid;
    int nd[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; ++d)
        if (nd[d] <= 0 || (nd[d] & (nd[d] - 1))) return PPS_ERR_BAD_SIZE;
    int lmg = 1 << 30;
    for (int d = 0; d < 3; ++d) {
        int l = 0;
        while ((1 << (l + 1)) <= nd[d]) ++l;
        if (l < lmg) lmg = l;
    }
    if (lmg < 2) return PPS_ERR_BAD_SIZE;
    RefState *r = new RefState();
    r->lmg = lmg;
    r->lv.resize(lmg + 1);
    r->data.resize(lmg + 1);
    r->pmg.resize(lmg + 1);
    r->emg.resize(lmg + 1);
    Dims D = dims_of(g);
    for (int l = lmg; l >= 1; --l) {
        r->lv[l] = lvlk(nd[0] >> (lmg - l), nd[1] >> (lmg - l), nd[2] >> (lmg - l), g.ng, D.per);
        int rc = r->data[l].ensure(sizeof(double) * 8 * lk_cells(r->lv[l]));
        if (rc == PPS_OK) rc = r->pmg[l].ensure(sizeof(double) * lk_padded(r->lv[l]));
        if (rc == PPS_OK) rc = r->emg[l].ensure(sizeof(double) * lk_padded(r->lv[l]));
        if (rc != PPS_OK) { ref_destroy(r); return rc; }
    }
    h->ref = r;
    return PPS_OK;
}

static void launch_zero(double *x, long long n, hipStream_t s) {
    k_zero<<<lin_blocks(n), 256, 0, s>>>(x, n);
}
static void launch_add(double *y, const double *x, long long n, hipStream_t s) {
    k_add<<<lin_blocks(n), 256, 0, s>>>(y, x, n);
}
static void launch_halo(const LvlK &L, double *p, hipStream_t s) {
    if (!(L.per0 | L.per1 | L.per2)) return;
    int a = L.nx > L.ny ? L.nx : L.ny;
    int b = L.ny > L.nz ? L.ny : L.nz;
    if (L.nz > b) b = L.nz;
    k_halo<<<dim3((unsigned)((a + 255) / 256), (unsigned)b), 256, 0, s>>>(L, p);
}
// RedBlackRelax: pass 0, halo, pass 1 (newsolver.f90:222-246)
static void launch_relax(const LvlK &L, const double *A, double *p, double beta, hipStream_t s) {
    dim3 grid((unsigned)(((L.nx + 1) / 2 + 63) / 64), (unsigned)((L.ny + 3) / 4), (unsigned)L.nz);
    k_rb_pass<<<grid, kCellBlock, 0, s>>>(L, A, p, beta, 0);
    launch_halo(L, p, s);
    k_rb_pass<<<grid, kCellBlock, 0, s>>>(L, A, p, beta, 1);
}
// The caller's A split by colour into h->asplit when nx is even (else
// nullptr: the interleaved kernels run).
static const double *split_colours(pps_handle *h, const LvlK &L, const double *A, hipStream_t s,
                                   int *rc) {
    *rc = PPS_OK;
    if (L.nx & 1) return nullptr;
    const long long n = (long long)L.nx * L.ny * L.nz;
    *rc = h->asplit.ensure(sizeof(double) * 8 * (size_t)n);
    if (*rc != PPS_OK) return nullptr;
    k_split_colours<<<cell_grid(L.nx, L.ny, L.nz), kCellBlock, 0, s>>>(L, A, h->asplit.as<double>());
    return h->asplit.as<double>();
}
static void launch_residual(const LvlK &L, double *A, const double *p, hipStream_t s) {
    k_residual_inplace<<<cell_grid(L.nx, L.ny, L.nz), kCellBlock, 0, s>>>(L, A, p);
}

// MG_iteration (MGsolver.f90:253-320)
static void mg_iteration(RefState *r, double beta, int nrelax, int sym, hipStream_t s) {
    const int lmg = r->lmg;
    for (int l = lmg; l >= 2; --l) {
        const LvlK &L = r->lv[l];
        double *A = r->data[l].as<double>(), *x = r->pmg[l].as<double>();
        launch_zero(x, lk_padded(L), s);
        for (int q = 0; q < nrelax; ++q) {
            launch_relax(L, A, x, beta, s);
            launch_halo(L, x, s);
        }
        launch_residual(L, A, x, s);
        const LvlK &C = r->lv[l - 1];
        k_coarse_from_fine<<<cell_grid(C.nx, C.ny, C.nz), kCellBlock, 0, s>>>(
            L.nx, L.ny, L.nz, A, r->data[l - 1].as<double>());
    }
    launch_relax(r->lv[2], r->data[2].as<double>(), r->pmg[2].as<double>(), beta, s);
    launch_halo(r->lv[2], r->pmg[2].as<double>(), s);
    for (int l = 3; l <= lmg; ++l) {
        const LvlK &L = r->lv[l], &C = r->lv[l - 1];
        double *A = r->data[l].as<double>();
        double *e = r->emg[l].as<double>(), *x = r->pmg[l].as<double>();
        const long long np = lk_padded(L);
        launch_zero(e, np, s);
        k_coarse_fine_interp<<<cell_grid(C.nx, C.ny, C.nz), kCellBlock, 0, s>>>(
            L, C, e, r->pmg[l - 1].as<double>(), sym);
        launch_halo(L, e, s);
        launch_residual(L, A, e, s);
        launch_add(x, e, np, s);
        launch_zero(e, np, s);
        for (int q = 0; q < nrelax; ++q) {
            launch_relax(L, A, e, beta, s);
            launch_halo(L, e, s);
        }
        launch_add(x, e, np, s);
        launch_residual(L, A, e, s);
    }
}

int ref_solve(pps_handle *h, const double *dA, double *dp, double maxError,
              double beta, int maxit, int norm, int *it, double *tres2,
              hipStream_t s) {
    PPS_TRY(ref_ensure(h));
    RefState *r = h->ref;
    const int lmg = r->lmg;
    const LvlK &F = r->lv[lmg];
    const Dims D = dims_of(h->grid);
    const long long n = lk_cells(F), np = lk_padded(F);
    const double N = (double)h->grid.nx * (double)h->grid.ny * (double)h->grid.nz;
    h->hist.clear();

    // get_MGmatrix_coef (MGsolver.f90:322-344)
    PPS_HIP(hipMemcpyAsync(r->data[lmg].ptr, dA, sizeof(double) * 8 * n, hipMemcpyDeviceToDevice, s));
    for (int l = lmg - 1; l >= 1; --l) {
        const LvlK &C = r->lv[l], &Fl = r->lv[l + 1];
        k_fill_coefficients<<<cell_grid(C.nx, C.ny, C.nz), kCellBlock, 0, s>>>(
            Fl.nx, Fl.ny, Fl.nz, r->data[l + 1].as<double>(), r->data[l].as<double>());
    }
    // initial residual into slot 8 (MGsolver.f90:370-377)
    double *x = r->pmg[lmg].as<double>();
    PPS_HIP(hipMemcpyAsync(x, dp, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
    double res = 0.0;
    PPS_TRY(compute_residual_norm(h, D, r->data[lmg].as<double>(), x, norm, &res, s));
    double t2 = res / N;
    launch_residual(F, r->data[lmg].as<double>(), x, s);
    launch_zero(x, np, s);
    PPS_HIP(hipGetLastError());
    int k;
    for (k = 1; k <= maxit; ++k) {
        mg_iteration(r, beta, h->opt.nrelax, h->opt.sym_mg, s);
        launch_add(dp, x, np, s);
        PPS_HIP(hipGetLastError());
        PPS_TRY(compute_residual_norm(h, D, dA, dp, norm, &res, s));
        t2 = res / N;
        if (norm == 2) t2 = std::sqrt(t2);
        h->hist.push_back(t2);
        if (t2 < maxError) break;
    }
    *it = k;
    *tres2 = t2;
    return PPS_OK;
}

// NewSolver_std (newsolver.f90:78-174) with catch_divergence folded into
// the residual norm: NaN in A or p makes the norm NaN.
int sor_solve(pps_handle *h, const double *dA, double *dp, double maxError,
              double beta, int maxit, int norm, int *it, double *tres2,
              hipStream_t s) {
    const pps_grid &g = h->grid;
    const Dims D = dims_of(g);
    LvlK L = lvlk(g.nx, g.ny, g.nz, g.ng, D.per);
    const double N = (double)g.nx * (double)g.ny * (double)g.nz;
    h->hist.clear();
    double t2 = 0.0, res = 0.0;
    int rc = PPS_OK;
    const double *As = split_colours(h, L, dA, s, &rc);
    PPS_TRY(rc);
    // split path: pass 0, halo, pass 1 + its cells' residuals, halo, the
    // residuals of pass 0's cells (the planes strided: ~2048 blocks each)
    const dim3 sgrid((unsigned)((L.nx / 2 + 63) / 64), (unsigned)((L.ny + 3) / 4), (unsigned)L.nz);
    const long long tiles = (long long)sgrid.x * sgrid.y;
    const dim3 tgrid(sgrid.x, sgrid.y, (unsigned)std::max<long long>(1, std::min<long long>(L.nz, 2048 / tiles)));
    const int nbk = (int)(tiles * tgrid.z);
    if (As && 2LL * nbk > kRedBlocks) As = nullptr;
    double *part = As ? stats_partials(h) : nullptr;
    if (As && !part) return PPS_ERR_HIP;
    int k;
    for (k = 1; k <= maxit; ++k) {
        if (As) {
            k_rb_pass_split<<<sgrid, kCellBlock, 0, s>>>(L, As, dp, nullptr, 0, beta, 0);
            launch_halo(L, dp, s);
            k_fs_pass_stats<true, false><<<tgrid, kCellBlock, 0, s>>>(L, As, dp, nullptr, 0, beta, 1, part, 0,
                                                                      norm);
            launch_halo(L, dp, s);
            k_fs_pass_stats<false, false><<<tgrid, kCellBlock, 0, s>>>(L, As, dp, nullptr, 0, beta, 0, part,
                                                                       nbk, norm);
            PPS_HIP(hipGetLastError());
            PPS_TRY(stats_finish(h, 2 * nbk, 1, 0, &res, s));
        } else {
            launch_relax(L, dA, dp, beta, s);
            launch_halo(L, dp, s);
            PPS_HIP(hipGetLastError());
            PPS_TRY(compute_residual_norm(h, D, dA, dp, norm, &res, s));
        }
        double res2 = res / N;
        if (!(res2 <= 1e16)) {  // > 1e16 or NaN
            *it = k;
            *tres2 = res2;
            return PPS_ERR_DIVERGED;
        }
        t2 = norm == 2 ? std::sqrt(res2) : res2;
        h->hist.push_back(t2);
        if (t2 < maxError) break;
    }
    *it = k;
    *tres2 = t2;
    return PPS_OK;
}

// ---------------------------------------------------------------------
// FreeSolver (freesurface.f90:700-855): RB-SOR on the cells the
// free-surface mask selects.  Active: pcmask == 0 (solver_flag 1), pcmask
// in {1, 2} (solver_flag 2); pcmask and P_gas are padded like p.
// ---------------------------------------------------------------------
__device__ __forceinline__ bool fs_active(int m, int flag) {
    return flag == 1 ? m == 0 : (m == 1 || m == 2);
}

// freesurface.f90:723-726: p = P_gas where pcmask != 0 (flag 1), p = 0
// where pcmask == 3 (flag 2), on the interior
__global__ void __launch_bounds__(256)
k_fs_init(LvlK L, const int *__restrict__ pcmask, const double *__restrict__ pgas,
          double *__restrict__ p, int flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long long q = ipad(L, i, j, k);
    const int m = pcmask[q];
    if (flag == 1 && m != 0) p[q] = pgas[q];
    if (flag == 2 && m == 3) p[q] = 0.0;
}

// RedBlackRelax_fs colour pass (freesurface.f90:871-893): the update of
// k_rb_pass on the active cells only
__global__ void __launch_bounds__(256)
k_rb_pass_fs(LvlK L, const double *__restrict__ A, double *__restrict__ p,
             const int *__restrict__ pcmask, int flag, double beta, int colour) {
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int i = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x) + ((j + k + colour) & 1);
    if (j >= L.ny || i >= L.nx) return;
    const long long q = ipad(L, i, j, k);
    if (!fs_active(pcmask[q], flag)) return;
    const long long n = lk_cells(L);
    const long long c = icel(L, i, j, k);
    const long long sy = L.X, sz = (long long)L.X * L.Y;
    const double s = A[c] * p[q - 1] + A[n + c] * p[q + 1] + A[2 * n + c] * p[q - sy] +
                     A[3 * n + c] * p[q + sy] + A[4 * n + c] * p[q - sz] +
                     A[5 * n + c] * p[q + sz] + A[7 * n + c];
    p[q] = (1.0 - beta) * p[q] + beta / A[6 * n + c] * s;
}

// FreeSolver: initial values, then per iteration one masked RB-SOR sweep
// (pass 0, halo, pass 1), the halo, and the residual statistics of the
// active cells; stops when max |r| < maxError (use_L_inf), reports
// tres2 = sqrt(sum r^2) / (active cells) (freesurface.f90:786-803);
// it = maxit + 1 when not converged.  NaN or sum r^2 > 1e18 returns
// PPS_ERR_DIVERGED (catch_divergence_fs with one rank).
int free_solve(pps_handle *h, const double *dA, double *dp, const int *dmask, const double *dpgas,
               int solver_flag, double maxError, double beta, int maxit, int *it, double *tres2,
               hipStream_t s) {
    if (solver_flag != 1 && solver_flag != 2) return PPS_ERR_BAD_ARG;   // freesurface.f90:722
    const pps_grid &g = h->grid;
    const Dims D = dims_of(g);
    LvlK L = lvlk(g.nx, g.ny, g.nz, g.ng, D.per);
    h->hist.clear();
    k_fs_init<<<cell_grid(L.nx, L.ny, L.nz), kCellBlock, 0, s>>>(L, dmask, dpgas, dp, solver_flag);
    launch_halo(L, dp, s);
    const dim3 grid((unsigned)(((L.nx + 1) / 2 + 63) / 64), (unsigned)((L.ny + 3) / 4), (unsigned)L.nz);
    double t2 = 0.0;
    int rc = PPS_OK;
    const double *As = split_colours(h, L, dA, s, &rc);
    PPS_TRY(rc);
    int k;
    // split path: pass 0, halo, pass 1 + its cells' statistics, halo, the
    // statistics of pass 0's cells
    // (the statistics kernels stride the planes so that each has at most
    // ~2048 blocks, one partial slot each)
    const dim3 sgrid((unsigned)((L.nx / 2 + 63) / 64), (unsigned)((L.ny + 3) / 4), (unsigned)L.nz);
    const long long tiles = (long long)sgrid.x * sgrid.y;
    const dim3 tgrid(sgrid.x, sgrid.y, (unsigned)std::max<long long>(1, std::min<long long>(L.nz, 2048 / tiles)));
    const int nbk = (int)(tiles * tgrid.z);
    if (As && 2LL * nbk > kRedBlocks) As = nullptr;   // too many row tiles: interleaved path
    double *part = As ? stats_partials(h) : nullptr;
    if (As && !part) return PPS_ERR_HIP;
    for (k = 1; k <= maxit; ++k) {
        double st[4];
        if (As) {
            k_rb_pass_split<<<sgrid, kCellBlock, 0, s>>>(L, As, dp, dmask, solver_flag, beta, 0);
            launch_halo(L, dp, s);
            k_fs_pass_stats<true><<<tgrid, kCellBlock, 0, s>>>(L, As, dp, dmask, solver_flag, beta, 1, part, 0);
            launch_halo(L, dp, s);
            k_fs_pass_stats<false><<<tgrid, kCellBlock, 0, s>>>(L, As, dp, dmask, solver_flag, beta, 0, part,
                                                                nbk);
            PPS_HIP(hipGetLastError());
            PPS_TRY(stats_finish(h, 2 * nbk, 4, 1 << 1, st, s));
        } else {
            k_rb_pass_fs<<<grid, kCellBlock, 0, s>>>(L, dA, dp, dmask, solver_flag, beta, 0);
            launch_halo(L, dp, s);
            k_rb_pass_fs<<<grid, kCellBlock, 0, s>>>(L, dA, dp, dmask, solver_flag, beta, 1);
            launch_halo(L, dp, s);
            PPS_HIP(hipGetLastError());
            PPS_TRY(fs_residual_stats(h, D, dA, dp, dmask, solver_flag, st, s));
        }
        if (st[3] > 0.0 || !(st[0] <= 1e18)) {
            *it = k;
            *tres2 = st[0];
            return PPS_ERR_DIVERGED;
        }
        t2 = std::sqrt(st[0]) / st[2];
        h->hist.push_back(t2);
        if (st[1] < maxError) break;
    }
    *it = k;
    *tres2 = t2;
    return PPS_OK;
}

}  // namespace pps

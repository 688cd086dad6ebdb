// pps_fast.hip -- multigrid-preconditioned CG for the PARIS pressure system
// (the system NewSolverMG solves, MGsolver.f90:346-405), built for gfx950.
//
// Storage.  A symmetric 7-point operator is kept per level as three
// coupling planes and the diagonal, x fastest, no ghost padding:
//   cx[c] = coupling of cell c to its +x neighbour (= A2(c) = A1(c+x)),
//   cy, cz likewise, dg[c] = A7(c).
// The -x coupling of a cell is cx of its -x neighbour.  Neighbour indices
// wrap in every direction: a wall face has coupling 0 (its A2 is folded
// into the right-hand side together with the fixed ghost value), so the
// wrapped read contributes nothing there.  This halves the coefficient
// bytes of the reference's 8-slot layout (4 planes instead of 7).
//
// Hierarchy.  Coarse couplings are 1/16 of the sum of the four fine faces
// that make up a coarse face (the rediscretised mean: 1/4 of the average
// for a face twice as wide, over a spacing twice as long); the diagonal is
// the sum of the coarse couplings plus the mean of the children's excess
// A7 - sum(couplings).  Restriction is the 8-child average of the
// residual (full weighting for cell-centred grids), prolongation is
// trilinear in cell-centred form (weights 3/4, 1/4 per direction, Neumann
// reflection at walls, wrap in periodic directions).  The coarsest level
// (<= 512 cells) is solved with a dense inverse built once per solve.
//
// Precision.  PPS_MG_FP64: everything FP64.  PPS_MG_MIXED: the V-cycle
// runs on FP32 copies of the operator with FP32 vectors and accumulation;
// the outer CG (residual, update, dot products, mean projection) is FP64.
#include <cmath>
#include <cstring>
#include <vector>

#include "pps_internal.h"

namespace pps {

// ---------------------------------------------------------------------
// kernel-side level view
// ---------------------------------------------------------------------
template <class T> struct LevelK {
    int nx, ny, nz;
    const T *cx, *cy, *cz, *dg;
};

__device__ __forceinline__ int wrap_m(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrap_p(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// Apply the stencil: returns b + sum coupling * neighbour  (the numerator
// of a Gauss-Seidel update) for cell (i,j,k).
template <class T, class V>
__device__ __forceinline__ T nb_sum(const LevelK<T> &L, const V *__restrict__ x, int i, int j,
                                    int k, long long c) {
    const long long sy = L.nx, sz = (long long)L.nx * L.ny;
    const int im = wrap_m(i, L.nx), ip = wrap_p(i, L.nx);
    const int jm = wrap_m(j, L.ny), jp = wrap_p(j, L.ny);
    const int km = wrap_m(k, L.nz), kp = wrap_p(k, L.nz);
    const long long cxm = c + (im - i), cxp = c + (ip - i);
    const long long cym = c + (long long)(jm - j) * sy, cyp = c + (long long)(jp - j) * sy;
    const long long czm = c + (long long)(km - k) * sz, czp = c + (long long)(kp - k) * sz;
    T s = L.cx[cxm] * (T)x[cxm] + L.cx[c] * (T)x[cxp];
    s += L.cy[cym] * (T)x[cym] + L.cy[c] * (T)x[cyp];
    s += L.cz[czm] * (T)x[czm] + L.cz[c] * (T)x[czp];
    return s;
}

// ---------------------------------------------------------------------
// red-black Gauss-Seidel colour pass.  colour 0 = (i+j+k) even.  FIRST:
// the iterate is zero before this pass (no read of x).
// ---------------------------------------------------------------------
// FINE does not change the code; it gives the finest level's launches
// their own symbol so profilers can separate them from coarse levels.
template <class T, class B, bool FIRST, bool FINE>
__global__ void __launch_bounds__(256)
k_rb(LevelK<T> L, const B *__restrict__ b, T *__restrict__ x, int colour) {
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int i = 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x) + ((j + k + colour) & 1);
    if (j >= L.ny || i >= L.nx) return;
    const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
    if (FIRST) {
        x[c] = (T)b[c] / L.dg[c];
    } else {
        x[c] = ((T)b[c] + nb_sum(L, x, i, j, k, c)) / L.dg[c];
    }
}

// residual of the 8 children of each coarse cell, averaged into bc
template <class T, class B>
__global__ void __launch_bounds__(256)
k_resid_restrict(LevelK<T> L, const B *__restrict__ b, const T *__restrict__ x,
                 T *__restrict__ bc) {
    const int cxn = L.nx >> 1, cyn = L.ny >> 1;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= cxn || J >= cyn) return;
    T acc = 0;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
                acc += ((T)b[c] + nb_sum(L, x, i, j, k, c)) - L.dg[c] * x[c];
            }
    bc[I + (long long)cxn * (J + (long long)cyn * K)] = acc * (T)0.125;
}

// x_f += trilinear interpolation of x_c (cell-centred)
template <class T>
__global__ void __launch_bounds__(256)
k_prolong_add(int fx, int fy, int fz, int px, int py, int pz, T *__restrict__ xf,
              const T *__restrict__ xc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= fx || j >= fy) return;
    const int cx = fx >> 1, cy = fy >> 1, cz = fz >> 1;
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    auto other = [](int f, int C, int nC, int per) {
        int o = (f & 1) ? C + 1 : C - 1;
        if (o < 0) o = per ? nC - 1 : 0;
        if (o >= nC) o = per ? 0 : nC - 1;
        return o;
    };
    const int I2 = other(i, I, cx, px), J2 = other(j, J, cy, py), K2 = other(k, K, cz, pz);
    const long long sy = cx, sz = (long long)cx * cy;
    const T w1 = (T)0.75, w2 = (T)0.25;
    auto X = [&](int a, int b_, int c_) { return xc[a + sy * b_ + sz * c_]; };
    const T v = w1 * (w1 * (w1 * X(I, J, K) + w2 * X(I2, J, K)) + w2 * (w1 * X(I, J2, K) + w2 * X(I2, J2, K))) +
                w2 * (w1 * (w1 * X(I, J, K2) + w2 * X(I2, J, K2)) + w2 * (w1 * X(I, J2, K2) + w2 * X(I2, J2, K2)));
    const long long c = i + (long long)fx * (j + (long long)fy * k);
    xf[c] += v;
}

// ---------------------------------------------------------------------
// operator packing: 8-slot FP64 A + padded p -> couplings, diagonal,
// folded right-hand side, interior iterate.  flags: bit0 NaN/tiny A7,
// bit1 asymmetric, bit2 nonzero excess (system not singular).
// ---------------------------------------------------------------------
struct PackK {
    int nx, ny, nz, ng, X, Y;
    int px, py, pz;
};

__global__ void __launch_bounds__(256)
k_pack(PackK d, const double *__restrict__ A, const double *__restrict__ p,
       double *__restrict__ cx, double *__restrict__ cy, double *__restrict__ cz,
       double *__restrict__ dg, double *__restrict__ b, double *__restrict__ x,
       int *__restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    const long long sy = d.nx, sz = (long long)d.nx * d.ny;
    const long long Py = d.X, Pz = (long long)d.X * d.Y;
    const long long q = (i + d.ng) + Py * (j + d.ng) + Pz * (k + d.ng);
    double a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = A[s * n + c];
    int f = 0;
    bool nan = false;
#pragma unroll
    for (int s = 0; s < 8; ++s) nan |= a[s] != a[s];
    if (nan || a[6] < 1e-55) f |= 1;
    double rhs = a[7];
    double kept = 0.0;
    // x direction
    const bool lastx = i == d.nx - 1, lasty = j == d.ny - 1, lastz = k == d.nz - 1;
    double vx = a[1], vy = a[3], vz = a[5];
    if (!d.px) {
        if (i == 0) rhs += a[0] * p[q - 1]; else kept += a[0];
        if (lastx) { rhs += a[1] * p[q + 1]; vx = 0.0; } else kept += a[1];
    } else {
        kept += a[0] + a[1];
    }
    if (!d.py) {
        if (j == 0) rhs += a[2] * p[q - Py]; else kept += a[2];
        if (lasty) { rhs += a[3] * p[q + Py]; vy = 0.0; } else kept += a[3];
    } else {
        kept += a[2] + a[3];
    }
    if (!d.pz) {
        if (k == 0) rhs += a[4] * p[q - Pz]; else kept += a[4];
        if (lastz) { rhs += a[5] * p[q + Pz]; vz = 0.0; } else kept += a[5];
    } else {
        kept += a[4] + a[5];
    }
    // symmetry: my + coupling equals the - coupling of my + neighbour
    if (!lastx || d.px) {
        const long long e = lastx ? c - (d.nx - 1) : c + 1;
        if (A[e] != a[1]) f |= 2;
    }
    if (!lasty || d.py) {
        const long long e = lasty ? c - (long long)(d.ny - 1) * sy : c + sy;
        if (A[2 * n + e] != a[3]) f |= 2;
    }
    if (!lastz || d.pz) {
        const long long e = lastz ? c - (long long)(d.nz - 1) * sz : c + sz;
        if (A[4 * n + e] != a[5]) f |= 2;
    }
    const double ex = a[6] - kept;
    if (fabs(ex) > 1e-12 * fabs(a[6])) f |= 4;
    cx[c] = vx; cy[c] = vy; cz[c] = vz; dg[c] = a[6];
    b[c] = rhs;
    x[c] = p[q];
    if (f) atomicOr(flags, f);
}

// excess A7 - sum(couplings) of a packed cell, with rounding noise
// (relative 1e-12) treated as zero
template <class T>
__device__ __forceinline__ double excess_of(const LevelK<T> &L, int i, int j, int k, long long c) {
    const long long sy = L.nx, sz = (long long)L.nx * L.ny;
    const long long cxm = c + (wrap_m(i, L.nx) - i);
    const long long cym = c + (long long)(wrap_m(j, L.ny) - j) * sy;
    const long long czm = c + (long long)(wrap_m(k, L.nz) - k) * sz;
    const double s = (double)L.cx[c] + (double)L.cx[cxm] + (double)L.cy[c] + (double)L.cy[cym] +
                     (double)L.cz[c] + (double)L.cz[czm];
    const double d = (double)L.dg[c];
    const double e = d - s;
    return fabs(e) > 1e-12 * fabs(d) ? e : 0.0;
}

// one coarse cell from its children; the coarse -x/-y/-z couplings are
// recomputed from the fine faces of the neighbouring coarse cell
template <class Tf, class Tc>
__global__ void __launch_bounds__(256)
k_coarsen(LevelK<Tf> F, Tc *__restrict__ ccx, Tc *__restrict__ ccy, Tc *__restrict__ ccz,
          Tc *__restrict__ cdg) {
    const int nx = F.nx >> 1, ny = F.ny >> 1, nz = F.nz >> 1;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= nx || J >= ny) return;
    const long long fsy = F.nx, fsz = (long long)F.nx * F.ny;
    auto fi = [&](int i, int j, int k) { return i + fsy * j + fsz * k; };
    // coupling of coarse cell (I,J,K) to +x: fine faces at i = 2I+1
    auto cpx = [&](int I_, int J_, int K_) {
        const int i = 2 * I_ + 1, j = 2 * J_, k = 2 * K_;
        return ((double)F.cx[fi(i, j, k)] + (double)F.cx[fi(i, j + 1, k)]) +
               ((double)F.cx[fi(i, j, k + 1)] + (double)F.cx[fi(i, j + 1, k + 1)]);
    };
    auto cpy = [&](int I_, int J_, int K_) {
        const int i = 2 * I_, j = 2 * J_ + 1, k = 2 * K_;
        return ((double)F.cy[fi(i, j, k)] + (double)F.cy[fi(i + 1, j, k)]) +
               ((double)F.cy[fi(i, j, k + 1)] + (double)F.cy[fi(i + 1, j, k + 1)]);
    };
    auto cpz = [&](int I_, int J_, int K_) {
        const int i = 2 * I_, j = 2 * J_, k = 2 * K_ + 1;
        return ((double)F.cz[fi(i, j, k)] + (double)F.cz[fi(i + 1, j, k)]) +
               ((double)F.cz[fi(i, j + 1, k)] + (double)F.cz[fi(i + 1, j + 1, k)]);
    };
    const double s = 1.0 / 16.0;
    const double vx = s * cpx(I, J, K), vy = s * cpy(I, J, K), vz = s * cpz(I, J, K);
    const double wx = s * cpx(wrap_m(I, nx), J, K), wy = s * cpy(I, wrap_m(J, ny), K),
                 wz = s * cpz(I, J, wrap_m(K, nz));
    double ex = 0.0;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                ex += excess_of(F, i, j, k, fi(i, j, k));
            }
    const long long c = I + (long long)nx * (J + (long long)ny * K);
    ccx[c] = (Tc)vx; ccy[c] = (Tc)vy; ccz[c] = (Tc)vz;
    cdg[c] = (Tc)(((vx + wx) + (vy + wy)) + (vz + wz) + 0.125 * ex);
}

template <class Ti, class To>
__global__ void __launch_bounds__(256)
k_convert(const Ti *__restrict__ a, To *__restrict__ o, long long n) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) o[q] = (To)a[q];
}

// Dense inverse of the coarsest operator (nb <= 512 cells), one workgroup.
// M = diag - couplings (+ alpha * 1 1^T when the system is singular, which
// makes M SPD while leaving mean-free solutions unchanged).  Gauss-Jordan
// without pivoting is stable for SPD matrices.  work: nb*nb doubles.
template <class T>
__global__ void __launch_bounds__(1024)
k_bottom_build(LevelK<T> L, int singular, double *__restrict__ M, T *__restrict__ Minv) {
    const int nb = L.nx * L.ny * L.nz;
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += 1024) M[e] = 0.0;
    __syncthreads();
    double dsum = 0.0;
    for (int c = 0; c < nb; ++c) dsum += (double)L.dg[c];
    const double alpha = singular ? dsum / ((double)nb * (double)nb) : 0.0;
    for (int c = tid; c < nb; c += 1024) {
        const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
        const long long sy = L.nx, sz = (long long)L.nx * L.ny;
        const int ip = wrap_p(i, L.nx), jp = wrap_p(j, L.ny), kp = wrap_p(k, L.nz);
        const int im = wrap_m(i, L.nx), jm = wrap_m(j, L.ny), km = wrap_m(k, L.nz);
        double *row = M + (long long)c * nb;
        row[c] += (double)L.dg[c];
        row[c + (ip - i)] -= (double)L.cx[c];
        row[c + (im - i)] -= (double)L.cx[c + (im - i)];
        row[c + (jp - j) * sy] -= (double)L.cy[c];
        row[c + (jm - j) * sy] -= (double)L.cy[c + (jm - j) * sy];
        row[c + (kp - k) * sz] -= (double)L.cz[c];
        row[c + (km - k) * sz] -= (double)L.cz[c + (km - k) * sz];
        for (int e = 0; e < nb; ++e) row[e] += alpha;
    }
    __syncthreads();
    // in-place Gauss-Jordan inversion
    for (int p = 0; p < nb; ++p) {
        const double piv = 1.0 / M[(long long)p * nb + p];
        __syncthreads();
        // scale row p (column p becomes the pivot reciprocal)
        for (int e = tid; e < nb; e += 1024) {
            if (e != p) M[(long long)p * nb + e] *= piv;
        }
        __syncthreads();
        if (tid == 0) M[(long long)p * nb + p] = piv;
        __syncthreads();
        for (int e = tid; e < nb * nb; e += 1024) {
            const int r = e / nb, col = e % nb;
            if (r == p) continue;
            const double f = M[(long long)r * nb + p];
            if (col == p) continue;
            M[e] -= f * M[(long long)p * nb + col];
        }
        __syncthreads();
        for (int r = tid; r < nb; r += 1024) {
            if (r != p) M[(long long)r * nb + p] = -M[(long long)r * nb + p] * piv;
        }
        __syncthreads();
    }
    for (int e = tid; e < nb * nb; e += 1024) Minv[e] = (T)M[e];
}

template <class T>
__global__ void __launch_bounds__(512)
k_bottom_solve(int nb, const T *__restrict__ Minv, const T *__restrict__ b, T *__restrict__ x) {
    __shared__ T sb[512];
    for (int e = threadIdx.x; e < nb; e += 512) sb[e] = b[e];
    __syncthreads();
    for (int r = threadIdx.x; r < nb; r += 512) {
        T acc = 0;
        const T *row = Minv + (long long)r * nb;
        for (int e = 0; e < nb; ++e) acc += row[e] * sb[e];
        x[r] = acc;
    }
}

// ---------------------------------------------------------------------
// FP64 outer iteration kernels.  Scalars live on the device:
// ---------------------------------------------------------------------
enum { S_RZ = 0, S_ZMEAN, S_BETA, S_DQ, S_ALPHA, S_RR, S_BB, S_FIRST, S_N };

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
// block of 256: writes the block sum of v to part[blockIdx.x]
__device__ __forceinline__ void block_partial(double v, double *part) {
    __shared__ double sm[4];
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    __syncthreads();
}
// fixed-order sum of nb partials in a 256-thread block, result in thread 0
__device__ __forceinline__ double sum_partials(const double *part, int nb) {
    __shared__ double sm[4];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += part[b];
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (sm[0] + sm[1]) + (sm[2] + sm[3]);
    __syncthreads();
    return t;
}

struct OpK {  // FP64 fine operator for the outer iteration
    int nx, ny, nz;
    const double *cx, *cy, *cz, *dg;
};

__device__ __forceinline__ double op_apply(const OpK &L, const double *__restrict__ v, int i, int j,
                                           int k, long long c) {
    const long long sy = L.nx, sz = (long long)L.nx * L.ny;
    const int im = wrap_m(i, L.nx), ip = wrap_p(i, L.nx);
    const int jm = wrap_m(j, L.ny), jp = wrap_p(j, L.ny);
    const int km = wrap_m(k, L.nz), kp = wrap_p(k, L.nz);
    const long long cxm = c + (im - i), cxp = c + (ip - i);
    const long long cym = c + (long long)(jm - j) * sy, cyp = c + (long long)(jp - j) * sy;
    const long long czm = c + (long long)(km - k) * sz, czp = c + (long long)(kp - k) * sz;
    double s = L.cx[cxm] * v[cxm] + L.cx[c] * v[cxp];
    s += L.cy[cym] * v[cym] + L.cy[c] * v[cyp];
    s += L.cz[czm] * v[czm] + L.cz[c] * v[czp];
    return L.dg[c] * v[c] - s;
}

// r = b - M x ; partials: 0 = sum r^2, 1 = sum b^2 ; optional FP32 copy
template <class TB>
__global__ void __launch_bounds__(256)
k_outer_resid(OpK L, const double *__restrict__ b, const double *__restrict__ x,
              double *__restrict__ r, TB *__restrict__ bt, double *__restrict__ part) {
    const long long n = (long long)L.nx * L.ny * L.nz;
    double rr = 0.0, bb = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const int i = (int)(c % L.nx);
        const long long t = c / L.nx;
        const int j = (int)(t % L.ny), k = (int)(t / L.ny);
        const double v = b[c] - op_apply(L, x, i, j, k, c);
        r[c] = v;
        if (bt) bt[c] = (TB)v;
        rr += v * v;
        bb += b[c] * b[c];
    }
    block_partial(rr, part);
    block_partial(bb, part + kRedBlocks);
}

// partials: 0 = sum r z, 1 = sum z, 2 = sum r
template <class TZ>
__global__ void __launch_bounds__(256)
k_dot_rz(long long n, const double *__restrict__ r, const TZ *__restrict__ z, double *__restrict__ part) {
    double rz = 0.0, zs = 0.0, rs = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const double zv = (double)z[c], rv = r[c];
        rz += rv * zv;
        zs += zv;
        rs += rv;
    }
    block_partial(rz, part);
    block_partial(zs, part + kRedBlocks);
    block_partial(rs, part + 2 * kRedBlocks);
}

// beta step: zmean, rho = r.(z - zmean), beta = rho / rho_old
__global__ void __launch_bounds__(256)
k_scal_beta(const double *__restrict__ part, int nb, long long n, int singular,
            double *__restrict__ S) {
    const double rz = sum_partials(part, nb);
    const double zs = sum_partials(part + kRedBlocks, nb);
    const double rs = sum_partials(part + 2 * kRedBlocks, nb);
    if (threadIdx.x == 0) {
        const double zm = singular ? zs / (double)n : 0.0;
        const double rho = rz - zm * rs;
        const double first = S[S_FIRST];
        S[S_BETA] = first != 0.0 ? 0.0 : rho / S[S_RZ];
        S[S_RZ] = rho;
        S[S_ZMEAN] = zm;
        S[S_FIRST] = 0.0;
    }
}

// d_new = (z - zmean) + beta d_old ; q = M d_new ; partial sum d.q
template <class TZ>
__global__ void __launch_bounds__(256)
k_dq(OpK L, const TZ *__restrict__ z, const double *__restrict__ dold, double *__restrict__ dnew,
     double *__restrict__ q, const double *__restrict__ S, double *__restrict__ part) {
    const long long n = (long long)L.nx * L.ny * L.nz;
    const double beta = S[S_BETA], zm = S[S_ZMEAN];
    const bool use_old = beta != 0.0;
    double dq = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const int i = (int)(c % L.nx);
        const long long t = c / L.nx;
        const int j = (int)(t % L.ny), k = (int)(t / L.ny);
        auto dv = [&](long long e) {
            const double v = (double)z[e] - zm;
            return use_old ? v + beta * dold[e] : v;
        };
        const long long sy = L.nx, sz = (long long)L.nx * L.ny;
        const int im = wrap_m(i, L.nx), ip = wrap_p(i, L.nx);
        const int jm = wrap_m(j, L.ny), jp = wrap_p(j, L.ny);
        const int km = wrap_m(k, L.nz), kp = wrap_p(k, L.nz);
        const long long cxm = c + (im - i), cxp = c + (ip - i);
        const long long cym = c + (long long)(jm - j) * sy, cyp = c + (long long)(jp - j) * sy;
        const long long czm = c + (long long)(km - k) * sz, czp = c + (long long)(kp - k) * sz;
        const double d0 = dv(c);
        double s = L.cx[cxm] * dv(cxm) + L.cx[c] * dv(cxp);
        s += L.cy[cym] * dv(cym) + L.cy[c] * dv(cyp);
        s += L.cz[czm] * dv(czm) + L.cz[c] * dv(czp);
        const double qv = L.dg[c] * d0 - s;
        dnew[c] = d0;
        q[c] = qv;
        dq += d0 * qv;
    }
    block_partial(dq, part);
}

__global__ void __launch_bounds__(256)
k_scal_alpha(const double *__restrict__ part, int nb, double *__restrict__ S) {
    const double dq = sum_partials(part, nb);
    if (threadIdx.x == 0) {
        S[S_DQ] = dq;
        S[S_ALPHA] = S[S_RZ] / dq;
    }
}

// x += alpha d ; r -= alpha q ; partial sum r^2 ; optional FP32 copy of r
template <class TB>
__global__ void __launch_bounds__(256)
k_update(long long n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ d,
         const double *__restrict__ q, TB *__restrict__ bt, const double *__restrict__ S,
         double *__restrict__ part) {
    const double alpha = S[S_ALPHA];
    double rr = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        x[c] += alpha * d[c];
        const double v = r[c] - alpha * q[c];
        r[c] = v;
        if (bt) bt[c] = (TB)v;
        rr += v * v;
    }
    block_partial(rr, part);
}

// refinement step: x += z - zmean ; partial sums of z
template <class TZ>
__global__ void __launch_bounds__(256)
k_zsum(long long n, const TZ *__restrict__ z, double *__restrict__ part) {
    double zs = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) zs += (double)z[c];
    block_partial(zs, part);
}
template <class TZ>
__global__ void __launch_bounds__(256)
k_add_z(long long n, double *__restrict__ x, const TZ *__restrict__ z, const double *__restrict__ part,
        int nb, int singular) {
    __shared__ double zm_s;
    const double zs = sum_partials(part, nb);
    if (threadIdx.x == 0) zm_s = singular ? zs / (double)n : 0.0;
    __syncthreads();
    const double zm = zm_s;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256)
        x[c] += (double)z[c] - zm;
}

__global__ void __launch_bounds__(256)
k_scal_rr(const double *__restrict__ part, int nb, int with_bb, double *__restrict__ S) {
    const double rr = sum_partials(part, nb);
    double bb = 0.0;
    if (with_bb) bb = sum_partials(part + kRedBlocks, nb);
    if (threadIdx.x == 0) {
        S[S_RR] = rr;
        if (with_bb) S[S_BB] = bb;
    }
}

// write the interior iterate back into padded p, then wrap the first
// ghost layer in periodic directions
__global__ void __launch_bounds__(256)
k_unpack(PackK d, const double *__restrict__ x, double *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    const long long Py = d.X, Pz = (long long)d.X * d.Y;
    const long long q = (i + d.ng) + Py * (j + d.ng) + Pz * (k + d.ng);
    const double v = x[c];
    p[q] = v;
    if (d.px && i == d.nx - 1) p[q - (long long)(d.nx - 1) - 1] = v;
    if (d.px && i == 0) p[q + d.nx] = v;
    if (d.py && j == d.ny - 1) p[q - (long long)d.ny * Py] = v;
    if (d.py && j == 0) p[q + (long long)d.ny * Py] = v;
    if (d.pz && k == d.nz - 1) p[q - (long long)d.nz * Pz] = v;
    if (d.pz && k == 0) p[q + (long long)d.nz * Pz] = v;
}

// adjoint right-hand side: interior of a padded vector
__global__ void __launch_bounds__(256)
k_interior(PackK d, const double *__restrict__ p, double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    x[c] = p[(i + d.ng) + (long long)d.X * (j + d.ng) + (long long)d.X * d.Y * (k + d.ng)];
}

__global__ void __launch_bounds__(256)
k_fill(double *__restrict__ a, long long n, double v) {
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) a[q] = v;
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
template <class T> struct Level {
    int nx, ny, nz;
    long long n;
    T *cx, *cy, *cz, *dg, *b, *x;
    LevelK<T> k() const { return LevelK<T>{nx, ny, nz, cx, cy, cz, dg}; }
};

template <class T> struct Hier {
    std::vector<Level<T>> lv;
    std::vector<DevBuf> bufs;
    DevBuf minv, work;
    int nb = 0;
    int per[3] = {0, 0, 0};
};

struct FastState {
    Dims d;
    int built = 0;
    // FP64 fine operator (outer iteration) and outer vectors
    DevBuf op;       // cx, cy, cz, dg (4n doubles)
    DevBuf vec;      // b, x, r, d0, d1, q, z64 (7n doubles)
    DevBuf red;      // partials (4 x kRedBlocks) + scalars
    DevBuf flags;
    int singular = 1;
    const double *built_for = nullptr;
    Hier<double> h64;
    Hier<float> h32;
    DevBuf b32;      // FP32 right-hand side of the mixed V-cycle
    int mixed_built = 0, fp64_built = 0;
};

void fast_destroy(FastState *f) {
    if (!f) return;
    f->op.release(); f->vec.release(); f->red.release(); f->flags.release(); f->b32.release();
    for (auto &b : f->h64.bufs) b.release();
    for (auto &b : f->h32.bufs) b.release();
    f->h64.minv.release(); f->h64.work.release();
    f->h32.minv.release(); f->h32.work.release();
    delete f;
}

static dim3 cgrid3(int nx, int ny, int nz) {
    return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz);
}
static const dim3 kB3(64, 4, 1);
static int nblk(long long n) {
    long long b = (n + 255) / 256;
    if (b > kRedBlocks) b = kRedBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

// level sizes: coarsen while every dim is even and the level has more
// than 64 cells; the coarsest must have at most 512 cells
static int plan_levels(const Dims &d, std::vector<int> &dims) {
    int nx = d.nx, ny = d.ny, nz = d.nz;
    dims.clear();
    dims.push_back(nx); dims.push_back(ny); dims.push_back(nz);
    while ((long long)nx * ny * nz > 64 && !(nx & 1) && !(ny & 1) && !(nz & 1)) {
        nx >>= 1; ny >>= 1; nz >>= 1;
        dims.push_back(nx); dims.push_back(ny); dims.push_back(nz);
    }
    if ((long long)nx * ny * nz > 512) return PPS_ERR_BAD_SIZE;
    return PPS_OK;
}

template <class T>
static int hier_alloc(Hier<T> &H, const std::vector<int> &dims, bool own_fine, const int per[3]) {
    const int nl = (int)dims.size() / 3;
    for (int a = 0; a < 3; ++a) H.per[a] = per[a];
    H.lv.resize(nl);
    H.bufs.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Level<T> &L = H.lv[l];
        L.nx = dims[3 * l]; L.ny = dims[3 * l + 1]; L.nz = dims[3 * l + 2];
        L.n = (long long)L.nx * L.ny * L.nz;
        const bool own = l > 0 || own_fine;
        // coefficient planes (4) + b + x, or only x when the fine operator
        // and right-hand side are borrowed from the outer iteration
        const int planes = own ? 6 : 1;
        PPS_TRY(H.bufs[l].ensure(sizeof(T) * planes * L.n));
        T *base = H.bufs[l].template as<T>();
        if (own) {
            L.cx = base; L.cy = base + L.n; L.cz = base + 2 * L.n; L.dg = base + 3 * L.n;
            L.b = base + 4 * L.n; L.x = base + 5 * L.n;
        } else {
            L.cx = L.cy = L.cz = L.dg = L.b = nullptr;
            L.x = base;
        }
    }
    H.nb = (int)H.lv.back().n;
    PPS_TRY(H.minv.ensure(sizeof(T) * H.nb * H.nb));
    PPS_TRY(H.work.ensure(sizeof(double) * H.nb * H.nb));
    return PPS_OK;
}

// coarse levels + bottom inverse from an already filled fine level
template <class T>
static void hier_build(Hier<T> &H, int singular, hipStream_t s) {
    const int nl = (int)H.lv.size();
    for (int l = 1; l < nl; ++l) {
        const Level<T> &F = H.lv[l - 1];
        Level<T> &C = H.lv[l];
        k_coarsen<T, T><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(F.k(), C.cx, C.cy, C.cz, C.dg);
    }
    k_bottom_build<T><<<1, 1024, 0, s>>>(H.lv.back().k(), singular, H.work.template as<double>(),
                                         H.minv.template as<T>());
}

// One colour pass of level l.  Algorithmic bytes of a pass on n cells:
// all three coupling planes (a cell's -x/-y/-z couplings live in its
// other-colour neighbours' planes), the diagonal and right-hand side of the
// updated colour, the other colour's iterate (read) and the updated
// colour's iterate (write; plus read unless FIRST):
//   FIRST: (3 + 1/2 + 1/2 + 1/2) n S ;  else (3 + 1/2 + 1/2 + 1/2 + 1/2 + 1/2) n S
template <class T>
static void rb_pass(const Level<T> &L, int l, const T *b, int colour, bool first, Prof *prof,
                    hipStream_t s) {
    const LevelK<T> K = L.k();
    const dim3 g2((unsigned)(((L.nx + 1) / 2 + 63) / 64), (unsigned)((L.ny + 3) / 4), (unsigned)L.nz);
    if (l != 0) {
        if (first) k_rb<T, T, true, false><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
        else k_rb<T, T, false, false><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
        return;
    }
    const double bytes = (double)L.n * sizeof(T) * (first ? 4.5 : 5.5);
    hipEvent_t *e = prof ? prof->begin(bytes) : nullptr;
    if (e) (void)hipEventRecord(e[0], s);
    if (first) k_rb<T, T, true, true><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
    else k_rb<T, T, false, true><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
    if (e) (void)hipEventRecord(e[1], s);
}

// one V-cycle: x_0 = V(b_0).  fineb: the fine right-hand side (the FP64
// residual in FP64 mode, its FP32 copy in mixed mode).
template <class T>
static void vcycle(Hier<T> &H, const T *fineb, int nu_pre, int nu_post, Prof *prof, hipStream_t s) {
    const int nl = (int)H.lv.size();
    for (int l = 0; l < nl - 1; ++l) {
        Level<T> &L = H.lv[l];
        const LevelK<T> K = L.k();
        const T *b = l == 0 ? fineb : L.b;
        rb_pass(L, l, b, 0, true, prof, s);
        rb_pass(L, l, b, 1, false, prof, s);
        for (int q = 1; q < nu_pre; ++q) {
            rb_pass(L, l, b, 0, false, prof, s);
            rb_pass(L, l, b, 1, false, prof, s);
        }
        Level<T> &C = H.lv[l + 1];
        k_resid_restrict<T, T><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(K, b, L.x, C.b);
    }
    {
        Level<T> &Bt = H.lv.back();
        const T *b = nl == 1 ? fineb : Bt.b;
        k_bottom_solve<T><<<1, 512, 0, s>>>(H.nb, H.minv.template as<T>(), b, Bt.x);
    }
    for (int l = nl - 2; l >= 0; --l) {
        Level<T> &L = H.lv[l];
        const Level<T> &C = H.lv[l + 1];
        const T *b = l == 0 ? fineb : L.b;
        k_prolong_add<T><<<cgrid3(L.nx, L.ny, L.nz), kB3, 0, s>>>(L.nx, L.ny, L.nz, H.per[0], H.per[1],
                                                                  H.per[2], L.x, C.x);
        for (int q = 0; q < nu_post; ++q) {
            rb_pass(L, l, b, 1, false, prof, s);
            rb_pass(L, l, b, 0, false, prof, s);
        }
    }
}

static PackK packk(const Dims &d) {
    PackK k;
    k.nx = d.nx; k.ny = d.ny; k.nz = d.nz; k.ng = d.ng; k.X = d.X(); k.Y = d.Y();
    k.px = d.per[0]; k.py = d.per[1]; k.pz = d.per[2];
    return k;
}

static int fast_ensure(pps_handle *h) {
    if (!h->fast) {
        FastState *f = new FastState();
        f->d = dims_of(h->grid);
        h->fast = f;
    }
    FastState *f = h->fast;
    const long long n = f->d.cells();
    PPS_TRY(f->op.ensure(sizeof(double) * 4 * n));
    PPS_TRY(f->vec.ensure(sizeof(double) * 7 * n));
    PPS_TRY(f->red.ensure(sizeof(double) * (4 * kRedBlocks + 64)));
    PPS_TRY(f->flags.ensure(64));
    if (!h->pinned) PPS_HIP(hipHostMalloc((void **)&h->pinned, 16 * sizeof(double), 0));
    return PPS_OK;
}

struct Vecs {
    double *b, *x, *r, *d[2], *q, *z;
};
static Vecs vecs(FastState *f) {
    const long long n = f->d.cells();
    double *v = f->vec.as<double>();
    Vecs V;
    V.b = v; V.x = v + n; V.r = v + 2 * n; V.d[0] = v + 3 * n; V.d[1] = v + 4 * n;
    V.q = v + 5 * n; V.z = v + 6 * n;
    return V;
}
static OpK opk(FastState *f) {
    const long long n = f->d.cells();
    double *o = f->op.as<double>();
    return OpK{f->d.nx, f->d.ny, f->d.nz, o, o + n, o + 2 * n, o + 3 * n};
}

// pack A, detect symmetry/singularity, build the hierarchy for `mixed`
static int fast_setup(pps_handle *h, const double *dA, const double *dp, bool mixed, hipStream_t s) {
    FastState *f = h->fast;
    const Dims &d = f->d;
    const long long n = d.cells();
    const OpK O = opk(f);
    Vecs V = vecs(f);
    int *flags = f->flags.as<int>();
    PPS_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
    k_pack<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), dA, dp, (double *)O.cx, (double *)O.cy,
                                                   (double *)O.cz, (double *)O.dg, V.b, V.x, flags);
    PPS_HIP(hipGetLastError());
    int fl = 0;
    PPS_HIP(hipMemcpyAsync(&fl, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    PPS_HIP(hipStreamSynchronize(s));
    if (fl & 1) return PPS_ERR_NAN_OR_TINY_DIAG;
    if (fl & 2) return PPS_ERR_NONSYMMETRIC;
    f->singular = (fl & 4) ? 0 : 1;

    std::vector<int> dims;
    PPS_TRY(plan_levels(d, dims));
    if (mixed) {
        Hier<float> &H = f->h32;
        if (!f->mixed_built) { PPS_TRY(hier_alloc(H, dims, true, d.per)); f->mixed_built = 1; }
        Level<float> &L0 = H.lv[0];
        const unsigned nbk = (unsigned)nblk(4 * n);
        // the four FP64 planes are contiguous, as are the FP32 ones
        k_convert<double, float><<<nbk, 256, 0, s>>>(O.cx, L0.cx, 4 * n);
        hier_build(H, f->singular, s);
        PPS_TRY(f->b32.ensure(sizeof(float) * n));
    } else {
        Hier<double> &H = f->h64;
        if (!f->fp64_built) { PPS_TRY(hier_alloc(H, dims, false, d.per)); f->fp64_built = 1; }
        Level<double> &L0 = H.lv[0];
        L0.cx = (double *)O.cx; L0.cy = (double *)O.cy; L0.cz = (double *)O.cz; L0.dg = (double *)O.dg;
        L0.b = nullptr;
        L0.x = V.z;
        hier_build(H, f->singular, s);
    }
    PPS_HIP(hipGetLastError());
    f->built_for = dA;
    return PPS_OK;
}

// The outer iteration on (b, x): PCG or refinement with the V-cycle of
// hierarchy H.  Convergence: sqrt(rr/N) < tol_abs (norm 2).  Returns the
// iteration count in *it (maxit+1 if not converged).
template <class T>
static int outer_iterate(pps_handle *h, Hier<T> &H, bool mixed, double tol_abs, int maxit,
                         int *it, double *last, bool record, hipStream_t s) {
    FastState *f = h->fast;
    const long long n = f->d.cells();
    const OpK O = opk(f);
    Vecs V = vecs(f);
    double *part = f->red.as<double>();
    double *S = part + 4 * kRedBlocks;
    const int nb = nblk(n);
    float *b32 = mixed ? f->b32.as<float>() : nullptr;
    const double N = (double)n;
    const bool cg = h->opt.outer == PPS_OUTER_CG;
    const int nu1 = h->opt.nu_pre < 1 ? 1 : h->opt.nu_pre, nu2 = h->opt.nu_post;

    // r = b - M x
    k_outer_resid<float><<<nb, 256, 0, s>>>(O, V.b, V.x, V.r, b32, part);
    k_scal_rr<<<1, 256, 0, s>>>(part, nb, 1, S);
    k_fill<<<1, 64, 0, s>>>(S + S_FIRST, 1, 1.0);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipMemcpyAsync(h->pinned, S + S_RR, sizeof(double), hipMemcpyDeviceToHost, s));
    PPS_HIP(hipStreamSynchronize(s));
    double t2 = std::sqrt(h->pinned[0] / N);
    *last = t2;
    int k = 1, cur = 0;
    if (t2 == 0.0) { *it = 1; return PPS_OK; }   // exact solution: nothing to do
    for (k = 1; k <= maxit; ++k) {
        // z = V(r)
        T *z = H.lv[0].x;
        vcycle<T>(H, mixed ? (const T *)(const void *)b32 : (const T *)(const void *)V.r, nu1, nu2,
                  &h->prof, s);
        if (cg) {
            k_dot_rz<T><<<nb, 256, 0, s>>>(n, V.r, z, part);
            k_scal_beta<<<1, 256, 0, s>>>(part, nb, n, f->singular, S);
            k_dq<T><<<nb, 256, 0, s>>>(O, z, V.d[cur], V.d[cur ^ 1], V.q, S, part);
            k_scal_alpha<<<1, 256, 0, s>>>(part, nb, S);
            k_update<float><<<nb, 256, 0, s>>>(n, V.x, V.r, V.d[cur ^ 1], V.q, b32, S, part);
            k_scal_rr<<<1, 256, 0, s>>>(part, nb, 0, S);
            cur ^= 1;
        } else {
            k_zsum<T><<<nb, 256, 0, s>>>(n, z, part);
            k_add_z<T><<<nb, 256, 0, s>>>(n, V.x, z, part, nb, f->singular);
            k_outer_resid<float><<<nb, 256, 0, s>>>(O, V.b, V.x, V.r, b32, part);
            k_scal_rr<<<1, 256, 0, s>>>(part, nb, 0, S);
        }
        PPS_HIP(hipGetLastError());
        PPS_HIP(hipMemcpyAsync(h->pinned, S + S_RR, sizeof(double), hipMemcpyDeviceToHost, s));
        PPS_HIP(hipStreamSynchronize(s));
        t2 = std::sqrt(h->pinned[0] / N);
        if (record) h->hist.push_back(t2);
        *last = t2;
        if (!(t2 == t2) || t2 > 1e300) { *it = k; return PPS_ERR_DIVERGED; }
        if (t2 < tol_abs) break;
    }
    *it = k;
    return PPS_OK;
}

int fast_solve(pps_handle *h, const double *dA, double *dp, double maxError, int maxit, int norm,
               int *it, double *tres2, hipStream_t s) {
    PPS_TRY(fast_ensure(h));
    FastState *f = h->fast;
    const bool mixed = h->opt.method == PPS_MG_MIXED;
    h->hist.clear();
    PPS_HIP(hipEventRecord(h->ev[0], s));
    PPS_TRY(fast_setup(h, dA, dp, mixed, s));
    PPS_HIP(hipEventRecord(h->ev[1], s));
    const Dims &d = f->d;
    const double N = (double)d.cells();
    int total = 0, k = 0;
    double t2 = 0.0, last = 0.0;
    int rc = PPS_OK;
    // iterate on the recursive residual, then confirm with the reference's
    // residual on the caller's A; restart from the true residual if the
    // recursion drifted
    for (int attempt = 0; attempt < 4 && total < maxit; ++attempt) {
        if (mixed) rc = outer_iterate<float>(h, f->h32, true, maxError, maxit - total, &k, &last, true, s);
        else rc = outer_iterate<double>(h, f->h64, false, maxError, maxit - total, &k, &last, true, s);
        if (rc != PPS_OK && rc != PPS_ERR_NOT_CONVERGED) break;
        total += (k > maxit - total) ? maxit - total : k;
        k_unpack<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), vecs(f).x, dp);
        PPS_HIP(hipGetLastError());
        double res = 0.0;
        PPS_TRY(compute_residual_norm(h, d, dA, dp, norm, &res, s));
        t2 = res / N;
        if (norm == 2) t2 = std::sqrt(t2);
        if (t2 < maxError) break;
        if (k > maxit) break;
    }
    PPS_HIP(hipEventRecord(h->ev[2], s));
    PPS_HIP(hipEventSynchronize(h->ev[2]));
    float ms_setup = 0.f, ms_all = 0.f;
    PPS_HIP(hipEventElapsedTime(&ms_setup, h->ev[0], h->ev[1]));
    PPS_HIP(hipEventElapsedTime(&ms_all, h->ev[0], h->ev[2]));
    h->setup_ms = ms_setup;
    h->solve_ms = ms_all;
    PPS_TRY(h->prof.collect());
    if (rc != PPS_OK) return rc;
    *tres2 = t2;
    *it = (t2 < maxError) ? (total < 1 ? 1 : total) : maxit + 1;
    return PPS_OK;
}

int fast_adjoint(pps_handle *h, const double *dA, const double *dg, double *dlam, double rtol,
                 int maxit, int *it, hipStream_t s) {
    PPS_TRY(fast_ensure(h));
    FastState *f = h->fast;
    const bool mixed = h->opt.method == PPS_MG_MIXED;
    const Dims &d = f->d;
    const long long n = d.cells();
    // rebuild only if the operator changed; the fold of wall ghosts into
    // the right-hand side is irrelevant here (the adjoint ghosts are zero)
    if (f->built_for != dA || (mixed ? !f->mixed_built : !f->fp64_built)) {
        PPS_TRY(fast_setup(h, dA, dlam, mixed, s));
    }
    Vecs V = vecs(f);
    k_interior<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), dg, V.b);
    // project the right-hand side mean-free when the operator is singular
    if (f->singular) {
        double *part = f->red.as<double>();
        const int nb = nblk(n);
        k_zsum<double><<<nb, 256, 0, s>>>(n, V.b, part);
        // b -= mean(b): reuse k_add_z with z = b and x = b would double it;
        // use a negated add on a zero vector instead
        k_fill<<<nb, 256, 0, s>>>(V.x, n, 0.0);
        k_add_z<double><<<nb, 256, 0, s>>>(n, V.x, V.b, part, nb, 1);
        PPS_HIP(hipMemcpyAsync(V.b, V.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    k_fill<<<nblk(n), 256, 0, s>>>(V.x, n, 0.0);
    // absolute tolerance from the right-hand side norm
    double *part = f->red.as<double>();
    double *S = part + 4 * kRedBlocks;
    const int nb = nblk(n);
    const OpK O = opk(f);
    k_outer_resid<float><<<nb, 256, 0, s>>>(O, V.b, V.x, V.r, nullptr, part);
    k_scal_rr<<<1, 256, 0, s>>>(part, nb, 1, S);
    PPS_HIP(hipMemcpyAsync(h->pinned, S + S_BB, sizeof(double), hipMemcpyDeviceToHost, s));
    PPS_HIP(hipStreamSynchronize(s));
    const double tol = rtol * std::sqrt(h->pinned[0] / (double)n);
    int k = 0;
    double last = 0.0;
    int rc;
    if (mixed) rc = outer_iterate<float>(h, f->h32, true, tol, maxit, &k, &last, false, s);
    else rc = outer_iterate<double>(h, f->h64, false, tol, maxit, &k, &last, false, s);
    if (rc != PPS_OK) return rc;
    // padded output, ghosts: wrapped (periodic) or zero (walls)
    PPS_HIP(hipMemsetAsync(dlam, 0, sizeof(double) * d.padded(), s));
    k_unpack<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), V.x, dlam);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    *it = k;
    PPS_TRY(h->prof.collect());
    return last < tol ? PPS_OK : PPS_ERR_NOT_CONVERGED;
}

}  // namespace pps

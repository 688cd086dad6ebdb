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
// residual, prolongation its transpose (each child gets its parent's
// value), which keeps the V-cycle symmetric for CG.  The coarsest level
// (<= 512 cells) is solved with a dense inverse built once per solve.
//
// Precision.  PPS_MG_FP64: everything FP64.  PPS_MG_MIXED: the V-cycle
// runs on FP32 copies of the operator with FP32 vectors and accumulation;
// the outer CG (residual, update, dot products, mean projection) is FP64.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "pps_internal.h"

#include <hip/hip_ext.h>
#include "pps_row.h"
#include "pps_xcd.h"

namespace pps {

// ---------------------------------------------------------------------
// kernel-side level view
// ---------------------------------------------------------------------
template <class T> struct LevelK {
    int nx, ny, nz;
    const T *cx, *cy, *cz, *dg;
};

// sum over the 64 lanes of a wavefront (result in lane 0)
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

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

// Diagonal of a row whose excess A7 - sum(couplings) is zero (every row
// of a singular system): the sum of its six couplings, in one fixed order
// shared by every kernel of the V-cycle.
template <class T>
__device__ __forceinline__ T diag_sum(T cxm, T cx, T cym, T cy, T czm, T cz) {
    return ((cxm + cx) + (cym + cy)) + (czm + cz);
}

// Diagonal of cell c: stored (DGI false) or the coupling sum (DGI true).
template <class T, bool DGI>
__device__ __forceinline__ T diag_of(const LevelK<T> &L, int i, int j, int k, long long c) {
    if (!DGI) return L.dg[c];
    const long long sy = L.nx, sz = (long long)L.nx * L.ny;
    const long long cxm = c + (wrap_m(i, L.nx) - i);
    const long long cym = c + (long long)(wrap_m(j, L.ny) - j) * sy;
    const long long czm = c + (long long)(wrap_m(k, L.nz) - k) * sz;
    return diag_sum(L.cx[cxm], L.cx[c], L.cy[cym], L.cy[c], L.cz[czm], L.cz[c]);
}

// ---------------------------------------------------------------------
// red-black Gauss-Seidel colour pass.  colour 0 = (i+j+k) even.  FIRST:
// the iterate is zero before this pass (no read of x).
// ---------------------------------------------------------------------
// FINE does not change the code; it gives the finest level's launches
// their own symbol so profilers can separate them from coarse levels.
template <class T, class B, bool FIRST, bool FINE, bool DGI>
__global__ void __launch_bounds__(256)
k_rb(LevelK<T> L, const B *__restrict__ b, T *__restrict__ x, int colour) {
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = 2 * t + ((j + k + colour) & 1);
    if (j >= L.ny) return;
    if (FIRST) {
        // the other cell of the pair starts at zero: on a wall axis of odd
        // size the wrapped neighbour across the wall (coupling 0) has the
        // cell's own colour, and the next pass reads it (times 0) before
        // writing it, so it must hold a finite value
        const int io = 2 * t + ((j + k + colour + 1) & 1);
        if (io < L.nx) x[io + (long long)L.nx * (j + (long long)L.ny * k)] = (T)0;
    }
    if (i >= L.nx) return;
    const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
    const T d = diag_of<T, DGI>(L, i, j, k, c);
    if (FIRST) {
        x[c] = (T)b[c] / d;
    } else {
        x[c] = ((T)b[c] + nb_sum(L, x, i, j, k, c)) / d;
    }
}

// ---------------------------------------------------------------------
// One whole red-black Gauss-Seidel sweep in one pass: xout = sweep(xin).
//
// Cells are handled in x-pairs (2m, 2m+1); in every plane each pair holds
// exactly one cell of each colour.  A block owns full rows of `rb` rows
// (thread rows 0 and rb-1 are a ring that only computes first-colour
// values for the neighbours of the rb-2 output rows) and marches through a
// chunk of planes.  At plane step s it
//   R: updates the first colour of plane s (all rows, ring included) from
//      the second-colour xin of planes s-1, s, s+1 staged in LDS, staging
//      the results in LDS,
//   K: updates the second colour of plane s-1 (output rows) from the
//      first-colour values of planes s-2, s-1 (LDS) and s, and stores both
//      cells of the pair at plane s-1,
//   S: stages the second-colour xin of plane s+2.
// Every first-colour value uses only old second-colour values and every
// second-colour value only new first-colour values, so the result equals
// the two colour passes of k_rb.  Only the second colour's xin is read (the
// first colour's old value is overwritten unread).
// PF: loads run ahead of their use: step s requests the operator of plane
// s+1 and the xin of plane s+3, which steps s+1 and s+2 consume.  The
// operator registers alternate between two sets by step parity (the loop
// is unrolled by two), so no register copy forces a loaded value to arrive
// within the step that requested it, and every load is unconditional and
// in bounds (past the chunk's end the last needed plane is loaded again;
// threads off the ring load their own row twice), so the waits the
// compiler places stay one step behind the requests.  Without PF a step
// loads the operator of plane s and the xin of plane s+2 itself.  The
// per-plane barrier orders LDS only.  Redundant work: the ring rows and one extra
// first-colour plane at each end of a chunk.
// c1: first colour ((x+y+z)&1 of the cells updated first).  FIRST: xin is
// identically zero (not read).  PROL: the coarse correction ec is added to
// every xin value before it is staged (xin + P ec, P the piecewise-constant
// prolongation: both cells of a pair share the parent m), so the sweep
// that follows a coarse-grid correction needs no separate prolongation
// pass.  PX pairs per thread.  DOT: the block's sums over its output
// cells of b x, x and b (x the new iterate) go to part[block],
// part[kRedBlocks + block] and part[2 kRedBlocks + block] (the block must
// be whole wavefronts).
// LOP (FP32, two pairs per thread, 768 threads): the couplings cx, cy, cz
// do not pass through registers.  Step s starts direct global-to-LDS loads
// of plane s+1's couplings into the LDS buffer that plane s-1 used, reads
// plane s's from the other buffer, and waits for its loads only before the
// step's barrier; b runs one plane ahead in registers.  The couplings of
// the next plane thus stream in while a plane computes, with no VGPRs held
// by the loads in flight (the register form without LOP can prefetch only
// in the first sweep of a cycle, which reads no xin).
// LDS: X[4][rb+2][np] (second-colour xin, plane k in slot k & 3),
// R[2][rb][np] (first colour new), and under LOP two static buffers of
// 3 rb + 1 coupling rows.
// ---------------------------------------------------------------------
template <int I> struct Ic { static constexpr int value = I; };

// values per LDS coupling buffer of the LOP sweep (2 x 50 KiB), and its
// threads per workgroup (3 wavefronts per SIMD leave 168 VGPRs a thread;
// 1024 threads leave 128, and the LOP form spills there)
constexpr int kLopCap = 12800;
constexpr int kLopThreads = 768;

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Writes the block sum of v to part[block linear id].  Any block of at most
// 1024 threads whose x extent is a multiple of 64 (each wavefront is 64
// consecutive linear thread ids); wavefront sums are added in a fixed order.
__device__ __forceinline__ void block_partial(double v, double *part) {
    __shared__ double sm[16];
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int nw = (blockDim.x * blockDim.y * blockDim.z) >> 6;
    v = wsum(v);
    if ((t & 63) == 0) sm[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double a = 0.0;
        for (int w = 0; w < nw; ++w) a += sm[w];
        part[b] = a;
    }
    __syncthreads();
}

template <class T> struct PairOp {   // operator of one pair at one plane
    T cxa, cxb, cxw;                  // +x couplings of 2m, 2m+1; of 2m-1
    T cya, cyb, csa, csb;             // +y couplings of the pair; of row j-1
    T cza, czb, dga, dgb, ba, bb;     // +z couplings, diagonals, rhs
};
template <class T> struct CellOp {   // operator of one cell
    T cx, cxm, cy, cym, cz, dg, b;
};
template <class T>
__device__ __forceinline__ CellOp<T> pick(const PairOp<T> &P, int e) {
    CellOp<T> c;
    c.cx = e ? P.cxb : P.cxa;
    c.cxm = e ? P.cxa : P.cxw;
    c.cy = e ? P.cyb : P.cya;
    c.cym = e ? P.csb : P.csa;
    c.cz = e ? P.czb : P.cza;
    c.dg = e ? P.dgb : P.dga;
    c.b = e ? P.bb : P.ba;
    return c;
}

// FINE does not change the code; it gives the finest level's launches
// their own symbol so profilers can separate them from coarse levels.
// DGI: the diagonal is the coupling sum (not read).
template <class T, bool FIRST, int PX, bool FINE, bool DGI, bool PROL, bool PF, bool DOT,
          bool LOP = false>
__global__ void __launch_bounds__(LOP ? kLopThreads : 1024)
k_rb_sweep(LevelK<T> L, const T *__restrict__ b, const T *__restrict__ xin, T *__restrict__ xout,
           int c1, int zc, int oy, int ytiles, const T *__restrict__ ec, double *__restrict__ part,
           const double *__restrict__ skip) {
    // a first sweep queued ahead of the outer iteration's convergence test
    // does nothing once that test has stopped the iteration
    if (FIRST && skip && *skip != 0.0) return;
    extern __shared__ unsigned char smem_raw[];
    T *sm = reinterpret_cast<T *>(smem_raw);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int np = nx >> 1;                 // pairs per row
    const int tpr = blockDim.x;             // threads per row (np / PX)
    const int rb = blockDim.y;              // thread rows (oy + 2)
    const int jj = threadIdx.y;
    const long long sy = nx, sz = (long long)nx * ny;
    T *Xs = sm;                             // [4][rb+2][np]
    T *Rs = sm + 4 * (rb + 2) * np;         // [2][rb][np]
    auto X = [&](int k, int row, int m) -> T & { return Xs[((k & 3) * (rb + 2) + row) * np + m]; };
    auto R = [&](int k, int row, int m) -> T & { return Rs[((k & 1) * rb + row) * np + m]; };
    // LOP: the couplings of planes s and s+1 in two LDS buffers, filled by
    // direct global-to-LDS loads so that the next plane's couplings stream
    // in while this plane computes (no registers held by loads in flight).
    // Rows of nx values: cx in rows 0..rb-1, cz in rb..2rb-1, cy in
    // 2rb..3rb (row 2rb: the row below thread row 0).  Two distinct arrays,
    // so the compiler can tell reads of one from loads into the other.
    __shared__ T lopA[LOP ? kLopCap : 1];
    __shared__ T lopB[LOP ? kLopCap : 1];
    // FP32: tiles numbered XCD by XCD (xcd_order), so that the tiles that
    // share ring rows run on one XCD and find them in its L2 (256^3: fetch
    // 474.7 -> 400.0 MB, 82.8 -> 78.9 us a sweep).  The FP64 sweep fetches
    // less the same way (1,212 -> 854 MB) but runs slower (196 -> 213 us),
    // so it keeps the launch order.
    const int bid = sizeof(T) == 4 ? xcd_order(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int tile = bid % ytiles, chunk = bid / ytiles;
    const int y0 = tile * oy;
    const int k0 = chunk * zc;
    const int k1 = min(k0 + zc, nz);
    auto wrapz = [&](int k) { k %= nz; return k < 0 ? k + nz : k; };
    auto wrapy = [&](int j) { j %= ny; return j < 0 ? j + ny : j; };
    const int j = wrapy(y0 - 1 + jj);
    const int jm = wrapy(j - 1);
    const long long dym = (long long)(jm - j) * sy;
    const bool out_row = jj >= 1 && jj <= oy && (y0 - 1 + jj) < ny;
    const bool ring_lo = jj == 0, ring_hi = jj == rb - 1;
    // the row of a thread's second xin load: the staged row beyond the ring
    // for ring threads, its own row (a duplicate load) otherwise
    const int jl = ring_lo ? wrapy(j - 1) : (ring_hi ? wrapy(j + 1) : j);
    auto col = [&](int jrow, int k, int m) { return 2LL * m + sy * jrow + sz * k; };
    // one row of nx values from src into LDS row lrow of buf: each lane of a
    // wavefront moves 4 consecutive values (tpr is a multiple of 64, so a
    // wavefront lies in one thread row and its LDS base is uniform)
    static_assert(!LOP || PX == 2, "the LDS-staged sweep runs two pairs a thread");
    auto dma_row = [&](T *buf, int lrow, const T *src) {
        const int t0 = threadIdx.x & ~63;
        __builtin_amdgcn_global_load_lds(
            (const void *)(src + 4 * threadIdx.x),
            (__attribute__((address_space(3))) void *)(buf + lrow * nx + 4 * t0), 16, 0, 0);
    };
    auto dma_op = [&](T *buf, int k) {
        const long long pk = sz * wrapz(k);
        dma_row(buf, jj, L.cx + pk + sy * j);
        dma_row(buf, rb + jj, L.cz + pk + sy * j);
        dma_row(buf, 2 * rb + 1 + jj, L.cy + pk + sy * j);
        if (jj == 0) dma_row(buf, 2 * rb, L.cy + pk + sy * jm);
    };
    auto lds_op = [&](const T *buf, int m, PairOp<T> &P) {
        const int c = 2 * m, cw = m == 0 ? nx - 1 : c - 1;
        const T *rx = buf + jj * nx, *rz = buf + (rb + jj) * nx, *ry = buf + (2 * rb + 1 + jj) * nx;
        P.cxa = rx[c]; P.cxb = rx[c + 1]; P.cxw = rx[cw];
        P.cya = ry[c]; P.cyb = ry[c + 1];
        P.csa = ry[c - nx]; P.csb = ry[c - nx + 1];
        P.cza = rz[c]; P.czb = rz[c + 1];
    };

    auto load_op = [&](int k, int m, PairOp<T> &P) {
        const long long c = col(j, wrapz(k), m);
        const long long cw = m == 0 ? c + (nx - 1) : c - 1;
        P.cxa = L.cx[c]; P.cxb = L.cx[c + 1]; P.cxw = L.cx[cw];
        P.cya = L.cy[c]; P.cyb = L.cy[c + 1];
        P.csa = L.cy[c + dym]; P.csb = L.cy[c + dym + 1];
        P.cza = L.cz[c]; P.czb = L.cz[c + 1];
        if (!DGI) { P.dga = L.dg[c]; P.dgb = L.dg[c + 1]; }
        P.ba = b[c]; P.bb = b[c + 1];
    };
    // parent of pair m of row jrow, plane k (wrapped) in the coarse grid
    auto par = [&](int jrow, int k, int m) {
        return (long long)m + (long long)(nx >> 1) * ((jrow >> 1) + (long long)(ny >> 1) * (k >> 1));
    };
    // second-colour xin of row jrow, plane q; its coarse correction
    auto ldx = [&](int jrow, int q, int m) -> T {
        // LOP: non-temporal loads for the streams read once per sweep (the
        // old iterate here, b below), so that they do not push the ring
        // rows out of the caches (256^3: 85.7 -> 81.1 us a sweep; the FP64
        // sweep, whose tiles re-read xin rows, is slower with them)
        if constexpr (LOP)
            return __builtin_nontemporal_load(xin + col(jrow, wrapz(q), m) + ((c1 + jrow + q + 1) & 1));
        else
            return xin[col(jrow, wrapz(q), m) + ((c1 + jrow + q + 1) & 1)];
    };
    auto lde = [&](int jrow, int q, int m) -> T { return ec[par(jrow, wrapz(q), m)]; };
    auto stage = [&](int k, int m, T xv, T ev) {
        X(k, jj + 1, m) = xv;
        if (ring_lo) X(k, 0, m) = ev;
        if (ring_hi) X(k, rb + 1, m) = ev;
    };

    // per pair: operator of planes s and s+1 (op[P], op[1-P] at a step of
    // parity P), the second-colour cell's operator of plane s-1 (kop: its
    // -z coupling is the first colour's at s), the first colour's +z
    // coupling of planes s-1 and s-2 (zr1, zr2: the second colour's -z
    // coupling one plane up), first colour new at s-2 .. s, and the xin of
    // own and extra row of plane s+2 (set P) and s+3 (set 1-P), with their
    // coarse corrections under PROL.
    PairOp<T> op[2][PX];
    CellOp<T> kop[PX];
    T zr1[PX], zr2[PX], r_m2[PX], r_m1[PX], r_0[PX];
    // xin run-ahead by one more plane (PF without LOP: LOP waits for all of
    // a step's loads at its end anyway)
    constexpr bool XRA = PF && !LOP;
    T xr[2][PX], er[2][PX], xq[2][PX], eq[2][PX];
    T bf1[PX];                       // DOT: first colour's b at plane s-1
    T bq[2][PX][2];                  // LOP: b of the pair, planes s and s+1
    double sbx = 0.0, sx = 0.0, sb = 0.0;
    const int s0 = k0 - 1;
    const int klast = k1 + 1;   // last xin plane a step reads

    // prologue: operator of plane s0, the +z couplings of s0-1; stage the
    // xin of planes s0-1 .. s0+1, request s0+2
#pragma unroll
    for (int p = 0; p < PX; ++p) {
        const int m = threadIdx.x + p * tpr;
        if (LOP) {
            const long long c = col(j, wrapz(s0), m);
            bq[0][p][0] = b[c];
            bq[0][p][1] = b[c + 1];
        } else if (PF) {
            load_op(s0, m, op[0][p]);
        }
        if (!FIRST) {
            for (int q = s0 - 1; q <= s0 + 1; ++q) {
                const T xv = ldx(j, q, m) + (PROL ? lde(j, q, m) : (T)0);
                const T ev = ldx(jl, q, m) + (PROL ? lde(jl, q, m) : (T)0);
                stage(q, m, xv, ev);
            }
            if (XRA) {
                xr[0][p] = ldx(j, s0 + 2, m);
                er[0][p] = ldx(jl, s0 + 2, m);
                if (PROL) {
                    xq[0][p] = lde(j, s0 + 2, m);
                    eq[0][p] = lde(jl, s0 + 2, m);
                }
            }
        }
        const int e1p = (c1 + j + s0) & 1;   // first colour at s0
        const long long cz2 = col(j, wrapz(s0 - 1), m);
        kop[p].cz = L.cz[cz2 + e1p];
        zr1[p] = L.cz[cz2 + (1 - e1p)];
        zr2[p] = (T)0;
        r_m2[p] = r_m1[p] = r_0[p] = (T)0;
        bf1[p] = (T)0;
    }
    if constexpr (LOP) {
        dma_op(lopA, s0);
        __builtin_amdgcn_s_waitcnt(0);
    }
    lds_barrier();

    auto step = [&](auto PC, int s) {
        // PF: this step uses set PU and requests into set PN for later
        // steps; else it loads what it uses into set 0
        constexpr int PU = PF ? decltype(PC)::value : 0;
        constexpr int PN = PF ? 1 - PU : 0;
        const int e1 = (c1 + j + s) & 1;   // first-colour offset at plane s
        // requests: operator of plane s+1 and xin of plane s+3 (PF), or
        // operator of plane s and xin of plane s+2
        const int kn = PF ? min(s + 1, k1) : s;
        const int qn = min(XRA ? s + 3 : s + 2, klast);
        constexpr int XU = XRA ? PU : 0, XN = XRA ? PN : 0;
        if constexpr (LOP) dma_op(PN ? lopB : lopA, kn);
#pragma unroll
        for (int p = 0; p < PX; ++p) {
            const int m = threadIdx.x + p * tpr;
            if constexpr (LOP) {
                const long long c = col(j, wrapz(kn), m);
                bq[PN][p][0] = __builtin_nontemporal_load(b + c);
                bq[PN][p][1] = __builtin_nontemporal_load(b + c + 1);
            } else {
                load_op(kn, m, op[PN][p]);
            }
            if (!FIRST) {
                xr[XN][p] = ldx(j, qn, m);
                er[XN][p] = ldx(jl, qn, m);
                if (PROL) {
                    xq[XN][p] = lde(j, qn, m);
                    eq[XN][p] = lde(jl, qn, m);
                }
            }
        }
        // R: first colour at plane s
        T zrs[PX], bft[PX];
#pragma unroll
        for (int p = 0; p < PX; ++p) {
            const int m = threadIdx.x + p * tpr;
            if constexpr (LOP) {
                lds_op(PU ? lopB : lopA, m, op[PU][p]);
                op[PU][p].ba = bq[PU][p][0];
                op[PU][p].bb = bq[PU][p][1];
            }
            const CellOp<T> o = pick(op[PU][p], e1);
            const T zm = kop[p].cz;
            zrs[p] = o.cz;
            if (DOT) {
                bft[p] = bf1[p];
                bf1[p] = o.b;
            }
            T v = o.b;
            if (!FIRST) {
                const int mm = m == 0 ? np - 1 : m - 1, mp = m == np - 1 ? 0 : m + 1;
                const T xw = X(s, jj + 1, e1 == 0 ? mm : m);
                const T xe = X(s, jj + 1, e1 == 0 ? m : mp);
                v += o.cxm * xw + o.cx * xe;
                v += o.cym * X(s, jj, m) + o.cy * X(s, jj + 2, m);
                v += zm * X(s - 1, jj + 1, m) + o.cz * X(s + 1, jj + 1, m);
            }
            v /= DGI ? diag_sum(o.cxm, o.cx, o.cym, o.cy, zm, o.cz) : o.dg;
            R(s, jj, m) = v;
            r_m2[p] = r_m1[p];
            r_m1[p] = r_0[p];
            r_0[p] = v;
        }
        // K: second colour at plane s-1 (offset e1 there), output
        const int t = s - 1;
        if (out_row && t >= k0 && t < k1) {
            const int e2 = e1;
#pragma unroll
            for (int p = 0; p < PX; ++p) {
                const int m = threadIdx.x + p * tpr;
                const int mm = m == 0 ? np - 1 : m - 1, mp = m == np - 1 ? 0 : m + 1;
                const CellOp<T> &o = kop[p];
                const T zm = zr2[p];
                const T d = DGI ? diag_sum(o.cxm, o.cx, o.cym, o.cy, zm, o.cz) : o.dg;
                const T xw = e2 == 0 ? R(t, jj, mm) : r_m1[p];
                const T xe = e2 == 0 ? r_m1[p] : R(t, jj, mp);
                T v = o.b;
                v += o.cxm * xw + o.cx * xe;
                v += o.cym * R(t, jj - 1, m) + o.cy * R(t, jj + 1, m);
                v += zm * r_m2[p] + o.cz * r_0[p];
                v /= d;
                const long long c0 = col(j, wrapz(t), m);
                // non-temporal: the new iterate is not read again by this
                // sweep, and keeping it out of the caches leaves them to
                // the ring rows that neighbouring tiles read again (256^3
                // FP32: 88.5 -> 84.6 us a sweep)
                __builtin_nontemporal_store(e2 ? r_m1[p] : v, xout + c0);
                __builtin_nontemporal_store(e2 ? v : r_m1[p], xout + c0 + 1);
                if (DOT) {
                    sbx += (double)o.b * (double)v + (double)bft[p] * (double)r_m1[p];
                    sx += (double)v + (double)r_m1[p];
                    sb += (double)o.b + (double)bft[p];
                }
            }
        }
        // S: stage the xin of plane s+2 (requested at step s-1); keep the
        // second colour's operator of plane s
#pragma unroll
        for (int p = 0; p < PX; ++p) {
            const int m = threadIdx.x + p * tpr;
            if (!FIRST)
                stage(s + 2, m, xr[XU][p] + (PROL ? xq[XU][p] : (T)0),
                      er[XU][p] + (PROL ? eq[XU][p] : (T)0));
            kop[p] = pick(op[PU][p], 1 - e1);
            zr2[p] = zr1[p];
            zr1[p] = zrs[p];
        }
        // LOP: this wavefront's loads into the other buffer have landed; the
        // barrier then publishes every wavefront's
        if constexpr (LOP) __builtin_amdgcn_s_waitcnt(0);
        lds_barrier();
    };
    for (int s = s0;; s += 2) {
        step(Ic<0>(), s);
        if (s + 1 > k1) break;
        step(Ic<1>(), s + 1);
        if (s + 2 > k1) break;
    }
    if (DOT) {
        block_partial(sbx, part);
        block_partial(sx, part + kRedBlocks);
        block_partial(sb, part + 2 * kRedBlocks);
    }
}

// Tiled launch shape of the z-marching stencil kernels: blocks of 64 x 16
// threads own a 64 x 16 column of cells and march zc planes (zc even when
// `even`, for kernels that pair planes).
struct TGrid {
    int tx, ty, zc, blocks;
};
static TGrid tgrid(int nx, int ny, int nz, int target = 1024, bool even = false) {
    TGrid g;
    g.tx = (nx + 63) / 64;
    g.ty = (ny + 15) / 16;
    const long long cols = (long long)g.tx * g.ty;
    long long chunks = target / cols;
    if (chunks * cols > kRedBlocks) chunks = kRedBlocks / cols;
    if (chunks < 1) chunks = 1;
    if (chunks > nz) chunks = nz;
    g.zc = (int)((nz + chunks - 1) / chunks);
    if (even && (g.zc & 1)) ++g.zc;
    chunks = (nz + g.zc - 1) / g.zc;
    g.blocks = (int)(cols * chunks);
    return g;
}

// The same restriction on LDS tiles (64 x 16 fine cells, zc fine planes
// per block, zc even; nx, ny, nz even): the iterate of each plane (own
// cells plus a one-cell halo ring, coordinates wrapped) goes to one of
// three LDS planes, the residual of a cell is added to its z-partner's in
// registers, then over its x/y partners through LDS.  The iterate of plane
// s+2 and the operator of plane s are requested at the start of step s.
template <class T, bool DGI>
__global__ void __launch_bounds__(1024)
k_resid_restrict_tiled(LevelK<T> L, TGrid g, const T *__restrict__ b, const T *__restrict__ x,
                       T *__restrict__ bc) {
    __shared__ T X[3][18][66];
    __shared__ T Rz[16][64];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int cnx = nx >> 1, cny = ny >> 1;
    const int bid = xcd_order(blockIdx.x, gridDim.x);
    const int tx = bid % g.tx, ty = (bid / g.tx) % g.ty, ch = bid / (g.tx * g.ty);
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = tx * 64 + li, j = ty * 16 + lj;
    const int k0 = ch * g.zc, k1 = min(k0 + g.zc, nz);
    const bool in = i < nx && j < ny;
    const bool emit = in && !(li & 1) && !(lj & 1);
    const long long sy = nx, sz = (long long)nx * ny;
    auto wrapn = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };
    auto wrapk = [&](int k) { return wrapn(k, nz); };
    const long long own = wrapn(i, nx) + sy * wrapn(j, ny);
    const int t = li + 64 * lj;
    int hx = 0, hy = 0;
    if (t < 66) { hx = t - 1; hy = -1; }
    else if (t < 132) { hx = t - 67; hy = 16; }
    else if (t < 148) { hx = -1; hy = t - 132; }
    else { hx = 64; hy = t - 148; }
    const bool halo = t < 164;
    const long long hcol = wrapn(tx * 64 + hx, nx) + sy * wrapn(ty * 16 + hy, ny);
    const long long dxm = in ? (long long)(wrap_m(i, nx) - i) : 0;
    const long long dym = in ? (long long)(wrap_m(j, ny) - j) * sy : 0;
    const long long cco = (long long)(i >> 1) + (long long)cnx * (j >> 1);   // coarse column
    const long long csz = (long long)cnx * cny;
    auto put = [&](int kc) {   // the 2x2 sum of the pair stored in Rz
        if (emit) {
            const T v = (Rz[lj][li] + Rz[lj][li + 1]) + (Rz[lj + 1][li] + Rz[lj + 1][li + 1]);
            bc[cco + csz * kc] = v * (T)0.125;
        }
    };

    T x_m = x[own + sz * wrapk(k0 - 1)];
    T x_0 = x[own + sz * k0];
    X[k0 % 3][lj + 1][li + 1] = x_0;
    if (halo) X[k0 % 3][hy + 1][hx + 1] = x[hcol + sz * k0];
    T xo1 = x[own + sz * wrapk(k0 + 1)];
    T xh1 = halo ? x[hcol + sz * wrapk(k0 + 1)] : (T)0;
    T czm = in ? L.cz[own + sz * wrapk(k0 - 1)] : (T)0;
    __syncthreads();
    T rpair = 0;
    for (int s = k0; s < k1; ++s) {
        T xo2 = 0, xh2 = 0;
        if (s + 1 < k1) {
            const long long p2 = sz * wrapk(s + 2);
            xo2 = x[own + p2];
            if (halo) xh2 = x[hcol + p2];
        }
        const long long c = own + sz * s;
        T cxw = 0, cxe = 0, cys = 0, cyn = 0, czp = 0, dgc = 0, bv = 0;
        if (in) {
            cxw = L.cx[c + dxm]; cxe = L.cx[c];
            cys = L.cy[c + dym]; cyn = L.cy[c];
            czp = L.cz[c]; bv = b[c];
            dgc = DGI ? diag_sum(cxw, cxe, cys, cyn, czm, czp) : L.dg[c];
        }
        const int wb = (s + 1) % 3, rb = s % 3;
        X[wb][lj + 1][li + 1] = xo1;
        if (halo) X[wb][hy + 1][hx + 1] = xh1;
        const bool odd = (s - k0) & 1;
        if (!odd && s > k0) put((s - 2) >> 1);
        T r = 0;
        if (in) {
            T v = bv;
            v += cxw * X[rb][lj + 1][li] + cxe * X[rb][lj + 1][li + 2];
            v += cys * X[rb][lj][li + 1] + cyn * X[rb][lj + 2][li + 1];
            v += czm * x_m + czp * xo1;
            r = v - dgc * x_0;
        }
        if (!odd) {
            rpair = r;
        } else {
            Rz[lj][li] = rpair + r;
        }
        x_m = x_0; x_0 = xo1; czm = czp;
        xo1 = xo2; xh1 = xh2;
        __syncthreads();
    }
    put((k1 - 2) >> 1);
}

// residual of the 8 children of each coarse cell, averaged into bc
template <class T, class B, bool DGI>
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
                acc += ((T)b[c] + nb_sum(L, x, i, j, k, c)) - diag_of<T, DGI>(L, i, j, k, c) * x[c];
            }
    bc[I + (long long)cxn * (J + (long long)cyn * K)] = acc * (T)0.125;
}

// x_f += x_c of the parent cell: piecewise-constant prolongation, the
// transpose of the 8-child average (times 8), so the V-cycle (symmetric
// smoothing, R = P^T / 8) is a symmetric operator and CG's recurrences hold
// for it.  (Trilinear interpolation breaks that symmetry: the forward
// solves still converge, but smooth right-hand sides such as adjoint
// sources stall for tens of iterations.)  One thread per coarse cell.
template <class T>
__global__ void __launch_bounds__(256)
k_prolong_add(int fx, int fy, int fz, T *__restrict__ xf, const T *__restrict__ xc) {
    const int cx = fx >> 1, cy = fy >> 1;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= cx || J >= cy) return;
    const T v = xc[I + (long long)cx * (J + (long long)cy * K)];
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const long long row = (long long)fx * ((2 * J + dj) + (long long)fy * (2 * K + dk));
            xf[row + 2 * I] += v;
            xf[row + 2 * I + 1] += v;
        }
}

// ---------------------------------------------------------------------
// Transfers of an irregular level: a factor of 1 or 2 per axis (semi-
// coarsening), and with factor 2 on an odd size the last coarse cell has
// a single child.  R = P^T / (fx fy fz) with the level's nominal factors
// for every coarse cell, so the V-cycle stays symmetric.
// ---------------------------------------------------------------------
struct Coarsening {
    int fx, fy, fz;      // 1 or 2 per axis
    int cnx, cny, cnz;   // coarse sizes, ceil(n / f)
};

// coarse b = (1 / (fx fy fz)) * the sum of the children's residuals; one
// thread per coarse cell
template <class T, class B, bool DGI>
__global__ void __launch_bounds__(256)
k_resid_restrict_g(LevelK<T> L, Coarsening g, const B *__restrict__ b, const T *__restrict__ x,
                   T *__restrict__ bc) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= g.cnx || J >= g.cny || K >= g.cnz) return;
    const int i0 = g.fx * I, j0 = g.fy * J, k0 = g.fz * K;
    const int i1 = min(i0 + g.fx, L.nx), j1 = min(j0 + g.fy, L.ny), k1 = min(k0 + g.fz, L.nz);
    T acc = 0;
    for (int k = k0; k < k1; ++k)
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) {
                const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
                acc += ((T)b[c] + nb_sum(L, x, i, j, k, c)) - diag_of<T, DGI>(L, i, j, k, c) * x[c];
            }
    bc[I + (long long)g.cnx * (J + (long long)g.cny * K)] = acc / (T)(g.fx * g.fy * g.fz);
}

// x_f += x_c of the parent cell; one thread per fine cell
template <class T>
__global__ void __launch_bounds__(256)
k_prolong_add_g(int nx, int ny, int nz, Coarsening g, T *__restrict__ xf, const T *__restrict__ xc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const long long p = i / g.fx + (long long)g.cnx * (j / g.fy + (long long)g.cny * (k / g.fz));
    xf[i + (long long)nx * (j + (long long)ny * k)] += xc[p];
}

// ---------------------------------------------------------------------
// Trilinear transfers (PPS_TRANSFER_TRILINEAR).  Along one axis, fine cell
// i lies a quarter of a coarse cell from the centre of its parent I = i/2,
// towards the coarse neighbour O = I - 1 (i even) or I + 1 (i odd):
// interpolation weighs the parent 3/4 and O 1/4.  O wraps on a periodic
// axis; beyond a wall it is the parent itself (constant extrapolation, the
// zero-gradient condition), which gives the wall cell the weight 1.  The 3-D
// prolongation P is the tensor product; the restriction is R = P^T / 8, so
// coarse cell I gathers fine cells 2I-1, 2I, 2I+1, 2I+2 with 1/4, 3/4, 3/4,
// 1/4 per axis (a tap beyond a wall falls on the fine cell next to it).
// ---------------------------------------------------------------------
struct TriK {
    int fx, fy, fz;     // fine level
    int px, py, pz;     // periodic axes
};
// coarse neighbour index of fine cell i (nc coarse cells on the axis)
__device__ __forceinline__ int tri_other(int i, int nc, int per) {
    const int I = i >> 1;
    int o = (i & 1) ? I + 1 : I - 1;
    if (o < 0) o = per ? nc - 1 : 0;
    else if (o >= nc) o = per ? 0 : nc - 1;
    return o;
}
// fine tap t (0..3) of coarse cell I: index 2I - 1 + t, wrapped or moved
// back inside at a wall
__device__ __forceinline__ int tri_tap(int I, int t, int nf, int per) {
    int i = 2 * I - 1 + t;
    if (i < 0) i = per ? nf - 1 : 0;
    else if (i >= nf) i = per ? 0 : nf - 1;
    return i;
}

// r = b + sum couplings x_nb - d x on one level (the residual the
// full-weighting restriction reads)
template <class T, class B, bool DGI>
__global__ void __launch_bounds__(256)
k_resid_field(LevelK<T> L, const B *__restrict__ b, const T *__restrict__ x, T *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= L.nx || j >= L.ny) return;
    const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
    r[c] = ((T)b[c] + nb_sum(L, x, i, j, k, c)) - diag_of<T, DGI>(L, i, j, k, c) * x[c];
}

// bc = R r: one thread per coarse cell, 64 taps
template <class T>
__global__ void __launch_bounds__(256)
k_restrict_tri(TriK t, const T *__restrict__ r, T *__restrict__ bc) {
    const int cx = t.fx >> 1, cy = t.fy >> 1, cz = t.fz >> 1;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= cx || J >= cy) return;
    const T w[4] = {(T)0.25, (T)0.75, (T)0.75, (T)0.25};
    int ii[4], jj[4], kk[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        ii[a] = tri_tap(I, a, t.fx, t.px);
        jj[a] = tri_tap(J, a, t.fy, t.py);
        kk[a] = tri_tap(K, a, t.fz, t.pz);
    }
    T acc = 0;
#pragma unroll
    for (int c3 = 0; c3 < 4; ++c3) {
        T accy = 0;
#pragma unroll
        for (int b3 = 0; b3 < 4; ++b3) {
            const long long row = (long long)t.fx * (jj[b3] + (long long)t.fy * kk[c3]);
            T accx = 0;
#pragma unroll
            for (int a = 0; a < 4; ++a) accx += w[a] * r[row + ii[a]];
            accy += w[b3] * accx;
        }
        acc += w[c3] * accy;
    }
    bc[I + (long long)cx * (J + (long long)cy * K)] = acc * (T)0.125;
}

// xf += P xc: one thread per fine cell
template <class T>
__global__ void __launch_bounds__(256)
k_prolong_tri_add(TriK t, T *__restrict__ xf, const T *__restrict__ xc) {
    const int cx = t.fx >> 1, cy = t.fy >> 1, cz = t.fz >> 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= t.fx || j >= t.fy) return;
    const int ia[2] = {i >> 1, tri_other(i, cx, t.px)};
    const int ja[2] = {j >> 1, tri_other(j, cy, t.py)};
    const int ka[2] = {k >> 1, tri_other(k, cz, t.pz)};
    const T w[2] = {(T)0.75, (T)0.25};
    T acc = 0;
#pragma unroll
    for (int c3 = 0; c3 < 2; ++c3) {
        T accy = 0;
#pragma unroll
        for (int b3 = 0; b3 < 2; ++b3) {
            const long long row = (long long)cx * (ja[b3] + (long long)cy * ka[c3]);
            accy += w[b3] * (w[0] * xc[row + ia[0]] + w[1] * xc[row + ia[1]]);
        }
        acc += w[c3] * accy;
    }
    xf[i + (long long)t.fx * (j + (long long)t.fy * k)] += acc;
}

// ---------------------------------------------------------------------
// operator packing: 8-slot FP64 A + padded p -> couplings, diagonal,
// folded right-hand side, interior iterate.  flags: bit0 NaN/tiny A7,
// bit1 asymmetric, bit2 nonzero excess (system not singular).
// The two couplings of a face differ when SetupPoisson's spacings do: on a
// uniform grid whose step is not a power of two, dx(i) = xh(i) - xh(i-1)
// varies in its last bits (paris.f90:2959-3102), so A2(i) and A1(i+1)
// differ by an ulp or so.  Such a face counts as symmetric, and the stored
// operator takes the + coupling A2(i) for both of its rows (a change of
// relative size 1e-15, far below any solve tolerance).
__device__ __forceinline__ bool asym(double a, double b) {
    return !(fabs(a - b) <= 1e-12 * fmax(fabs(a), fabs(b)));
}
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
    // symmetry: my + coupling equals the - coupling of my + neighbour (to
    // rounding: see asym)
    if (!lastx || d.px) {
        const long long e = lastx ? c - (d.nx - 1) : c + 1;
        if (asym(A[e], a[1])) f |= 2;
    }
    if (!lasty || d.py) {
        const long long e = lasty ? c - (long long)(d.ny - 1) * sy : c + sy;
        if (asym(A[2 * n + e], a[3])) f |= 2;
    }
    if (!lastz || d.pz) {
        const long long e = lastz ? c - (long long)(d.nz - 1) * sz : c + sz;
        if (asym(A[4 * n + e], a[5])) f |= 2;
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
// recomputed from the fine faces of the neighbouring coarse cell.  DGI: the
// V-cycle forms every diagonal from the couplings (singular systems), so
// the coarse diagonal is not written.
template <class Tf, class Tc, bool DGI>
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
    const long long c = I + (long long)nx * (J + (long long)ny * K);
    ccx[c] = (Tc)vx; ccy[c] = (Tc)vy; ccz[c] = (Tc)vz;
    if (DGI) return;
    const double wx = s * cpx(wrap_m(I, nx), J, K), wy = s * cpy(I, wrap_m(J, ny), K),
                 wz = s * cpz(I, J, wrap_m(K, nz));
    double ex = 0.0;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                ex += excess_of(F, i, j, k, fi(i, j, k));
            }
    cdg[c] = (Tc)(((vx + wx) + (vy + wy)) + (vz + wz) + 0.125 * ex);
}

// Coarse operator of an irregular level (the rule of k_coarsen with
// per-axis factors): the coupling across a coarse x face is the sum of the
// fine x faces on it (the +x couplings of the coarse cell's last children
// in x) times 1 / (fy fz fx^2), i.e. their mean over the face area divided
// by the square of the spacing factor (likewise y, z); the diagonal is the
// coupling sum plus 1 / (fx fy fz) times the children's excess.
template <class Tf, class Tc, bool DGI>
__global__ void __launch_bounds__(256)
k_coarsen_g(LevelK<Tf> F, Coarsening g, Tc *__restrict__ ccx, Tc *__restrict__ ccy, Tc *__restrict__ ccz,
            Tc *__restrict__ cdg) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= g.cnx || J >= g.cny || K >= g.cnz) return;
    const long long fsy = F.nx, fsz = (long long)F.nx * F.ny;
    auto fi = [&](int i, int j, int k) { return i + fsy * j + fsz * k; };
    auto lo = [](int C, int f) { return f * C; };
    auto hi = [](int C, int f, int n) { return min(f * C + f, n); };
    auto cpx = [&](int I_, int J_, int K_) {
        const int i = hi(I_, g.fx, F.nx) - 1;
        double a = 0.0;
        for (int k = lo(K_, g.fz); k < hi(K_, g.fz, F.nz); ++k)
            for (int j = lo(J_, g.fy); j < hi(J_, g.fy, F.ny); ++j) a += (double)F.cx[fi(i, j, k)];
        return a;
    };
    auto cpy = [&](int I_, int J_, int K_) {
        const int j = hi(J_, g.fy, F.ny) - 1;
        double a = 0.0;
        for (int k = lo(K_, g.fz); k < hi(K_, g.fz, F.nz); ++k)
            for (int i = lo(I_, g.fx); i < hi(I_, g.fx, F.nx); ++i) a += (double)F.cy[fi(i, j, k)];
        return a;
    };
    auto cpz = [&](int I_, int J_, int K_) {
        const int k = hi(K_, g.fz, F.nz) - 1;
        double a = 0.0;
        for (int j = lo(J_, g.fy); j < hi(J_, g.fy, F.ny); ++j)
            for (int i = lo(I_, g.fx); i < hi(I_, g.fx, F.nx); ++i) a += (double)F.cz[fi(i, j, k)];
        return a;
    };
    const double sx = 1.0 / (double)(g.fy * g.fz * g.fx * g.fx);
    const double sy = 1.0 / (double)(g.fx * g.fz * g.fy * g.fy);
    const double sz = 1.0 / (double)(g.fx * g.fy * g.fz * g.fz);
    const double vx = sx * cpx(I, J, K), vy = sy * cpy(I, J, K), vz = sz * cpz(I, J, K);
    const long long c = I + (long long)g.cnx * (J + (long long)g.cny * K);
    ccx[c] = (Tc)vx; ccy[c] = (Tc)vy; ccz[c] = (Tc)vz;
    if (DGI) return;
    const double wx = sx * cpx(wrap_m(I, g.cnx), J, K), wy = sy * cpy(I, wrap_m(J, g.cny), K),
                 wz = sz * cpz(I, J, wrap_m(K, g.cnz));
    double ex = 0.0;
    for (int k = lo(K, g.fz); k < hi(K, g.fz, F.nz); ++k)
        for (int j = lo(J, g.fy); j < hi(J, g.fy, F.ny); ++j)
            for (int i = lo(I, g.fx); i < hi(I, g.fx, F.nx); ++i) ex += excess_of(F, i, j, k, fi(i, j, k));
    cdg[c] = (Tc)(((vx + wx) + (vy + wy)) + (vz + wz) + ex / (double)(g.fx * g.fy * g.fz));
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
    for (int c = 0; c < nb; ++c) {
        const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
        dsum += singular ? (double)diag_of<T, true>(L, i, j, k, c) : (double)L.dg[c];
    }
    const double alpha = singular ? dsum / ((double)nb * (double)nb) : 0.0;
    for (int c = tid; c < nb; c += 1024) {
        const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
        const long long sy = L.nx, sz = (long long)L.nx * L.ny;
        const int ip = wrap_p(i, L.nx), jp = wrap_p(j, L.ny), kp = wrap_p(k, L.nz);
        const int im = wrap_m(i, L.nx), jm = wrap_m(j, L.ny), km = wrap_m(k, L.nz);
        double *row = M + (long long)c * nb;
        row[c] += singular ? (double)diag_of<T, true>(L, i, j, k, c) : (double)L.dg[c];
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

// The same inverse in LDS for a coarsest level of nb <= 90 cells (the
// matrix and copies of the pivot row and column fit 64 KB): per pivot, the
// pivot row and column are copied, then every entry is updated from the
// copies (the arithmetic of k_bottom_build), two barriers per pivot.
// Requires nb <= 1024 threads (true for nb <= 90).
template <class T>
__global__ void __launch_bounds__(1024)
k_bottom_build_lds(LevelK<T> L, int singular, T *__restrict__ Minv) {
    extern __shared__ double Ms[];
    __shared__ double red[16];
    const int nb = L.nx * L.ny * L.nz;
    const int tid = threadIdx.x;
    constexpr int NT = 1024;
    double *M = Ms, *rowp = Ms + nb * nb, *colp = rowp + nb;
    for (int e = tid; e < nb * nb; e += NT) M[e] = 0.0;
    // alpha = mean diagonal / nb (block reduction of the diagonals)
    double dpart = 0.0;
    for (int c = tid; c < nb; c += NT) {
        const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
        dpart += singular ? (double)diag_of<T, true>(L, i, j, k, c) : (double)L.dg[c];
    }
    dpart = wsum(dpart);
    if ((tid & 63) == 0) red[tid >> 6] = dpart;
    __syncthreads();
    double dsum = 0.0;
    for (int w = 0; w < NT / 64; ++w) dsum += red[w];
    const double alpha = singular ? dsum / ((double)nb * (double)nb) : 0.0;
    for (int c = tid; c < nb; c += NT) {
        const int i = c % L.nx, j = (c / L.nx) % L.ny, k = c / (L.nx * L.ny);
        const int sy = L.nx, sz = L.nx * L.ny;
        const int ip = wrap_p(i, L.nx), jp = wrap_p(j, L.ny), kp = wrap_p(k, L.nz);
        const int im = wrap_m(i, L.nx), jm = wrap_m(j, L.ny), km = wrap_m(k, L.nz);
        double *row = M + c * nb;
        row[c] += singular ? (double)diag_of<T, true>(L, i, j, k, c) : (double)L.dg[c];
        row[c + (ip - i)] -= (double)L.cx[c];
        row[c + (im - i)] -= (double)L.cx[c + (im - i)];
        row[c + (jp - j) * sy] -= (double)L.cy[c];
        row[c + (jm - j) * sy] -= (double)L.cy[c + (jm - j) * sy];
        row[c + (kp - k) * sz] -= (double)L.cz[c];
        row[c + (km - k) * sz] -= (double)L.cz[c + (km - k) * sz];
        for (int e = 0; e < nb; ++e) row[e] += alpha;
    }
    __syncthreads();
    // thread tid owns column tid % nb of rows tid / nb, tid / nb + rs, ...
    const int col = tid % nb, r0 = tid / nb, rs = NT / nb;
    for (int p = 0; p < nb; ++p) {
        if (tid < nb) {
            rowp[tid] = M[p * nb + tid];
            colp[tid] = M[tid * nb + p];
        }
        __syncthreads();
        const double piv = 1.0 / rowp[p];
        if (r0 < rs) {
            const double rc = rowp[col] * piv;
            for (int r = r0; r < nb; r += rs) {
                double *m = M + r * nb + col;
                if (r == p) *m = col == p ? piv : rc;
                else if (col == p) *m = -colp[r] * piv;
                else *m -= colp[r] * rc;
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < nb * nb; e += NT) Minv[e] = (T)M[e];
}

// x = Minv b.  Minv is the inverse of the symmetric coarsest operator, so
// thread r reads column r (consecutive threads, consecutive addresses)
// rather than row r.
template <class T>
__global__ void __launch_bounds__(512)
k_bottom_solve(int nb, const T *__restrict__ Minv, const T *__restrict__ b, T *__restrict__ x) {
    __shared__ T sb[512];
    for (int e = threadIdx.x; e < nb; e += 512) sb[e] = b[e];
    __syncthreads();
    for (int r = threadIdx.x; r < nb; r += 512) {
        T acc = 0;
        for (int e = 0; e < nb; ++e) acc += Minv[(long long)e * nb + r] * sb[e];
        x[r] = acc;
    }
}

// ---------------------------------------------------------------------
// The coarse end of a V-cycle in one workgroup, in LDS.  Levels 0 .. nl-1
// of the tail (level nl-1 is the coarsest, solved with the dense inverse)
// are staged into LDS at the start: their operators, the right-hand side
// of level 0 and the inverse.  Colour passes, restriction, the bottom
// solve and prolongation then run back to back on LDS with workgroup
// barriers in between, and the iterate of level 0 is written back at the
// end: one launch instead of several per level, and LDS latency instead of
// L2 latency per pass.  The arithmetic is that of k_rb / k_resid_restrict /
// k_bottom_solve / k_prolong_add (in-place colour passes: pre-smoothing
// colour 0 first, post-smoothing colour 1 first).
// LDS layout: off[l] = {cx, cy, cz, dg (DGI false only), b, x} of level l,
// off_minv = the inverse; all offsets in elements of T.
// ---------------------------------------------------------------------
constexpr int kTailMax = 8;
template <class T> struct TailK {
    int nl, nb, nu_pre, nu_post;
    LevelK<T> L[kTailMax];   // operators in global memory
    const T *b0;             // right-hand side of level 0 (global)
    T *x0;                   // iterate of level 0 (global, written at the end)
    const T *minv;
    int off[kTailMax][6];
    int off_minv;
    int off_r;               // residual of a level (trilinear transfers)
    int px, py, pz;          // periodic axes (trilinear transfers)
};

// TRI: trilinear transfers (k_resid_field + k_restrict_tri, k_prolong_tri_add
// on LDS) instead of the 8-child average and piecewise-constant prolongation
template <class T, bool DGI, bool TRI>
__global__ void __launch_bounds__(1024)
k_tail(TailK<T> K) {
    extern __shared__ unsigned char smem_raw[];
    T *sm = reinterpret_cast<T *>(smem_raw);
    const int tid = threadIdx.x;
    constexpr int NT = 1024;
    // stage operators, the right-hand side of level 0 and the inverse: per
    // level, every thread issues all its loads (up to 4 elements of each
    // array) before its LDS stores, so a level costs one memory latency
    // rather than one per element and array
    for (int l = 0; l < K.nl; ++l) {
        const LevelK<T> &G = K.L[l];
        const int n = G.nx * G.ny * G.nz;
        const T *src[5] = {G.cx, G.cy, G.cz, DGI ? G.cx : G.dg, l == 0 ? K.b0 : G.cx};
        T *dst[5] = {sm + K.off[l][0], sm + K.off[l][1], sm + K.off[l][2], sm + K.off[l][3],
                     sm + K.off[0][4]};
        const int na = 3 + (DGI ? 0 : 1);
        for (int e0 = tid; e0 < n; e0 += 4 * NT) {
            T v[5][4];
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                const bool use = a < 3 || (a == 3 && na == 4) || (a == 4 && l == 0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = e0 + u * NT;
                    v[a][u] = (use && e < n) ? src[a][e] : (T)0;
                }
            }
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                const bool use = a < 3 || (a == 3 && na == 4) || (a == 4 && l == 0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = e0 + u * NT;
                    if (use && e < n) dst[a][e] = v[a][u];
                }
            }
        }
    }
    {
        T *mi = sm + K.off_minv;
        const int nm = K.nb * K.nb;
        for (int e0 = tid; e0 < nm; e0 += 4 * NT) {
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = e0 + u * NT < nm ? K.minv[e0 + u * NT] : (T)0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (e0 + u * NT < nm) mi[e0 + u * NT] = v[u];
        }
    }
    __syncthreads();
    auto lev = [&](int l) {
        LevelK<T> L;
        L.nx = K.L[l].nx; L.ny = K.L[l].ny; L.nz = K.L[l].nz;
        L.cx = sm + K.off[l][0]; L.cy = sm + K.off[l][1]; L.cz = sm + K.off[l][2];
        L.dg = sm + K.off[l][3];
        return L;
    };
    // (a, b, c) of linear index e in an na x nb x nc box: shifts and masks
    // when na and nb are powers of two (every level of a power-of-two grid),
    // divisions otherwise
    auto split = [](int e, int na, int nb, int &a, int &b, int &c) {
        if (((na & (na - 1)) | (nb & (nb - 1))) == 0) {
            const int la = __builtin_ctz(na), lb = __builtin_ctz(nb);
            a = e & (na - 1);
            b = (e >> la) & (nb - 1);
            c = e >> (la + lb);
        } else {
            a = e % na;
            b = (e / na) % nb;
            c = e / (na * nb);
        }
    };
    auto cpass = [&](int l, int colour, bool first) {
        const LevelK<T> L = lev(l);
        const T *b = sm + K.off[l][4];
        T *x = sm + K.off[l][5];
        const int nxh = (L.nx + 1) >> 1;
        const int total = nxh * L.ny * L.nz;
        for (int e = tid; e < total; e += NT) {
            int ih, j, k;
            split(e, nxh, L.ny, ih, j, k);
            const int i = 2 * ih + ((j + k + colour) & 1);
            if (i >= L.nx) continue;
            const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
            const T d = diag_of<T, DGI>(L, i, j, k, c);
            x[c] = first ? b[c] / d : (b[c] + nb_sum(L, x, i, j, k, c)) / d;
        }
        __syncthreads();
    };
    // pre-smoothing and restriction down to the coarsest level
    for (int l = 0; l < K.nl - 1; ++l) {
        cpass(l, 0, true);
        cpass(l, 1, false);
        for (int q = 1; q < K.nu_pre; ++q) {
            cpass(l, 0, false);
            cpass(l, 1, false);
        }
        const LevelK<T> L = lev(l);
        const T *b = sm + K.off[l][4];
        const T *x = sm + K.off[l][5];
        const int cxn = L.nx >> 1, cyn = L.ny >> 1, czn = L.nz >> 1;
        T *bc = sm + K.off[l + 1][4];
        if constexpr (TRI) {
            T *r = sm + K.off_r;
            for (int e = tid; e < L.nx * L.ny * L.nz; e += NT) {
                int i, j, k;
                split(e, L.nx, L.ny, i, j, k);
                r[e] = (b[e] + nb_sum(L, x, i, j, k, e)) - diag_of<T, DGI>(L, i, j, k, e) * x[e];
            }
            __syncthreads();
            const T w[4] = {(T)0.25, (T)0.75, (T)0.75, (T)0.25};
            for (int e = tid; e < cxn * cyn * czn; e += NT) {
                int I, J, Kz;
                split(e, cxn, cyn, I, J, Kz);
                T acc = 0;
                for (int c3 = 0; c3 < 4; ++c3) {
                    const int kf = tri_tap(Kz, c3, L.nz, K.pz);
                    T accy = 0;
                    for (int b3 = 0; b3 < 4; ++b3) {
                        const int jf = tri_tap(J, b3, L.ny, K.py);
                        const T *row = r + L.nx * (jf + L.ny * kf);
                        T accx = 0;
                        for (int a = 0; a < 4; ++a) accx += w[a] * row[tri_tap(I, a, L.nx, K.px)];
                        accy += w[b3] * accx;
                    }
                    acc += w[c3] * accy;
                }
                bc[e] = acc * (T)0.125;
            }
            __syncthreads();
            continue;
        }
        for (int e = tid; e < cxn * cyn * czn; e += NT) {
            int I, J, Kz;
            split(e, cxn, cyn, I, J, Kz);
            T acc = 0;
            for (int dk = 0; dk < 2; ++dk)
                for (int dj = 0; dj < 2; ++dj)
                    for (int di = 0; di < 2; ++di) {
                        const int i = 2 * I + di, j = 2 * J + dj, k = 2 * Kz + dk;
                        const long long c = i + (long long)L.nx * (j + (long long)L.ny * k);
                        acc += (b[c] + nb_sum(L, x, i, j, k, c)) - diag_of<T, DGI>(L, i, j, k, c) * x[c];
                    }
            bc[e] = acc * (T)0.125;
        }
        __syncthreads();
    }
    // dense solve on the coarsest level
    {
        const int nb = K.nb;
        const T *b = sm + K.off[K.nl - 1][4];
        T *x = sm + K.off[K.nl - 1][5];
        const T *mi = sm + K.off_minv;
        // column r of the (symmetric) inverse: no LDS bank conflicts
        for (int r = tid; r < nb; r += NT) {
            T acc = 0;
            for (int e = 0; e < nb; ++e) acc += mi[e * nb + r] * b[e];
            x[r] = acc;
        }
        __syncthreads();
    }
    // prolongation and post-smoothing back up
    for (int l = K.nl - 2; l >= 0; --l) {
        const int fx = K.L[l].nx, fy = K.L[l].ny, fz = K.L[l].nz;
        const int cx = fx >> 1, cy = fy >> 1;
        T *xf = sm + K.off[l][5];
        const T *xc = sm + K.off[l + 1][5];
        for (int e = tid; e < fx * fy * fz; e += NT) {
            int i, j, k;
            split(e, fx, fy, i, j, k);
            if constexpr (TRI) {
                const int cz = fz >> 1;
                const int ia[2] = {i >> 1, tri_other(i, cx, K.px)};
                const int ja[2] = {j >> 1, tri_other(j, cy, K.py)};
                const int ka[2] = {k >> 1, tri_other(k, cz, K.pz)};
                const T w[2] = {(T)0.75, (T)0.25};
                T acc = 0;
                for (int c3 = 0; c3 < 2; ++c3) {
                    T accy = 0;
                    for (int b3 = 0; b3 < 2; ++b3) {
                        const T *row = xc + cx * (ja[b3] + cy * ka[c3]);
                        accy += w[b3] * (w[0] * row[ia[0]] + w[1] * row[ia[1]]);
                    }
                    acc += w[c3] * accy;
                }
                xf[e] += acc;
            } else {
                xf[e] += xc[(i >> 1) + cx * ((j >> 1) + cy * (k >> 1))];
            }
        }
        __syncthreads();
        for (int q = 0; q < K.nu_post; ++q) {
            cpass(l, 1, false);
            cpass(l, 0, false);
        }
    }
    // the iterate of level 0 back to global memory
    {
        const int n0 = K.L[0].nx * K.L[0].ny * K.L[0].nz;
        const T *x = sm + K.off[0][5];
        for (int e = tid; e < n0; e += NT) K.x0[e] = x[e];
    }
}

// ---------------------------------------------------------------------
// FP64 outer iteration kernels.  Scalars live on the device:
// ---------------------------------------------------------------------
enum { S_RZ = 0, S_ZMEAN, S_BETA, S_DQ, S_ALPHA, S_RR, S_BB, S_FIRST, S_STOP, S_N };
// alpha of the direction in ring slot j: S[S_AH + j] (j < kRingSlots;
// the scalars take 64 doubles)
constexpr int kRingSlots = 9;   // directions held for deferred x updates
constexpr int S_AH = S_N;
static_assert(S_AH + kRingSlots <= 64, "ring alphas fit the scalars");

// 2.5-D launch shape for z-marching kernels: (64, 4) threads over (x, y),
// each thread walks `zc` consecutive planes; at most kRedBlocks blocks so
// that one partial slot per block suffices.
struct ZGrid {
    dim3 grid;
    int zc;
    int blocks;
};
static ZGrid zgrid(int nx, int ny, int nz, int target = 16384) {
    const int gx = (nx + 63) / 64, gy = (ny + 3) / 4;
    const long long cols = (long long)gx * gy;
    long long chunks = (target < kRedBlocks ? target : kRedBlocks) / cols;
    if (chunks < 1) chunks = 1;
    if (chunks > nz) chunks = nz;
    const int zc = (int)((nz + chunks - 1) / chunks);
    chunks = (nz + zc - 1) / zc;
    ZGrid z;
    z.grid = dim3((unsigned)gx, (unsigned)gy, (unsigned)chunks);
    z.zc = zc;
    z.blocks = (int)(cols * chunks);
    return z;
}
// fixed-order sum of nb partials, the same value in every thread (any
// block of at most 1024 threads whose x extent is a multiple of 64)
__device__ __forceinline__ double sum_partials(const double *part, int nb) {
    __shared__ double sm[16];
    const int nt = blockDim.x * blockDim.y * blockDim.z;
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nw = nt >> 6;
    double v = 0.0;
    for (int b = tid; b < nb; b += nt) v += part[b];
    v = wsum(v);
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += sm[w];
    __syncthreads();
    return t;
}


// FP64 fine operator for the outer iteration: +x/+y/+z couplings and the
// diagonal, either the compact planes the setup pass writes (0 on wall
// faces) or the caller's A slots 2, 4, 6, 7 (mixed mode), whose wall-face
// coefficients the setup folded into b: op_x/op_y/op_z read them as 0 on
// a wall face (px/py/pz: periodic directions), so both give the same
// operator.
struct OpK {
    int nx, ny, nz;
    const double *cx, *cy, *cz, *dg;
    int px, py, pz;
};
__device__ __forceinline__ double op_x(const OpK &L, long long c, int i) {
    return (i == L.nx - 1 && !L.px) ? 0.0 : L.cx[c];
}
__device__ __forceinline__ double op_y(const OpK &L, long long c, int j) {
    return (j == L.ny - 1 && !L.py) ? 0.0 : L.cy[c];
}
__device__ __forceinline__ double op_z(const OpK &L, long long c, int k) {
    return (k == L.nz - 1 && !L.pz) ? 0.0 : L.cz[c];
}

// Diagonal of the FP64 outer operator at cell c from its six couplings
// (-x, +x, -y, +y, -z, +z): stored, or (DGI, used when the setup pass
// found every A7 equal to it) the six added in the order of the assembly's
// sum(A(i,j,k,1:6)) (modules.f90:2667), which reproduces A7 bit for bit.
template <bool DGI>
__device__ __forceinline__ double outer_diag(const OpK &L, long long c, double vxm, double vx, double vym,
                                             double vy, double vzm, double vz) {
    if (!DGI) return L.dg[c];
    return ((((vxm + vx) + vym) + vy) + vzm) + vz;
}

// r = b - M x ; partials: 0 = sum r^2, 1 = sum b^2 ; optional FP32 copy.
// z-marching over zc planes per thread (launch with zgrid()).
template <class TB, bool DGI>
__global__ void __launch_bounds__(256)
k_outer_resid(OpK L, int zc, const double *__restrict__ b, const double *__restrict__ x,
              double *__restrict__ r, TB *__restrict__ bt, double *__restrict__ part) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * zc;
    const int k1 = min(k0 + zc, L.nz);
    double rr = 0.0, bb = 0.0;
    if (i < L.nx && j < L.ny) {
        const long long sy = L.nx, sz = (long long)L.nx * L.ny;
        const int dxm = wrap_m(i, L.nx) - i, dxp = wrap_p(i, L.nx) - i;
        const long long dym = (long long)(wrap_m(j, L.ny) - j) * sy, dyp = (long long)(wrap_p(j, L.ny) - j) * sy;
        const long long col = i + sy * j;
        double xm = x[col + sz * wrap_m(k0, L.nz)];
        double x0 = x[col + sz * k0];
        for (int k = k0; k < k1; ++k) {
            const long long c = col + sz * k;
            const long long czm = col + sz * wrap_m(k, L.nz);
            const double xp = x[col + sz * wrap_p(k, L.nz)];
            const double vxm = op_x(L, c + dxm, wrap_m(i, L.nx)), vx = op_x(L, c, i);
            const double vym = op_y(L, c + dym, wrap_m(j, L.ny)), vy = op_y(L, c, j);
            const double vzm = op_z(L, czm, wrap_m(k, L.nz)), vz = op_z(L, c, k);
            double s = vxm * x[c + dxm] + vx * x[c + dxp];
            s += vym * x[c + dym] + vy * x[c + dyp];
            s += vzm * xm + vz * xp;
            const double bv = b[c];
            const double v = bv - (outer_diag<DGI>(L, c, vxm, vx, vym, vy, vzm, vz) * x0 - s);
            r[c] = v;
            if (bt) bt[c] = (TB)v;
            rr += v * v;
            bb += bv * bv;
            xm = x0;
            x0 = xp;
        }
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



// d_new = (z - zmean) + beta d_old, rounded to TZ (the direction is
// stored in the precision of z) ; q = M d_new from the rounded values, so
// that x += alpha d and r -= alpha q change r by exactly -M (alpha d) in
// FP64 ; partial sum d.q.  Given x, the previous iteration's
// x += alpha_prev d_old (alpha_prev = S[S_ALPHA], 0 for none) is applied
// here, where d_old is read anyway (the no-allocation solve; the others
// pass no x and defer the updates to k_flush).  On LDS tiles: d_new of every plane (own cells plus a
// one-cell halo ring, coordinates wrapped) is written to one of three LDS
// planes; the in-plane neighbours of q = M d_new are read there, the
// vertical ones from registers.  The raw z / d_old of plane s+2 and the
// operator of plane s are requested at the start of step s, so the loads
// overlap the arithmetic of the step; one barrier per plane (three
// buffers: the plane written at step s+1 was last read at step s-2).
// Registers held to 64 a thread (two blocks a CU): the FP64 form needs 66
// otherwise, one block a CU, and ran 239 against 223 us at 256^3.
template <class TZ, bool DGI>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
k_dq_tiled(OpK L, TGrid g, const TZ *__restrict__ z, const TZ *__restrict__ dold,
           TZ *__restrict__ dnew, double *__restrict__ q, double *__restrict__ x,
           const double *__restrict__ S, double *__restrict__ part) {
    __shared__ double D[3][18][66];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int bid = xcd_order(blockIdx.x, gridDim.x);
    const int tx = bid % g.tx, ty = (bid / g.tx) % g.ty, ch = bid / (g.tx * g.ty);
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = tx * 64 + li, j = ty * 16 + lj;
    const int k0 = ch * g.zc, k1 = min(k0 + g.zc, nz);
    const bool in = i < nx && j < ny;
    const double beta = S[S_BETA], zm = S[S_ZMEAN], ax = S[S_ALPHA];
    // x null: the previous direction is not applied here (the caller
    // defers it)
    const bool use_old = beta != 0.0, upd = x && ax != 0.0;
    const bool rd_old = use_old || upd;
    const long long sy = nx, sz = (long long)nx * ny;
    auto wrapn = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };
    auto wrapk = [&](int k) { return wrapn(k, nz); };
    // own slot (li+1, lj+1) holds cell (wrap i, wrap j); threads 0..163 also
    // fill one halo slot: rows -1 / 16 (66 slots each), columns -1 / 64
    const long long own = wrapn(i, nx) + sy * wrapn(j, ny);
    const int t = li + 64 * lj;
    int hx = 0, hy = 0;
    if (t < 66) { hx = t - 1; hy = -1; }
    else if (t < 132) { hx = t - 67; hy = 16; }
    else if (t < 148) { hx = -1; hy = t - 132; }
    else { hx = 64; hy = t - 148; }
    const bool halo = t < 164;
    const long long hcol = wrapn(tx * 64 + hx, nx) + sy * wrapn(ty * 16 + hy, ny);
    auto dval = [&](TZ zr, TZ dr) -> double {
        const double v = (double)zr - zm;
        return (double)(TZ)(use_old ? v + beta * (double)dr : v);
    };
    // in-plane operator offsets of the own cell
    const long long dxm = in ? (long long)(wrap_m(i, nx) - i) : 0;
    const long long dym = in ? (long long)(wrap_m(j, ny) - j) * sy : 0;

    // prologue: d_new at k0-1 (register), k0 (register + LDS); raw of k0+1;
    // dx0: d_old of the own cell at the plane being output
    double d_m = 0.0, d_0 = 0.0;
    TZ zo1 = 0, zh1 = 0;
    TZ do1 = 0, dh1 = 0, dx0 = 0;
    {
        const long long pm = sz * wrapk(k0 - 1), p0 = sz * wrapk(k0), p1 = sz * wrapk(k0 + 1);
        d_m = dval(z[own + pm], rd_old ? dold[own + pm] : (TZ)0);
        dx0 = rd_old ? dold[own + p0] : (TZ)0;
        d_0 = dval(z[own + p0], dx0);
        D[k0 % 3][lj + 1][li + 1] = d_0;
        if (halo) D[k0 % 3][hy + 1][hx + 1] = dval(z[hcol + p0], rd_old ? dold[hcol + p0] : (TZ)0);
        zo1 = z[own + p1];
        if (rd_old) do1 = dold[own + p1];
        if (halo) {
            zh1 = z[hcol + p1];
            if (rd_old) dh1 = dold[hcol + p1];
        }
    }
    double czm = in ? op_z(L, own + sz * wrapk(k0 - 1), wrapk(k0 - 1)) : 0.0;
    __syncthreads();
    double dq = 0.0;
    for (int s = k0; s < k1; ++s) {
        // requests: raw z / d_old of plane s+2, operator of plane s
        TZ zo2 = 0, zh2 = 0;
        TZ do2 = 0, dh2 = 0;
        if (s + 1 < k1) {
            const long long p2 = sz * wrapk(s + 2);
            zo2 = z[own + p2];
            if (rd_old) do2 = dold[own + p2];
            if (halo) {
                zh2 = z[hcol + p2];
                if (rd_old) dh2 = dold[hcol + p2];
            }
        }
        const long long c = own + sz * s;
        double cxw = 0.0, cxe = 0.0, cys = 0.0, cyn = 0.0, czp = 0.0, dgc = 0.0;
        if (in) {
            cxw = op_x(L, c + dxm, wrap_m(i, nx)); cxe = op_x(L, c, i);
            cys = op_y(L, c + dym, wrap_m(j, ny)); cyn = op_y(L, c, j);
            czp = op_z(L, c, s);
            dgc = outer_diag<DGI>(L, c, cxw, cxe, cys, cyn, czm, czp);
        }
        // d_new of plane s+1 into LDS
        const int wb = (s + 1) % 3, rb = s % 3;
        const double d_p = dval(zo1, do1);
        D[wb][lj + 1][li + 1] = d_p;
        if (halo) D[wb][hy + 1][hx + 1] = dval(zh1, dh1);
        if (in) {
            const double sx = cxw * D[rb][lj + 1][li] + cxe * D[rb][lj + 1][li + 2];
            const double syv = cys * D[rb][lj][li + 1] + cyn * D[rb][lj + 2][li + 1];
            const double szv = czm * d_m + czp * d_p;
            const double qv = dgc * d_0 - ((sx + syv) + szv);
            // non-temporal: d, q and x are read again only by the next
            // outer kernels, after the V-cycle has streamed its own arrays;
            // keeping them out of the caches leaves those to the V-cycle's
            // (with k_update's: 7.59 -> 7.46 ms a 256^3 mixed solve)
            __builtin_nontemporal_store((TZ)d_0, dnew + c);
            __builtin_nontemporal_store(qv, q + c);
            dq += d_0 * qv;
            if (upd) __builtin_nontemporal_store(x[c] + ax * (double)dx0, x + c);
        }
        d_m = d_0; d_0 = d_p; czm = czp;
        dx0 = do1;
        zo1 = zo2; do1 = do2; zh1 = zh2; dh1 = dh2;
        __syncthreads();
    }
    block_partial(dq, part);
}


// r -= alpha q ; partial sum r^2 ; optional FP32 copy of r (x += alpha d
// follows in the next k_dq_tiled, or in k_axpy_d after the last iteration)
template <class TB>
__global__ void __launch_bounds__(256)
k_update(long long n, double *__restrict__ r, const double *__restrict__ q, TB *__restrict__ bt,
         const double *__restrict__ S, double *__restrict__ part) {
    const double alpha = S[S_ALPHA];
    double rr = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        // r and q non-temporal; the FP32 copy of r, which the next V-cycle's
        // first sweep reads at once, is cached (that sweep: 73.7 -> 63.8 us
        // at 256^3)
        const double v = __builtin_nontemporal_load(r + c) - alpha * __builtin_nontemporal_load(q + c);
        __builtin_nontemporal_store(v, r + c);
        if (bt) bt[c] = (TB)v;
        rr += v * v;
    }
    block_partial(rr, part);
}

// x += alpha d with alpha = S[S_ALPHA] (the last iteration's deferred
// solution update)
template <class TD>
__global__ void __launch_bounds__(256)
k_axpy_d(long long n, double *__restrict__ x, const TD *__restrict__ d, const double *__restrict__ S) {
    const double alpha = S[S_ALPHA];
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256)
        x[c] = __fma_rn(alpha, (double)d[c], x[c]);
}

// The residual of x = 0: r = b and its FP32 copy, partial sums of r^2 (0)
// and of b^2 (1), which are the same.
template <class TB>
__global__ void __launch_bounds__(256)
k_resid_zero(long long n, const double *__restrict__ b, double *__restrict__ r, TB *__restrict__ bt,
             double *__restrict__ part) {
    double rr = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        const double v = b[c];
        r[c] = v;
        if (bt) bt[c] = (TB)v;
        rr += v * v;
    }
    block_partial(rr, part);
    block_partial(rr, part + kRedBlocks);
}

// the outer iteration's scalars at its start: first iteration, no pending
// x update, not stopped
__global__ void __launch_bounds__(64) k_outer_start(double *__restrict__ S) {
    if (threadIdx.x == 0) {
        S[S_FIRST] = 1.0;
        S[S_ALPHA] = 0.0;
        S[S_STOP] = 0.0;
    }
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

// the beta step from k_dot_rz's partials: zmean, rho = r.(z - zmean),
// beta = rho / rho_old
__global__ void __launch_bounds__(1024)
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

// the alpha step from k_dq's partials
__global__ void __launch_bounds__(1024)
k_scal_alpha(const double *__restrict__ part, int nb, double *__restrict__ S, int slot) {
    const double dq = sum_partials(part, nb);
    if (threadIdx.x == 0) {
        S[S_DQ] = dq;
        S[S_ALPHA] = S[S_RZ] / dq;
        S[S_AH + slot] = S[S_ALPHA];
    }
}

// x += alpha_j d_j for the directions of ring slots first, first+1, ...
// (count of them, modulo R, alphas S[S_AH + slot]) one after another in
// that order, each an fma as k_dq_tiled's and k_check_tiled's update: x is
// bit for bit what applying them one iteration at a time gives.  W cells a
// thread (16-byte loads of the directions when W = 16 / sizeof(T); n a
// multiple of W); the directions are read once (non-temporal).
template <class T, int W>
__global__ void __launch_bounds__(256)
k_flush(long long n, double *__restrict__ x, const T *__restrict__ ring, const double *__restrict__ S,
        int first, int count, int R) {
    using Vt = T __attribute__((ext_vector_type(W)));
    const long long nv = n / W;
    for (long long e = blockIdx.x * 256LL + threadIdx.x; e < nv; e += (long long)gridDim.x * 256) {
        Vt dv[kRingSlots];
#pragma unroll
        for (int m = 0; m < kRingSlots; ++m)
            if (m < count)
                dv[m] = __builtin_nontemporal_load(
                    reinterpret_cast<const Vt *>(ring + (long long)((first + m) % R) * n) + e);
        double v[W];
#pragma unroll
        for (int w = 0; w < W; ++w) v[w] = x[e * W + w];
#pragma unroll
        for (int m = 0; m < kRingSlots; ++m) {
            if (m < count) {
                const double al = S[S_AH + (first + m) % R];
#pragma unroll
                for (int w = 0; w < W; ++w) v[w] = __fma_rn(al, (double)dv[m][w], v[w]);
            }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) x[e * W + w] = v[w];
    }
}

// b -= mean(b), the mean from the partial sums of k_zsum
__global__ void __launch_bounds__(256)
k_sub_mean(long long n, double *__restrict__ b, const double *__restrict__ part, int nb) {
    __shared__ double m_s;
    const double bs = sum_partials(part, nb);
    if (threadIdx.x == 0) m_s = bs / (double)n;
    __syncthreads();
    const double m = m_s;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) b[c] -= m;
}

__global__ void __launch_bounds__(1024)
k_scal_rr(const double *__restrict__ part, int nb, int with_bb, double *__restrict__ S,
          Mailbox *mb = nullptr, unsigned long long seq = 0, const int *flags = nullptr,
          double N = 1.0, double tol = 0.0) {
    const double rr = sum_partials(part, nb);
    double bb = 0.0;
    if (with_bb) bb = sum_partials(part + kRedBlocks, nb);
    if (threadIdx.x == 0) {
        S[S_RR] = rr;
        if (with_bb) S[S_BB] = bb;
        // the outer iteration's stopping test, sqrt(rr / N) < tol (tol > 0),
        // for the first sweep already queued behind this kernel; the host
        // takes the same decision from the mailbox
        const double stop = (tol > 0.0 && sqrt(rr / N) < tol) ? 1.0 : 0.0;
        if (tol > 0.0) S[S_STOP] = stop;
        if (mb) {
            // the values, then seq with a system-scope release: a host that
            // reads seq (acquire) sees them
            mb->rr = rr;
            mb->bb = bb;
            mb->flags = flags ? (double)*flags : 0.0;
            mb->pad = stop;
            __hip_atomic_store(&mb->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
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

// Setup in one pass (z-marching, launch with zgrid()): from the 8-slot A
// and the padded initial p, write the FP64 operator (couplings, diagonal),
// the folded right-hand side b, the interior iterate x, the initial
// residual r = A8 - A7 p + sum A_nb p_nb (the reference's own expression
// on the padded p, MGsolver.f90:97-100), and in mixed mode the FP32 copies
// of the operator and of r.  Partials: 0 = sum r^2, 1 = sum b^2.  flags as
// k_pack, plus bit3: some A7 differs from the coupling sum outer_diag forms.
// W (wx, wy, wz not null): every row is first multiplied by the weight
// w = wx[i] wy[j] wz[k] (normalised cell volumes), which makes the system
// of a stretched grid symmetric: SetupPoisson's couplings of a face are
// 2 dt mask / (dx dxh (rho + rho)) with the width dx of the row's own cell
// (modules.f90:2737-2742), so dx(i) A2(i) = dx(i+1) A1(i+1).  b, r and the
// operator are those of the weighted rows; the b^2 partial is unweighted.
//
// The rows come from a row source: RowsA reads the caller's 8-slot A;
// RowsF assembles each row from the fields (SetupPoisson + Poisson_BCs +
// the solid rule, the expressions the assembly kernel uses), so that A is
// never written.  A source gives the row of a cell, and for the symmetry
// check the -x / -y / -z coupling of its +x / +y / +z neighbour (wrapped on
// a periodic axis; only asked for where that neighbour exists).
struct CellAt {
    int i, j, k;
    long long c, q;     // interior and padded index
    bool lastx, lasty, lastz;
};
struct RowsA {
    static constexpr bool kDiagStored = true;
    const double *A;
    long long n;
    double a4n;   // the -z coupling of the next plane, carried down the march
    __device__ void begin(long long c0) { a4n = A[4 * n + c0]; }
    __device__ int row(const CellAt &e, double a[8]) {
#pragma unroll
        for (int s = 0; s < 8; ++s) a[s] = s == 4 ? a4n : A[s * n + e.c];
        return 0;
    }
    __device__ double nb_x(const CellAt &e, long long ex) const { return A[e.c + ex]; }
    __device__ double nb_y(const CellAt &e, long long ey) const { return A[2 * n + e.c + ey]; }
    // the next plane's -z coupling (wrapped): read every step and carried
    __device__ double nb_z(const CellAt &e, long long ez) {
        a4n = A[4 * n + e.c + ez];
        return a4n;
    }
};
struct RowsF {
    static constexpr bool kDiagStored = true;
    RowK R;
    int nx, ny, nz, ng;
    long long Py, Pz;
    double raw[3];   // the row's own +x, +y, +z couplings before Poisson_BCs
    __device__ void begin(long long) {}
    __device__ int row(const CellAt &e, double a[8]) {
        return setup_row<true>(R, nx, ny, nz, ng, Py, Pz, e.i, e.j, e.k, e.q, a, raw);
    }
    // For boundary codes 0-4 a neighbour's lower coupling is SetupPoisson's
    // expression: Poisson_BCs changes only couplings across the domain
    // boundary (codes 3/4), which this check never asks for, and the solid
    // rule changes only the diagonal.  Inside the domain that expression
    // differs from the row's own coupling of the same face only in the width
    // of the cell (face_coupling(dt, mask, d, dh, r_nb, r_c) with d the
    // neighbour's width and r_nb + r_c added the other way round, which
    // rounds the same): with equal widths it is that coupling, bit for bit,
    // and is not evaluated again.  Across a periodic boundary the neighbour's
    // coupling reads the ghost mask and densities, so it is always formed.
    __device__ double nb_x(const CellAt &e, long long) const {
        const int I = (e.lastx ? 0 : e.i + 1) + ng;
        if (!e.lastx && R.ax[I] == R.ax[e.i + ng]) return raw[0];
        const long long q = e.lastx ? e.q - (nx - 1) : e.q + 1;
        return face_coupling(R.dt, R.um[q - 1], R.ax[I], R.ax[nx + 2 * ng + I - 1], R.rho[q - 1], R.rho[q]);
    }
    __device__ double nb_y(const CellAt &e, long long) const {
        const int J = (e.lasty ? 0 : e.j + 1) + ng;
        if (!e.lasty && R.ay[J] == R.ay[e.j + ng]) return raw[1];
        const long long q = e.lasty ? e.q - (long long)(ny - 1) * Py : e.q + Py;
        return face_coupling(R.dt, R.vm[q - Py], R.ay[J], R.ay[ny + 2 * ng + J - 1], R.rho[q - Py], R.rho[q]);
    }
    __device__ double nb_z(const CellAt &e, long long) {
        const int K = (e.lastz ? 0 : e.k + 1) + ng;
        if (!e.lastz && R.az[K] == R.az[e.k + ng]) return raw[2];
        const long long q = e.lastz ? e.q - (long long)(nz - 1) * Pz : e.q + Pz;
        return face_coupling(R.dt, R.wm[q - Pz], R.az[K], R.az[nz + 2 * ng + K - 1], R.rho[q - Pz], R.rho[q]);
    }
};

// Rows from face values (the map of k_faces_to_A in pps_adjoint.hip, the
// same products and the same sum, bit for bit): A1..A6 the weighted values
// w c of the cell's six faces (0 through a wall), A7 their sum, A8 = b.
// The system is symmetric by construction (a neighbour's lower coupling is
// the same product as the row's upper one), so the check gets the row's
// own value.  It is also singular, and every A7 is the coupling sum in
// outer_diag's order, by construction: the diagonals are formed from the
// couplings wherever they are used, and none is stored.  The -z coupling
// of a row is the +z coupling of the row below, carried down the march.
struct RowsC {
    static constexpr bool kDiagStored = false;
    FacesK F;
    int nx, ny, nz, px, py, pz;
    double up[3];
    bool have_dn;   // up[2] holds the -z coupling of the next row
    __device__ void begin(long long) { have_dn = false; }
    __device__ int row(const CellAt &e, double a[8]) {
#pragma clang fp contract(off)
        const long long n = (long long)nx * ny * nz, sy = nx, sz = (long long)nx * ny;
        const double *cx = F.cf, *cy = F.cf + n, *cz = F.cf + 2 * n;
        const int i = e.i, j = e.j, k = e.k;
        const long long c = e.c;
        a[0] = i > 0 ? F.wx[i - 1] * cx[c - 1] : (px ? F.wx[nx - 1] * cx[c + nx - 1] : 0.0);
        a[1] = (i < nx - 1 || px) ? F.wx[i] * cx[c] : 0.0;
        a[2] = j > 0 ? F.wy[j - 1] * cy[c - sy] : (py ? F.wy[ny - 1] * cy[c + (long long)(ny - 1) * sy] : 0.0);
        a[3] = (j < ny - 1 || py) ? F.wy[j] * cy[c] : 0.0;
        a[4] = have_dn ? up[2]
                       : (k > 0 ? F.wz[k - 1] * cz[c - sz] : (pz ? F.wz[nz - 1] * cz[c + (long long)(nz - 1) * sz] : 0.0));
        a[5] = (k < nz - 1 || pz) ? F.wz[k] * cz[c] : 0.0;
        have_dn = true;
        a[6] = ((((a[0] + a[1]) + a[2]) + a[3]) + a[4]) + a[5];
        a[7] = F.b ? F.b[c] : 0.0;
        up[0] = a[1];
        up[1] = a[3];
        up[2] = a[5];
        return 0;
    }
    __device__ double nb_x(const CellAt &, long long) const { return up[0]; }
    __device__ double nb_y(const CellAt &, long long) const { return up[1]; }
    __device__ double nb_z(const CellAt &, long long) { return up[2]; }
};

template <class Src>
__global__ void __launch_bounds__(256)
k_setup(PackK d, int zc, Src src, const double *__restrict__ p,
        double *__restrict__ cx, double *__restrict__ cy, double *__restrict__ cz,
        double *__restrict__ dg, double *__restrict__ bo, double *__restrict__ x,
        double *__restrict__ r, float *__restrict__ f32, float *__restrict__ r32,
        double *__restrict__ part, int *__restrict__ flags, const double *__restrict__ wx,
        const double *__restrict__ wy, const double *__restrict__ wz) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * zc;
    const int k1 = min(k0 + zc, d.nz);
    const long long n = (long long)d.nx * d.ny * d.nz;
    double rr = 0.0, bb = 0.0;
    int f = 0;
    if (i < d.nx && j < d.ny) {
        const long long sy = d.nx, sz = (long long)d.nx * d.ny;
        const long long Py = d.X, Pz = (long long)d.X * d.Y;
        const bool lastx = i == d.nx - 1, lasty = j == d.ny - 1;
        const long long ex_ = lastx ? -(long long)(d.nx - 1) : 1;
        const long long ey_ = lasty ? -(long long)(d.ny - 1) * sy : sy;
        // the column's p at k-1 and k, and the -z coupling of plane k (the
        // symmetry check of plane k-1 loaded it), carried down the march
        const long long q0 = (i + d.ng) + Py * (j + d.ng) + Pz * (k0 + d.ng);
        double pb = p[q0 - Pz], pc = p[q0];
        src.begin(i + sy * j + sz * k0);
        for (int k = k0; k < k1; ++k) {
            const long long c = i + sy * j + sz * k;
            const long long q = (i + d.ng) + Py * (j + d.ng) + Pz * (k + d.ng);
            const bool lastz = k == d.nz - 1;
            const CellAt e{i, j, k, c, q, lastx, lasty, lastz};
            double a[8];
            if (src.row(e, a)) f |= 1;
            bool nan = false;
#pragma unroll
            for (int s = 0; s < 8; ++s) nan |= a[s] != a[s];
            if (nan || a[6] < 1e-55) f |= 1;
            // row weight and the weights of the +x, +y, +z neighbours
            double wc = 1.0, wxn = 1.0, wyn = 1.0, wzn = 1.0;
            if (wx) {
                const double yz = wy[j] * wz[k];
                wc = wx[i] * yz;
                wxn = wx[lastx ? 0 : i + 1] * yz;
                wyn = wx[i] * wy[lasty ? 0 : j + 1] * wz[k];
                wzn = wx[i] * wy[j] * wz[k == d.nz - 1 ? 0 : k + 1];
#pragma unroll
                for (int s = 0; s < 8; ++s) a[s] *= wc;
            }
            const double pw = p[q - 1], pe = p[q + 1], ps = p[q - Py], pn = p[q + Py], pt = p[q + Pz];
            const double rv = a[7] - a[6] * pc + a[0] * pw + a[1] * pe + a[2] * ps + a[3] * pn +
                              a[4] * pb + a[5] * pt;
            double rhs = a[7], kept = 0.0;
            double vx = a[1], vy = a[3], vz = a[5];
            if (!d.px) {
                if (i == 0) rhs += a[0] * pw; else kept += a[0];
                if (lastx) { rhs += a[1] * pe; vx = 0.0; } else kept += a[1];
            } else {
                kept += a[0] + a[1];
            }
            if (!d.py) {
                if (j == 0) rhs += a[2] * ps; else kept += a[2];
                if (lasty) { rhs += a[3] * pn; vy = 0.0; } else kept += a[3];
            } else {
                kept += a[2] + a[3];
            }
            if (!d.pz) {
                if (k == 0) rhs += a[4] * pb; else kept += a[4];
                if (lastz) { rhs += a[5] * pt; vz = 0.0; } else kept += a[5];
            } else {
                kept += a[4] + a[5];
            }
            if (!lastx || d.px) { if (asym(src.nb_x(e, ex_) * wxn, a[1])) f |= 2; }
            if (!lasty || d.py) { if (asym(src.nb_y(e, ey_) * wyn, a[3])) f |= 2; }
            {
                // the +z neighbour's -z coupling (the next plane's a[4])
                const double zn = src.nb_z(e, lastz ? -(long long)(d.nz - 1) * sz : sz);
                if ((!lastz || d.pz) && asym(zn * wzn, a[5])) f |= 2;
            }
            if (fabs(a[6] - kept) > 1e-12 * fabs(a[6])) f |= 4;
            {
                // is A7 the coupling sum of the stored operator, bit for bit?
                const double vxm = (i == 0 && !d.px) ? 0.0 : a[0];
                const double vym = (j == 0 && !d.py) ? 0.0 : a[2];
                const double vzm = (k == 0 && !d.pz) ? 0.0 : a[4];
                if (a[6] != ((((vxm + vx) + vym) + vy) + vzm) + vz) f |= 8;
            }
            if (cx) {
                cx[c] = vx; cy[c] = vy; cz[c] = vz;
                if (Src::kDiagStored) dg[c] = a[6];
            }
            bo[c] = rhs;
            x[c] = pc;
            r[c] = rv;
            if (f32) {
                f32[c] = (float)vx; f32[n + c] = (float)vy; f32[2 * n + c] = (float)vz;
                if (Src::kDiagStored) f32[3 * n + c] = (float)a[6];
                r32[c] = (float)rv;
            }
            rr += rv * rv;
            bb += wx ? (rhs / wc) * (rhs / wc) : rhs * rhs;
            pb = pc;
            pc = pt;
        }
    }
    if (f) atomicOr(flags, f);
    block_partial(rr, part);
    block_partial(bb, part + kRedBlocks);
}

// project_velocity's face update (paris.f90:899-909), operand order and
// rounding as the reference (no contraction)
__device__ __forceinline__ double proj_face(double vel, double mask, double dh, double pe, double pc,
                                            double re, double rc, double dt) {
#pragma clang fp contract(off)
    return vel - dt * (2.0 * mask / dh) * (pe - pc) / (re + rc);
}

// Final check fused with the write-back, on LDS tiles (64 x 16 cells, zc
// planes a block; launch with tgrid()): x_f = x + alpha d, the last
// iteration's deferred update (alpha = S[S_ALPHA]; dd null: x_f = x), then
// r = b - M x_f with the FP64 operator, partial sum r^2, and p = x_f on the
// interior with the first ghost layer wrapped in periodic directions.  x is
// left as it was (a restart applies the update first).  x_f of each plane
// (own cells plus a one-cell halo ring, coordinates wrapped) goes to one of
// three LDS planes: the in-plane neighbours come from there, the vertical
// ones from registers, and the raw x / d of plane s+2 are requested at the
// start of step s.  PROJ: also the velocity correction of project_velocity
// into P.uo / P.vo / P.wo from x_f (the +x / +y neighbours from LDS, the
// +z one of the march; on a periodic axis the wrapped value is the ghost
// the reference reads); faces left alone are copied.  Without PROJ the
// registers are held to 64 a thread, two blocks a CU, so that one block's
// loads overlap the other's plane step (256^3: 230 -> 214 us, a few bytes
// of spill); the PROJ form would spill far more and keeps its registers.
template <bool DGI, bool PROJ, class TD>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(PROJ ? 1 : 8)))
k_check_tiled(OpK L, PackK d, TGrid g, const double *__restrict__ b, const double *__restrict__ x,
              const TD *__restrict__ dd, const double *__restrict__ S, double *__restrict__ p,
              double *__restrict__ part, ProjK P, const double *__restrict__ wx,
              const double *__restrict__ wy, const double *__restrict__ wz) {
    __shared__ double X[3][18][66];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int bid = xcd_order(blockIdx.x, gridDim.x);
    const int tx = bid % g.tx, ty = (bid / g.tx) % g.ty, ch = bid / (g.tx * g.ty);
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = tx * 64 + li, j = ty * 16 + lj;
    const int k0 = ch * g.zc, k1 = min(k0 + g.zc, nz);
    const bool in = i < nx && j < ny;
    const double al = dd ? S[S_ALPHA] : 0.0;
    const bool upd = dd != nullptr;
    const long long sy = nx, sz = (long long)nx * ny;
    const long long Py = d.X, Pz = (long long)d.X * d.Y;
    auto wrapn = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };
    auto wrapk = [&](int k) { return wrapn(k, nz); };
    // own slot (li+1, lj+1) holds cell (wrap i, wrap j); threads 0..163 also
    // fill one halo slot: rows -1 / 16 (66 slots each), columns -1 / 64
    const long long own = wrapn(i, nx) + sy * wrapn(j, ny);
    const int t = li + 64 * lj;
    int hx = 0, hy = 0;
    if (t < 66) { hx = t - 1; hy = -1; }
    else if (t < 132) { hx = t - 67; hy = 16; }
    else if (t < 148) { hx = -1; hy = t - 132; }
    else { hx = 64; hy = t - 148; }
    const bool halo = t < 164;
    const long long hcol = wrapn(tx * 64 + hx, nx) + sy * wrapn(ty * 16 + hy, ny);
    auto ldd = [&](long long c) -> TD { return upd ? dd[c] : (TD)0; };
    auto xf = [&](double xv, TD dv) -> double { return upd ? __fma_rn(al, (double)dv, xv) : xv; };
    const long long dxm = in ? (long long)(wrap_m(i, nx) - i) : 0;
    const long long dym = in ? (long long)(wrap_m(j, ny) - j) * sy : 0;

    // prologue: x_f at k0-1 (register), k0 (register + LDS); raw of k0+1
    double x_m = 0.0, x_0 = 0.0, xo1 = 0.0, xh1 = 0.0;
    TD do1 = 0, dh1 = 0;
    {
        const long long pm = sz * wrapk(k0 - 1), p0 = sz * wrapk(k0), p1 = sz * wrapk(k0 + 1);
        x_m = xf(x[own + pm], ldd(own + pm));
        x_0 = xf(x[own + p0], ldd(own + p0));
        X[k0 % 3][lj + 1][li + 1] = x_0;
        if (halo) X[k0 % 3][hy + 1][hx + 1] = xf(x[hcol + p0], ldd(hcol + p0));
        xo1 = x[own + p1];
        do1 = ldd(own + p1);
        if (halo) {
            xh1 = x[hcol + p1];
            dh1 = ldd(hcol + p1);
        }
    }
    double czm = in ? op_z(L, own + sz * wrapk(k0 - 1), wrapk(k0 - 1)) : 0.0;
    // PROJ: the cell's density, carried up the march
    double rc = (PROJ && in) ? P.rho[(i + d.ng) + Py * (j + d.ng) + Pz * (k0 + d.ng)] : 0.0;
    __syncthreads();
    double rr = 0.0;
    for (int s = k0; s < k1; ++s) {
        // requests: raw x / d of plane s+2, operator of plane s
        double xo2 = 0.0, xh2 = 0.0;
        TD do2 = 0, dh2 = 0;
        if (s + 1 < k1) {
            const long long p2 = sz * wrapk(s + 2);
            xo2 = x[own + p2];
            do2 = ldd(own + p2);
            if (halo) {
                xh2 = x[hcol + p2];
                dh2 = ldd(hcol + p2);
            }
        }
        const long long c = own + sz * s;
        double cxw = 0.0, cxe = 0.0, cys = 0.0, cyn = 0.0, czp = 0.0, dgc = 0.0, bv = 0.0;
        if (in) {
            cxw = op_x(L, c + dxm, wrap_m(i, nx)); cxe = op_x(L, c, i);
            cys = op_y(L, c + dym, wrap_m(j, ny)); cyn = op_y(L, c, j);
            czp = op_z(L, c, s);
            dgc = outer_diag<DGI>(L, c, cxw, cxe, cys, cyn, czm, czp);
            bv = b[c];
        }
        // x_f of plane s+1 into LDS
        const int wb = (s + 1) % 3, rb = s % 3;
        const double x_p = xf(xo1, do1);
        X[wb][lj + 1][li + 1] = x_p;
        if (halo) X[wb][hy + 1][hx + 1] = xf(xh1, dh1);
        if (in) {
            const double xe = X[rb][lj + 1][li + 2], xn = X[rb][lj + 2][li + 1];
            double sm = cxw * X[rb][lj + 1][li] + cxe * xe;
            sm += cys * X[rb][lj][li + 1] + cyn * xn;
            sm += czm * x_m + czp * x_p;
            double v = bv - (dgc * x_0 - sm);
            if (wx) v /= wx[i] * wy[j] * wz[s];   // the caller's (unweighted) row
            rr += v * v;
            const long long q = (i + d.ng) + Py * (j + d.ng) + Pz * (s + d.ng);
            if constexpr (PROJ) {
                const double re = P.rho[q + 1], rn = P.rho[q + Py], rt = P.rho[q + Pz];
                const double ui = P.u[q], vi = P.v[q], wi = P.w[q];
                P.uo[q] = i < nx - P.lx ? proj_face(ui, P.um[q], P.dxh[i + d.ng], xe, x_0, re, rc, P.dt) : ui;
                P.vo[q] = j < ny - P.ly ? proj_face(vi, P.vm[q], P.dyh[j + d.ng], xn, x_0, rn, rc, P.dt) : vi;
                P.wo[q] = s < nz - P.lz ? proj_face(wi, P.wm[q], P.dzh[s + d.ng], x_p, x_0, rt, rc, P.dt) : wi;
                rc = rt;
            }
            p[q] = x_0;
            if (d.px && i == d.nx - 1) p[q - d.nx] = x_0;
            if (d.px && i == 0) p[q + d.nx] = x_0;
            if (d.py && j == d.ny - 1) p[q - (long long)d.ny * Py] = x_0;
            if (d.py && j == 0) p[q + (long long)d.ny * Py] = x_0;
            if (d.pz && s == d.nz - 1) p[q - (long long)d.nz * Pz] = x_0;
            if (d.pz && s == 0) p[q + (long long)d.nz * Pz] = x_0;
        }
        x_m = x_0; x_0 = x_p; czm = czp;
        xo1 = xo2; do1 = do2; xh1 = xh2; dh1 = dh2;
        __syncthreads();
    }
    block_partial(rr, part);
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

// x *= wx[i] wy[j] wz[k] on the compact interior (the adjoint of a
// row-weighted system: lam = W mu)
__global__ void __launch_bounds__(256)
k_weight(PackK d, double *__restrict__ x, const double *__restrict__ wx, const double *__restrict__ wy,
         const double *__restrict__ wz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    x[c] *= wx[i] * wy[j] * wz[k];
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
    T *cx, *cy, *cz, *dg, *b;
    T *x, *xt;   // current iterate and the other buffer of the out-of-place sweeps
    // coarsening to the next level: factor 1 or 2 per axis, and whether it
    // is the regular one (every factor 2, every size even: the fused
    // restriction, prolongation and tail kernels apply)
    Coarsening g{2, 2, 2, 1, 1, 1};
    bool reg = true;
    LevelK<T> k() const { return LevelK<T>{nx, ny, nz, cx, cy, cz, dg}; }
};

template <class T> struct Hier {
    std::vector<Level<T>> lv;
    std::vector<DevBuf> bufs;
    DevBuf minv, work;
    int nb = 0;
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
    int exact = 0;   // every A7 is the coupling sum of outer_diag, bit for bit
    const double *built_for = nullptr;
    const double *built_w[3] = {nullptr, nullptr, nullptr};   // face weights (face-value systems)
    int built_mixed = 0;   // precision of the hierarchy last built for built_for
    Hier<double> h64;
    Hier<float> h32;
    DevBuf b32;      // FP32 right-hand side of the mixed V-cycle
    DevBuf ring;     // the outer CG's directions, kRingSlots slots (deferred x updates)
    int mixed_built = 0, fp64_built = 0;
    // the outer operator is read from the caller's A (mixed mode: the FP64
    // compact operator is not written)
    int op_in_A = 0;
    int have_r = 0;  // the setup pass already wrote r (and its norms) for x
    // deferred setup: the setup pass's flags are still to be read (at the
    // next synchronisation), and the coarse hierarchy is being built on
    // the handle's side stream (the solve stream waits for sev[1] before
    // the first coarse-level kernel)
    int flags_pending = 0, coarse_pending = 0;
    size_t lds_max = 65536;  // LDS bytes per workgroup on this device
    int cus = 256;           // compute units on this device
    // no-allocation solve: every device buffer is a piece of this arena
    // (the scratch array) instead of a hipMalloc
    Arena *arena = nullptr;
};

void fast_invalidate(FastState *f) {
    if (f) f->built_for = nullptr;
}

void fast_destroy(FastState *f) {
    if (!f) return;
    f->op.release(); f->vec.release(); f->red.release(); f->flags.release(); f->b32.release();
    f->ring.release();
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
// Blocks of the grid-stride streaming kernels (one partial each): 32 per CU
// at most (k_update at 256^3: 81.3 us with 2048 blocks, 75.7 us with 8192;
// the r.r kernel then sums 8192 partials, 6.3 -> 7.2 us).
static int nblk(long long n) {
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

// Level sizes of the fast hierarchy.  From each level every axis is halved
// (ceil(n / 2) coarse cells; on an odd size the last coarse cell has a
// single child) except
//   - an axis whose cells are more than 1.5 times as long as along the
//     shortest one (semi-coarsening: the strongly coupled directions are
//     coarsened first, which makes the coarse cells more nearly cubic);
//   - a periodic axis of odd size (an odd periodic ring cannot be coloured
//     red-black), and an axis of one cell.
// A level with a periodic axis of odd size cannot be smoothed, so it is
// the coarsest.  Coarsening stops at 64 cells or fewer; the coarsest level
// must have at most 512 cells (its exact dense solve).  lengths: the box
// (cell sizes length / n; a non-positive length counts as n).
static int plan_levels(const Dims &d, const double lengths[3], std::vector<int> &dims) {
    int n[3] = {d.nx, d.ny, d.nz};
    // z-marching reductions keep one partial slot per (64 x 4) column block
    if ((long long)((n[0] + 63) / 64) * ((n[1] + 3) / 4) > kRedBlocks) return PPS_ERR_BAD_SIZE;
    double h[3];
    for (int a = 0; a < 3; ++a) h[a] = lengths[a] > 0.0 ? lengths[a] / n[a] : 1.0;
    dims.clear();
    for (int a = 0; a < 3; ++a) dims.push_back(n[a]);
    while ((long long)n[0] * n[1] * n[2] > 64) {
        bool podd = false;
        for (int a = 0; a < 3; ++a) podd |= d.per[a] && (n[a] & 1);
        if (podd) break;
        const double hmin = std::fmin(h[0], std::fmin(h[1], h[2]));
        bool any = false;
        int f[3];
        for (int a = 0; a < 3; ++a) {
            f[a] = (n[a] > 1 && h[a] <= 1.5 * hmin) ? 2 : 1;
            any |= f[a] == 2;
        }
        if (!any) break;
        for (int a = 0; a < 3; ++a) {
            if (f[a] == 2) { n[a] = (n[a] + 1) / 2; h[a] *= 2.0; }
            dims.push_back(n[a]);
        }
    }
    if ((long long)n[0] * n[1] * n[2] > 512) return PPS_ERR_BAD_SIZE;
    return PPS_OK;
}

template <class T>
static int hier_alloc(Hier<T> &H, const std::vector<int> &dims, bool own_fine, Arena *arena) {
    const int nl = (int)dims.size() / 3;
    H.lv.resize(nl);
    H.bufs.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Level<T> &L = H.lv[l];
        L.nx = dims[3 * l]; L.ny = dims[3 * l + 1]; L.nz = dims[3 * l + 2];
        L.n = (long long)L.nx * L.ny * L.nz;
        if (l + 1 < nl) {
            const int cx = dims[3 * l + 3], cy = dims[3 * l + 4], cz = dims[3 * l + 5];
            L.g = Coarsening{cx == L.nx ? 1 : 2, cy == L.ny ? 1 : 2, cz == L.nz ? 1 : 2, cx, cy, cz};
            L.reg = L.g.fx == 2 && L.g.fy == 2 && L.g.fz == 2 && !(L.nx & 1) && !(L.ny & 1) && !(L.nz & 1);
        }
        const bool own = l > 0 || own_fine;
        // coefficient planes (4) + b + x + xt, or only xt when the fine
        // operator, right-hand side and iterate are borrowed from the outer
        // iteration
        const int planes = own ? 7 : 1;
        PPS_TRY(H.bufs[l].ensure(sizeof(T) * planes * L.n, arena));
        T *base = H.bufs[l].template as<T>();
        if (own) {
            L.cx = base; L.cy = base + L.n; L.cz = base + 2 * L.n; L.dg = base + 3 * L.n;
            L.b = base + 4 * L.n; L.x = base + 5 * L.n; L.xt = base + 6 * L.n;
        } else {
            L.cx = L.cy = L.cz = L.dg = L.b = L.x = nullptr;
            L.xt = base;
        }
    }
    H.nb = (int)H.lv.back().n;
    PPS_TRY(H.minv.ensure(sizeof(T) * H.nb * H.nb, arena));
    PPS_TRY(H.work.ensure(sizeof(double) * H.nb * H.nb, arena));
    return PPS_OK;
}

// coarse levels + bottom inverse from an already filled fine level
template <class T>
static void hier_build(Hier<T> &H, int singular, hipStream_t s) {
    const int nl = (int)H.lv.size();
    for (int l = 1; l < nl; ++l) {
        const Level<T> &F = H.lv[l - 1];
        Level<T> &C = H.lv[l];
        if (!F.reg) {
            if (singular) k_coarsen_g<T, T, true><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(F.k(), F.g, C.cx, C.cy, C.cz, C.dg);
            else k_coarsen_g<T, T, false><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(F.k(), F.g, C.cx, C.cy, C.cz, C.dg);
        } else if (singular) {
            k_coarsen<T, T, true><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(F.k(), C.cx, C.cy, C.cz, C.dg);
        } else {
            k_coarsen<T, T, false><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(F.k(), C.cx, C.cy, C.cz, C.dg);
        }
    }
    const size_t lds = sizeof(double) * ((size_t)H.nb * H.nb + 2 * (size_t)H.nb);
    if (lds <= 65536)
        k_bottom_build_lds<T><<<1, 1024, lds, s>>>(H.lv.back().k(), singular, H.minv.template as<T>());
    else
        k_bottom_build<T><<<1, 1024, 0, s>>>(H.lv.back().k(), singular, H.work.template as<double>(),
                                             H.minv.template as<T>());
}

// One colour pass of level l.  Algorithmic bytes of a pass on n cells:
// all three coupling planes (a cell's -x/-y/-z couplings live in its
// other-colour neighbours' planes), the diagonal and right-hand side of the
// updated colour, the other colour's iterate (read) and the updated
// colour's iterate (write; plus read unless FIRST):
//   FIRST: (3 + 1/2 + 1/2 + 1/2) n S ;  else (3 + 1/2 + 1/2 + 1/2 + 1/2 + 1/2) n S
template <class T, bool DGI>
static void rb_pass_t(const Level<T> &L, int l, const T *b, int colour, bool first, Prof *prof,
                      hipStream_t s) {
    const LevelK<T> K = L.k();
    const dim3 g2((unsigned)(((L.nx + 1) / 2 + 63) / 64), (unsigned)((L.ny + 3) / 4), (unsigned)L.nz);
    if (l != 0) {
        if (first) k_rb<T, T, true, false, DGI><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
        else k_rb<T, T, false, false, DGI><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
        return;
    }
    const double bytes = (double)L.n * sizeof(T) * ((first ? 4.5 : 5.5) - (DGI ? 0.5 : 0.0));
    hipEvent_t *e = prof ? prof->begin(bytes) : nullptr;
    if (e) (void)hipEventRecord(e[0], s);
    if (first) k_rb<T, T, true, true, DGI><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
    else k_rb<T, T, false, true, DGI><<<g2, kB3, 0, s>>>(K, b, L.x, colour);
    if (e) (void)hipEventRecord(e[1], s);
}
template <class T>
static void rb_pass(const Level<T> &L, int l, const T *b, int colour, bool first, bool dgi,
                    Prof *prof, hipStream_t s) {
    if (dgi) rb_pass_t<T, true>(L, l, b, colour, first, prof, s);
    else rb_pass_t<T, false>(L, l, b, colour, first, prof, s);
}

// Launch shape of the fused sweep on a level: pairs per thread, thread
// rows per block, planes per block.  Blocks hold at most 1024 threads and
// 64 KB of LDS; one runs per CU.
// Pairs per thread: 2 in FP32 when the diagonals are the coupling sums
// (dgi) and a row has a multiple of 128 pairs (whole wavefronts per row;
// twice the rows per block, so half the ring-row overhead), without the
// run-ahead loads except in the first sweep, which would not fit the
// registers; else 1.  (Two pairs per thread on rows of 64 pairs, 128^3,
// ran at 21 against 18 us.)
// Planes per block: with one pair per thread a level of 2^23 cells or
// more takes about 1024 blocks (256^3 FP64: 11 planes, within 1% of the
// best of 2..32); otherwise the count minimises waves of blocks x steps
// per block, the sweep being bound by the latency of its plane steps
// (128^3: 6 planes, 220 blocks, 18 us against 30 us for 640 blocks of 2
// planes; 256^3 FP32 with two pairs per thread: 20 planes, 247 blocks,
// 109 us against 124 us for one pair per thread).
// Thread rows: at most 8 on levels of at most 32 pairs a row (64^3, 32^3 at
// 256^3), whose small blocks then run several to a CU and overlap their
// plane steps: the 256^3 mixed solve 7.46 -> 7.35 ms (32 and 30 rows before).
struct SweepShape {
    int px, tpr, rb, oy, zc, ytiles;
    bool lop;   // couplings staged through LDS (k_rb_sweep<..., LOP>)
    dim3 grid, block;
    size_t lds;
};
template <class T>
static SweepShape sweep_shape1(int nx, int ny, int nz, int max_rows, int planes, int cus, bool dgi,
                               bool lop) {
    SweepShape sh;
    const int np = nx / 2;
    sh.px = (sizeof(T) == 4 && dgi && np % 128 == 0) ? 2 : 1;
    // LOP: two pairs a thread (the one-pair form on 128-wide levels, with
    // the couplings staged the same way, ran slower than the register form:
    // the 256^3 mixed solve 7.46 against 7.49 ms)
    sh.lop = lop && dgi && sizeof(T) == 4 && sh.px == 2;
    sh.tpr = np / sh.px;
    int rb = (sh.lop ? kLopThreads : 1024) / sh.tpr;
    if (max_rows > 0 && rb > max_rows) rb = max_rows;
    if (max_rows <= 0 && np <= 32 && rb > 8) rb = 8;
    const int lds_rows = (int)((65536 / ((long long)np * sizeof(T)) - 8) / 6);
    if (rb > lds_rows) rb = lds_rows;
    if (rb > ny + 2) rb = ny + 2;
    sh.rb = rb;                      // < 3: the caller falls back to colour passes
    sh.oy = rb - 2 >= ny ? ny : (rb > 2 ? rb - 2 : 1);
    const int ytiles = (ny + sh.oy - 1) / sh.oy;
    int zc = planes;
    if (zc <= 0 && sh.px == 1 && (long long)nx * ny * nz >= (1LL << 23)) {
        zc = (int)(((long long)nz * ytiles + 1023) / 1024);
    } else if (zc <= 0) {
        long long best = -1;
        for (int c = 1; c <= nz; ++c) {
            const long long blocks = (long long)ytiles * ((nz + c - 1) / c);
            const long long cost = ((blocks + cus - 1) / cus) * (c + 2);
            if (best < 0 || cost < best) { best = cost; zc = c; }
        }
    }
    if (zc < 1) zc = 1;
    if (zc > nz) zc = nz;
    sh.zc = zc;
    sh.ytiles = ytiles;
    sh.grid = dim3((unsigned)(ytiles * ((nz + zc - 1) / zc)), 1, 1);
    sh.block = dim3((unsigned)sh.tpr, (unsigned)(rb > 0 ? rb : 1), 1);
    sh.lds = sizeof(T) * (size_t)np * (4 * (rb + 2) + 2 * rb);
    return sh;
}
// The LOP form where its two coupling buffers (3 rb + 1 rows each) and the
// X / R planes fit the device's LDS per workgroup (160 KiB on MI355X), else
// the plain one.
template <class T>
static SweepShape sweep_shape(int nx, int ny, int nz, int max_rows, int planes, int cus, bool dgi,
                              size_t lds_max) {
    if (sizeof(T) == 4) {
        const SweepShape sh = sweep_shape1<T>(nx, ny, nz, max_rows, planes, cus, dgi, true);
        if (sh.lop && sh.rb >= 3 && (long long)(3 * sh.rb + 1) * nx <= kLopCap &&
            sh.lds + 2 * sizeof(float) * kLopCap <= lds_max)
            return sh;
    }
    return sweep_shape1<T>(nx, ny, nz, max_rows, planes, cus, dgi, false);
}

// LDS bytes of a tail that starts at level lt (and its offsets in K)
template <class T>
static size_t tail_layout(const Hier<T> &H, int lt, bool dgi, bool tri, TailK<T> *K) {
    const int nl = (int)H.lv.size();
    size_t off = 0;
    if (K) K->off_r = 0;
    if (tri && lt < nl) off += (size_t)H.lv[lt].n;   // residual of the tail's top level
    for (int l = lt; l < nl; ++l) {
        const size_t n = (size_t)H.lv[l].n;
        int o[6];
        o[0] = (int)off; off += n;
        o[1] = (int)off; off += n;
        o[2] = (int)off; off += n;
        o[3] = (int)off; if (!dgi) off += n;
        o[4] = (int)off; off += n;
        o[5] = (int)off; off += n;
        if (K) for (int f = 0; f < 6; ++f) K->off[l - lt][f] = o[f];
    }
    if (K) K->off_minv = (int)off;
    off += (size_t)H.nb * H.nb;
    return off * sizeof(T);
}

// How a V-cycle runs (from the handle's options and the system):
//   nu_pre, nu_post  smoothing sweeps;
//   dgi              every diagonal is the coupling sum (singular systems);
//   fused            one kernel per red-black sweep (k_rb_sweep) instead of
//                    one per colour (k_rb);
//   rows, planes     fused-sweep block shape overrides (0 = automatic);
//   tail_cells       levels below the finest with at most this many cells
//                    form the one-workgroup tail (0: no tail).
struct Cycle {
    int nu_pre = 2, nu_post = 2;
    bool dgi = false, fused = true;
    bool tri = false;          // trilinear transfers
    int per[3] = {0, 0, 0};    // periodic axes (trilinear transfers)
    int rows = 0, planes = 0;
    long long tail_cells = 4096;
    size_t lds_max = 65536;   // LDS bytes one workgroup may use
    int cus = 256;            // compute units (fused-sweep plane chunks)
};

// One fused red-black sweep of level l (out of place, the iterate pointers
// swap).  Algorithmic bytes: the operator planes (3 couplings, plus the
// diagonal unless dgi), b, the old iterate (unless FIRST) and the new
// iterate, once each.
// Launch of a fused sweep; with ev, the kernel's own start and end are
// recorded in ev[0], ev[1] (hipExtLaunchKernelGGL: the timestamps of the
// dispatch itself, no marker packets between kernels).
template <class... KArgs, class... Args>
static void launch_sweep(void (*k)(KArgs...), const SweepShape &sh, hipStream_t s, hipEvent_t *ev,
                         Args... args) {
    if (ev)
        hipExtLaunchKernelGGL(k, sh.grid, sh.block, (std::uint32_t)sh.lds, s, ev[0], ev[1], 0, args...);
    else
        hipLaunchKernelGGL(k, sh.grid, sh.block, sh.lds, s, args...);
}

template <class T, bool FIRST, bool DGI, bool PROL, int PX>
static void sweep_lop(const SweepShape &sh, const LevelK<T> &K, const Level<T> &L, int l, const T *b,
                      int c1, const T *ec, double *dot, hipStream_t s, hipEvent_t *ev, const double *skip) {
    if constexpr (!FIRST) {
        if (l == 0 && dot) {
            launch_sweep(k_rb_sweep<T, FIRST, PX, true, DGI, PROL, true, true, true>, sh, s, ev,
                K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, dot, skip);
            return;
        }
    }
    if (l == 0)
        launch_sweep(k_rb_sweep<T, FIRST, PX, true, DGI, PROL, true, false, true>, sh, s, ev,
            K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, nullptr, skip);
    else
        launch_sweep(k_rb_sweep<T, FIRST, PX, false, DGI, PROL, true, false, true>, sh, s, ev,
            K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, nullptr, skip);
}
template <class T, bool FIRST, bool DGI, bool PROL, bool PF>
static void sweep_launch2(const SweepShape &sh, const Level<T> &L, int l, const T *b, int c1,
                          const T *ec, double *dot, hipStream_t s, hipEvent_t *ev, const double *skip) {
    const LevelK<T> K = L.k();
    if constexpr (sizeof(T) == 4 && DGI) {
        // couplings staged through LDS
        if (sh.lop) {
            sweep_lop<T, FIRST, DGI, PROL, 2>(sh, K, L, l, b, c1, ec, dot, s, ev, skip);
            return;
        }
        // two pairs per thread, the LDS buffers not fitting (no run-ahead
        // loads but in the first sweep)
        if (sh.px == 2) {
            if constexpr (!FIRST) {
                if (l == 0 && dot) {
                    launch_sweep(k_rb_sweep<T, FIRST, 2, true, DGI, PROL, false, true>, sh, s, ev,
                        K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, dot, skip);
                    return;
                }
            }
            // the first sweep reads no iterate, which leaves registers for
            // the run-ahead loads (115 VGPRs; 89 against 95 us at 256^3)
            if (l == 0)
                launch_sweep(k_rb_sweep<T, FIRST, 2, true, DGI, PROL, FIRST, false>, sh, s, ev,
                    K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, nullptr, skip);
            else
                launch_sweep(k_rb_sweep<T, FIRST, 2, false, DGI, PROL, false, false>, sh, s, ev,
                    K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, nullptr, skip);
            return;
        }
    }
    if constexpr (!FIRST) {
        if (l == 0 && dot) {
            launch_sweep(k_rb_sweep<T, FIRST, 1, true, DGI, PROL, PF, true>, sh, s, ev,
                K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, dot, skip);
            return;
        }
    }
    if (l == 0)
        launch_sweep(k_rb_sweep<T, FIRST, 1, true, DGI, PROL, PF, false>, sh, s, ev,
            K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, nullptr, skip);
    else
        launch_sweep(k_rb_sweep<T, FIRST, 1, false, DGI, PROL, PF, false>, sh, s, ev,
            K, b, L.x, L.xt, c1, sh.zc, sh.oy, sh.ytiles, ec, nullptr, skip);
}
// Loads run ahead in FP32 only: the FP64 kernel with both operator sets
// exceeds 128 VGPRs and spills (226 us against 198 us at 256^3).
template <class T, bool FIRST, bool DGI, bool PROL>
static void sweep_launch1(const SweepShape &sh, const Level<T> &L, int l, const T *b, int c1,
                          const T *ec, double *dot, hipStream_t s, hipEvent_t *ev, const double *skip) {
    sweep_launch2<T, FIRST, DGI, PROL, sizeof(T) == 4>(sh, L, l, b, c1, ec, dot, s, ev, skip);
}
template <class T, bool DGI>
static void sweep_launch(const SweepShape &sh, const Level<T> &L, int l, const T *b, int c1, bool first,
                         const T *ec, double *dot, hipStream_t s, hipEvent_t *ev, const double *skip) {
    if (first) sweep_launch1<T, true, DGI, false>(sh, L, l, b, c1, nullptr, nullptr, s, ev, skip);
    else if (ec) sweep_launch1<T, false, DGI, true>(sh, L, l, b, c1, ec, dot, s, ev, skip);
    else sweep_launch1<T, false, DGI, false>(sh, L, l, b, c1, nullptr, dot, s, ev, skip);
}

// Whether level l's sweeps run as one fused kernel (else two colour passes).
template <class T>
static bool sweep_fused(const Level<T> &L, const Cycle &cy) {
    const SweepShape sh = sweep_shape<T>(L.nx, L.ny, L.nz, cy.rows, cy.planes, cy.cus, cy.dgi, cy.lds_max);
    return cy.fused && !(sh.rb < 3 || sh.lds > 65536 || (L.nx & 1) || sh.tpr * sh.rb > 1024);
}

// Whether the fused sweep of the finest level can also form the block sums
// of b x, x and b (whole wavefronts per block, one partial slot per block).
template <class T>
static bool sweep_dot_ok(const Level<T> &L, const Cycle &cy) {
    if (!sweep_fused(L, cy)) return false;
    const SweepShape sh = sweep_shape<T>(L.nx, L.ny, L.nz, cy.rows, cy.planes, cy.cus, cy.dgi, cy.lds_max);
    return (sh.tpr * sh.rb) % 64 == 0 && (long long)sh.grid.x <= kRedBlocks;
}

// Blocks of the fused sweep of level l (its partial sums under DOT).
template <class T>
static int sweep_blocks(const Level<T> &L, const Cycle &cy) {
    return (int)sweep_shape<T>(L.nx, L.ny, L.nz, cy.rows, cy.planes, cy.cus, cy.dgi, cy.lds_max).grid.x;
}

// ec: the coarse correction to add to the iterate first (fused sweeps
// only); dot: partial sums of b x, x, b of the result (sweep_dot_ok only).
template <class T>
static void rb_sweep(Level<T> &L, int l, const T *b, int c1, bool first, const Cycle &cy, Prof *prof,
                     hipStream_t s, const T *ec = nullptr, double *dot = nullptr,
                     const double *skip = nullptr) {
    const bool dgi = cy.dgi;
    const SweepShape sh = sweep_shape<T>(L.nx, L.ny, L.nz, cy.rows, cy.planes, cy.cus, cy.dgi, cy.lds_max);
    if (!sweep_fused(L, cy)) {
        // rows too long for the LDS planes: two colour passes instead
        rb_pass(L, l, b, c1, first, dgi, prof, s);
        rb_pass(L, l, b, 1 - c1, false, dgi, prof, s);
        return;
    }
    // the profiled kernel: finest level, neither the first sweep of a cycle
    // nor the one that adds the coarse correction
    const double bytes = (double)L.n * sizeof(T) * (dgi ? 6.0 : 7.0);
    hipEvent_t *e = (l == 0 && !first && !ec && !dot && prof) ? prof->begin(bytes) : nullptr;
    if (dgi) sweep_launch<T, true>(sh, L, l, b, c1, first, ec, dot, s, e, skip);
    else sweep_launch<T, false>(sh, L, l, b, c1, first, ec, dot, s, e, skip);
    T *t = L.x;
    L.x = L.xt;
    L.xt = t;
}

// The finest level's first pre-sweep of a V-cycle on its own, queued ahead
// by the outer iteration; it returns at once when *skip is non-zero.
template <class T>
static void queue_first_sweep(Hier<T> &H, const T *fineb, const Cycle &cy, hipStream_t s, const double *skip) {
    rb_sweep<T>(H.lv[0], 0, fineb, 0, true, cy, nullptr, s, nullptr, nullptr, skip);
}
// Whether the look-ahead applies: the finest level runs fused sweeps (one
// kernel per sweep) and the cycle has a level below it.
template <class T>
static bool can_queue_first(const Hier<T> &H, const Cycle &cy) {
    return cy.fused && H.lv.size() > 1 && sweep_fused(H.lv[0], cy);
}

// one V-cycle: x_0 = V(b_0).  fineb: the fine right-hand side (the FP64
// residual in FP64 mode, its FP32 copy in mixed mode).  dot: when not null
// and the last fine sweep is a fused one that can, that sweep writes the
// block sums of b_0 x_0, x_0 and b_0 there (as k_dot_rz lays them out);
// returns their block count, 0 if it did not.  first_done: the finest
// level's first pre-sweep is already queued (queue_first_sweep).
template <class T>
static int vcycle(Hier<T> &H, const T *fineb, const Cycle &cy, Prof *prof, hipStream_t s,
                  double *dot = nullptr, hipEvent_t coarse_ready = nullptr, bool first_done = false) {
    const int nl = (int)H.lv.size();
    const int nu_pre = cy.nu_pre, nu_post = cy.nu_post;
    const bool dgi = cy.dgi, fused = cy.fused;
    int ndot = 0;
    // levels lt .. nl-1 run in one workgroup (k_tail): the first level
    // below the finest with at most cy.tail_cells cells
    int lt = nl;
    if (cy.tail_cells > 0) {
        for (int l = 1; l < nl; ++l) {
            if (H.lv[l].n <= cy.tail_cells) { lt = l; break; }
        }
        // the tail must fit the workgroup's LDS, and its transfers are the
        // regular ones
        auto tail_regular = [&](int from) {
            for (int l = from; l + 1 < nl; ++l)
                if (!H.lv[l].reg) return false;
            return true;
        };
        while (lt < nl && (nl - lt > kTailMax || !tail_regular(lt) ||
                           tail_layout(H, lt, dgi, cy.tri, (TailK<T> *)nullptr) > cy.lds_max))
            ++lt;
    }
    for (int l = 0; l < (lt < nl ? lt : nl - 1); ++l) {
        Level<T> &L = H.lv[l];
        const T *b = l == 0 ? fineb : L.b;
        if (fused) {
            if (!(l == 0 && first_done)) rb_sweep(L, l, b, 0, true, cy, prof, s);
            for (int q = 1; q < nu_pre; ++q) rb_sweep(L, l, b, 0, false, cy, prof, s);
        } else {
            rb_pass(L, l, b, 0, true, dgi, prof, s);
            rb_pass(L, l, b, 1, false, dgi, prof, s);
            for (int q = 1; q < nu_pre; ++q) {
                rb_pass(L, l, b, 0, false, dgi, prof, s);
                rb_pass(L, l, b, 1, false, dgi, prof, s);
            }
        }
        Level<T> &C = H.lv[l + 1];
        if (cy.tri) {
            // residual into the free iterate buffer, then full weighting
            const TriK t{L.nx, L.ny, L.nz, cy.per[0], cy.per[1], cy.per[2]};
            if (dgi) k_resid_field<T, T, true><<<cgrid3(L.nx, L.ny, L.nz), kB3, 0, s>>>(L.k(), b, L.x, L.xt);
            else k_resid_field<T, T, false><<<cgrid3(L.nx, L.ny, L.nz), kB3, 0, s>>>(L.k(), b, L.x, L.xt);
            k_restrict_tri<T><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(t, L.xt, C.b);
        } else if (!L.reg) {
            if (dgi) k_resid_restrict_g<T, T, true><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(L.k(), L.g, b, L.x, C.b);
            else k_resid_restrict_g<T, T, false><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(L.k(), L.g, b, L.x, C.b);
        } else if (L.nx >= 32 && L.ny >= 16) {
            const TGrid rg = tgrid(L.nx, L.ny, L.nz, 1024, true);
            if (dgi) k_resid_restrict_tiled<T, true><<<rg.blocks, dim3(64, 16, 1), 0, s>>>(L.k(), rg, b, L.x, C.b);
            else k_resid_restrict_tiled<T, false><<<rg.blocks, dim3(64, 16, 1), 0, s>>>(L.k(), rg, b, L.x, C.b);
        } else {
            if (dgi) k_resid_restrict<T, T, true><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(L.k(), b, L.x, C.b);
            else k_resid_restrict<T, T, false><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(L.k(), b, L.x, C.b);
        }
        // the coarse operators (built on the side stream) are needed from
        // here on: the finest level's smoothing and restriction only wrote
        // the coarse right-hand side
        if (l == 0 && coarse_ready) {
            (void)hipStreamWaitEvent(s, coarse_ready, 0);
            coarse_ready = nullptr;
        }
    }
    if (coarse_ready) (void)hipStreamWaitEvent(s, coarse_ready, 0);
    if (lt < nl) {
        TailK<T> K;
        K.nl = nl - lt;
        K.nb = H.nb;
        K.nu_pre = nu_pre;
        K.nu_post = nu_post;
        for (int l = lt; l < nl; ++l) K.L[l - lt] = H.lv[l].k();
        K.b0 = H.lv[lt].b;
        K.x0 = H.lv[lt].x;
        K.minv = H.minv.template as<T>();
        K.px = cy.per[0]; K.py = cy.per[1]; K.pz = cy.per[2];
        const size_t lds = tail_layout(H, lt, dgi, cy.tri, &K);
        if (cy.tri) {
            if (dgi) k_tail<T, true, true><<<1, 1024, lds, s>>>(K);
            else k_tail<T, false, true><<<1, 1024, lds, s>>>(K);
        } else {
            if (dgi) k_tail<T, true, false><<<1, 1024, lds, s>>>(K);
            else k_tail<T, false, false><<<1, 1024, lds, s>>>(K);
        }
    } else {
        Level<T> &Bt = H.lv.back();
        const T *b = nl == 1 ? fineb : Bt.b;
        k_bottom_solve<T><<<1, 512, 0, s>>>(H.nb, H.minv.template as<T>(), b, Bt.x);
    }
    for (int l = (lt < nl ? lt : nl - 1) - 1; l >= 0; --l) {
        Level<T> &L = H.lv[l];
        const Level<T> &C = H.lv[l + 1];
        const T *b = l == 0 ? fineb : L.b;
        // the first fused post-sweep adds a piecewise-constant correction
        // as it loads x; the trilinear one is a pass of its own
        const bool fold = nu_post > 0 && sweep_fused(L, cy) && !cy.tri && L.reg;
        if (cy.tri) {
            const TriK t{L.nx, L.ny, L.nz, cy.per[0], cy.per[1], cy.per[2]};
            k_prolong_tri_add<T><<<cgrid3(L.nx, L.ny, L.nz), kB3, 0, s>>>(t, L.x, C.x);
        } else if (!L.reg) {
            k_prolong_add_g<T><<<cgrid3(L.nx, L.ny, L.nz), kB3, 0, s>>>(L.nx, L.ny, L.nz, L.g, L.x, C.x);
        } else if (!fold) {
            k_prolong_add<T><<<cgrid3(C.nx, C.ny, C.nz), kB3, 0, s>>>(L.nx, L.ny, L.nz, L.x, C.x);
        }
        const bool with_dot = l == 0 && dot && fused && sweep_dot_ok(L, cy);
        for (int q = 0; q < nu_post; ++q) {
            if (fused) {
                rb_sweep(L, l, b, 1, false, cy, prof, s, (fold && q == 0) ? (const T *)C.x : nullptr,
                         (with_dot && q == nu_post - 1) ? dot : nullptr);
            } else {
                rb_pass(L, l, b, 1, false, dgi, prof, s);
                rb_pass(L, l, b, 0, false, dgi, prof, s);
            }
        }
        if (with_dot && nu_post > 0) ndot = sweep_blocks(L, cy);
    }
    return ndot;
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
    Arena *a = f->arena;
    PPS_TRY(f->op.ensure(sizeof(double) * 4 * n, a));
    PPS_TRY(f->vec.ensure(sizeof(double) * 5 * n, a));
    PPS_TRY(f->red.ensure(sizeof(double) * (4 * kRedBlocks + 64), a));
    PPS_TRY(f->flags.ensure(64, a));
    if (!h->pinned) PPS_HIP(hipHostMalloc((void **)&h->pinned, 16 * sizeof(double), 0));
    if (!f->built) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) == hipSuccess && v > 0)
            f->lds_max = (size_t)v;
        v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && v > 0)
            f->cus = v;
        // the tail kernel may use all of it
        const int mx = (int)f->lds_max;
        const void *tails[8] = {(const void *)k_tail<float, true, false>, (const void *)k_tail<float, false, false>,
                                (const void *)k_tail<double, true, false>, (const void *)k_tail<double, false, false>,
                                (const void *)k_tail<float, true, true>, (const void *)k_tail<float, false, true>,
                                (const void *)k_tail<double, true, true>, (const void *)k_tail<double, false, true>};
        for (const void *kt : tails) (void)hipFuncSetAttribute(kt, hipFuncAttributeMaxDynamicSharedMemorySize, mx);
        (void)hipGetLastError();
        f->built = 1;
    }
    return PPS_OK;
}

struct Vecs {
    double *b, *x, *r, *q, *z;
};
static Vecs vecs(FastState *f) {
    const long long n = f->d.cells();
    double *v = f->vec.as<double>();
    Vecs V;
    V.b = v; V.x = v + n; V.r = v + 2 * n; V.q = v + 3 * n; V.z = v + 4 * n;
    return V;
}
static OpK opk(FastState *f) {
    const long long n = f->d.cells();
    const int px = f->d.per[0], py = f->d.per[1], pz = f->d.per[2];
    if (f->op_in_A) {
        const double *A = f->built_for;   // slots 2, 4, 6, 7
        return OpK{f->d.nx, f->d.ny, f->d.nz, A + n, A + 3 * n, A + 5 * n, A + 6 * n, px, py, pz};
    }
    double *o = f->op.as<double>();
    return OpK{f->d.nx, f->d.ny, f->d.nz, o, o + n, o + 2 * n, o + 3 * n, px, py, pz};
}

// Waits until the mailbox carries the last requested seq (a k_scal_rr on
// stream s wrote it).  Spins on the coherent host word; now and then asks
// whether the stream has drained, so that a failed launch ends the wait.
static int wait_mailbox(pps_handle *h, hipStream_t s) {
    const unsigned long long want = h->mseq;
    for (long long spin = 1;; ++spin) {
        if (__atomic_load_n(&h->mbox->seq, __ATOMIC_ACQUIRE) == want) return PPS_OK;
        if ((spin & 255) == 0) {
            const hipError_t e = hipStreamQuery(s);
            if (e == hipSuccess) {
                if (__atomic_load_n(&h->mbox->seq, __ATOMIC_ACQUIRE) == want) return PPS_OK;
                std::fprintf(stderr, "libpps: the stream drained without the expected reduction\n");
                return PPS_ERR_HIP;
            }
            if (e != hipErrorNotReady) PPS_HIP(e);
        }
        __builtin_ia32_pause();
    }
}

// pack A (or, given rows, the system assembled from the fields), detect
// symmetry/singularity, build the hierarchy for `mixed`
//
// defer: do not wait for the setup pass.  Its flags are copied to the
// pinned scalars and read at the next synchronisation (outer_iterate), and
// the coarse hierarchy is built on the side stream for a singular system
// (every PARIS wall / periodic system is), overlapping the start of the
// first V-cycle; outer_iterate builds it again on the solve stream if the
// flags say the system is not singular.
static int fast_setup(pps_handle *h, const double *dA, const double *dp, bool mixed, hipStream_t s,
                      const RowK *rows = nullptr, bool defer = false, const FacesK *faces = nullptr) {
    FastState *f = h->fast;
    const Dims &d = f->d;
    const long long n = d.cells();
    Vecs V = vecs(f);
    int *flags = f->flags.as<int>();
    std::vector<int> dims;
    const double lengths[3] = {h->grid.xlength, h->grid.ylength, h->grid.zlength};
    PPS_TRY(plan_levels(d, lengths, dims));
    // the trilinear transfers need every level halved in every direction
    if (h->opt.transfer == PPS_TRANSFER_TRILINEAR) {
        for (size_t l = 0; l + 3 < dims.size(); l += 3)
            for (int a = 0; a < 3; ++a)
                if ((dims[l + a] & 1) || dims[l + 3 + a] * 2 != dims[l + a]) return PPS_ERR_UNSUPPORTED;
    }
    float *f32 = nullptr, *r32 = nullptr;
    if (mixed) {
        if (!f->mixed_built) { PPS_TRY(hier_alloc(f->h32, dims, true, f->arena)); f->mixed_built = 1; }
        PPS_TRY(f->b32.ensure(sizeof(float) * n, f->arena));
        f32 = f->h32.lv[0].cx;   // cx, cy, cz, dg are contiguous
        r32 = f->b32.as<float>();
    }
    // operator, right-hand side, iterate and initial residual in one pass
    double *part = f->red.as<double>();
    double *S = part + 4 * kRedBlocks;
    const ZGrid zg = zgrid(d.nx, d.ny, d.nz);
    PPS_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
    // mixed mode: the outer kernels read the operator from A itself, unless
    // the rows are weighted (cell widths set) or there is no A
    double *o64 = f->op.as<double>();
    const double *w = h->widths.as<double>();
    const double *wx = h->have_widths ? w : nullptr;
    const double *wy = wx ? w + d.nx : nullptr, *wz = wx ? w + d.nx + d.ny : nullptr;
    const bool own = !mixed || wx || rows || faces;
    double *ocx = own ? o64 : nullptr, *ocy = own ? o64 + n : nullptr;
    double *ocz = own ? o64 + 2 * n : nullptr, *odg = own ? o64 + 3 * n : nullptr;
    if (rows) {
        const RowsF src{*rows, d.nx, d.ny, d.nz, d.ng, (long long)d.X(), (long long)d.X() * d.Y()};
        k_setup<RowsF><<<zg.grid, dim3(64, 4, 1), 0, s>>>(packk(d), zg.zc, src, dp, ocx, ocy, ocz, odg, V.b,
                                                          V.x, V.r, f32, r32, part, flags, wx, wy, wz);
    } else if (faces) {
        const RowsC src{*faces, d.nx, d.ny, d.nz, d.per[0], d.per[1], d.per[2], {0.0, 0.0, 0.0}};
        k_setup<RowsC><<<zg.grid, dim3(64, 4, 1), 0, s>>>(packk(d), zg.zc, src, dp, ocx, ocy, ocz, odg, V.b,
                                                          V.x, V.r, f32, r32, part, flags, wx, wy, wz);
    } else {
        const RowsA src{dA, n, 0.0};
        k_setup<RowsA><<<zg.grid, dim3(64, 4, 1), 0, s>>>(packk(d), zg.zc, src, dp, ocx, ocy, ocz, odg, V.b,
                                                          V.x, V.r, f32, r32, part, flags, wx, wy, wz);
    }
    // deferred: r.r, b.b and the flags go to the mailbox as well
    if (defer) k_scal_rr<<<1, 1024, 0, s>>>(part, zg.blocks, 1, S, h->mbox, ++h->mseq, flags);
    else k_scal_rr<<<1, 1024, 0, s>>>(part, zg.blocks, 1, S);
    PPS_HIP(hipGetLastError());
    // the hierarchy of the chosen precision, built on stream st
    auto build = [&](int singular, hipStream_t st) -> int {
        if (mixed) {
            hier_build(f->h32, singular, st);
        } else {
            Hier<double> &H = f->h64;
            if (!f->fp64_built) { PPS_TRY(hier_alloc(H, dims, false, f->arena)); f->fp64_built = 1; }
            Level<double> &L0 = H.lv[0];
            L0.cx = o64; L0.cy = o64 + n; L0.cz = o64 + 2 * n; L0.dg = o64 + 3 * n;
            L0.b = nullptr;
            L0.x = V.z;
            hier_build(H, singular, st);
        }
        return PPS_OK;
    };
    if (defer) {
        PPS_HIP(hipEventRecord(h->sev[0], s));
        PPS_HIP(hipStreamWaitEvent(h->side, h->sev[0], 0));
        PPS_TRY(build(1, h->side));
        PPS_HIP(hipEventRecord(h->sev[1], h->side));
        f->singular = 1;   // provisional until the flags are read
        f->exact = 1;
        f->flags_pending = 1;
        f->coarse_pending = 1;
    } else {
        int fl = 0;
        PPS_HIP(hipMemcpyAsync(&fl, flags, sizeof(int), hipMemcpyDeviceToHost, s));
        PPS_HIP(hipStreamSynchronize(s));
        if (fl & 1) return PPS_ERR_NAN_OR_TINY_DIAG;
        if (fl & 2) return PPS_ERR_NONSYMMETRIC;
        f->singular = (fl & 4) ? 0 : 1;
        f->exact = (fl & 8) ? 0 : 1;
        PPS_TRY(build(f->singular, s));
    }
    f->have_r = 1;
    PPS_HIP(hipGetLastError());
    // an operator assembled from fields is not kept for a later adjoint; one
    // from face values is known by their address
    f->built_for = rows ? nullptr : (faces ? faces->cf : dA);
    f->built_w[0] = faces ? faces->wx : nullptr;
    f->built_w[1] = faces ? faces->wy : nullptr;
    f->built_w[2] = faces ? faces->wz : nullptr;
    f->built_mixed = mixed ? 1 : 0;
    f->op_in_A = own ? 0 : 1;
    return PPS_OK;
}

// The outer iteration on (b, x): PCG or refinement with the V-cycle of
// hierarchy H.  Convergence: sqrt(rr/N) < tol_abs (norm 2).  Returns the
// iteration count in *it (maxit+1 if not converged).  The last iteration's
// x += alpha d is applied here, or, given pend, left to the caller: *pend is
// then the direction d (null when nothing is pending), alpha S[S_ALPHA].
template <class T>
static int outer_iterate(pps_handle *h, Hier<T> &H, bool mixed, double tol_abs, int maxit,
                         int *it, double *last, bool record, hipStream_t s,
                         const void **pend = nullptr) {
    if (pend) *pend = nullptr;
    FastState *f = h->fast;
    const long long n = f->d.cells();
    const OpK O = opk(f);
    Vecs V = vecs(f);
    double *part = f->red.as<double>();
    double *S = part + 4 * kRedBlocks;
    const int nb = nblk(n);
    const ZGrid zg = zgrid(f->d.nx, f->d.ny, f->d.nz);
    const TGrid tg = tgrid(f->d.nx, f->d.ny, f->d.nz);
    const dim3 zb(64, 4, 1);
    float *b32 = mixed ? f->b32.as<float>() : nullptr;
    const double N = (double)n;
    const bool cg = h->opt.outer == PPS_OUTER_CG;

    // r = b - M x (unless the setup pass just computed it)
    if (!f->have_r) {
        if (f->exact) k_outer_resid<float, true><<<zg.grid, zb, 0, s>>>(O, zg.zc, V.b, V.x, V.r, b32, part);
        else k_outer_resid<float, false><<<zg.grid, zb, 0, s>>>(O, zg.zc, V.b, V.x, V.r, b32, part);
        k_scal_rr<<<1, 1024, 0, s>>>(part, zg.blocks, 1, S, h->mbox, ++h->mseq);
    }
    f->have_r = 0;
    k_outer_start<<<1, 64, 0, s>>>(S);
    PPS_HIP(hipGetLastError());
    // r.r of the setup pass or of the residual above
    PPS_TRY(wait_mailbox(h, s));
    const double rr0 = h->mbox->rr;
    if (f->flags_pending) {
        // a deferred setup: its flags have arrived with r.r
        f->flags_pending = 0;
        const int fl = (int)h->mbox->flags;
        if (fl & 3) {
            PPS_HIP(hipStreamSynchronize(h->side));
            f->coarse_pending = 0;
            return (fl & 1) ? PPS_ERR_NAN_OR_TINY_DIAG : PPS_ERR_NONSYMMETRIC;
        }
        f->singular = (fl & 4) ? 0 : 1;
        f->exact = (fl & 8) ? 0 : 1;
        if (!f->singular) {
            // built for a singular system on the side stream: build again
            // with the stored diagonals, after that build
            PPS_HIP(hipStreamWaitEvent(s, h->sev[1], 0));
            hier_build(H, 0, s);
            PPS_HIP(hipGetLastError());
            f->coarse_pending = 0;
        }
    }
    Cycle cy;
    cy.nu_pre = h->opt.nu_pre < 1 ? 1 : h->opt.nu_pre;
    cy.nu_post = h->opt.nu_post;
    cy.dgi = f->singular != 0;
    cy.fused = h->opt.smoother != PPS_SMOOTH_COLOUR_PASSES;
    cy.rows = h->opt.sweep_rows;
    cy.planes = h->opt.sweep_planes;
    cy.tail_cells = h->opt.tail_cells == 0 ? 4096 : (h->opt.tail_cells < 0 ? 0 : h->opt.tail_cells);
    cy.lds_max = f->lds_max;
    cy.cus = f->cus;
    cy.tri = h->opt.transfer == PPS_TRANSFER_TRILINEAR;
    for (int a = 0; a < 3; ++a) cy.per[a] = f->d.per[a];
    double t2 = std::sqrt(rr0 / N);
    *last = t2;
    int k = 1;
    // The direction d_k goes to ring slot k % R.  With R = kRingSlots the
    // x += alpha_k d_k updates are deferred: the directions stay in the ring
    // with their alphas (S[S_AH + slot]) and are applied together by
    // k_flush, when the ring is full and after the last iteration, in
    // iteration order and each as the same fma, so x is bit for bit that of
    // one update per iteration, without x being read and written in every
    // k_dq_tiled (256^3: 16 of its 60 bytes a cell).  The no-allocation
    // solve keeps two slots (taken last from its arena), and k_dq_tiled
    // applies the previous direction itself.
    const bool lazy = cg && !f->arena;
    const int R = lazy ? kRingSlots : 2;
    if (cg) PPS_TRY(f->ring.ensure(sizeof(T) * (size_t)R * (size_t)n, f->arena));
    auto dslot = [&](int kk) -> T * { return f->ring.as<T>() + (long long)(kk % R) * n; };
    int p0 = 1;   // the first direction not yet applied to x
    auto flush = [&](int upto) {   // apply d_p0 .. d_upto
        constexpr int W = 16 / sizeof(T);
        if (upto >= p0 && n % W == 0)
            k_flush<T, W><<<nblk(n / W), 256, 0, s>>>(n, V.x, f->ring.as<T>(), S, p0 % R, upto - p0 + 1, R);
        else if (upto >= p0)
            k_flush<T, 1><<<nb, 256, 0, s>>>(n, V.x, f->ring.as<T>(), S, p0 % R, upto - p0 + 1, R);
        p0 = upto + 1;
    };
    if (t2 == 0.0) {   // exact solution: nothing to do
        if (f->coarse_pending) PPS_HIP(hipStreamSynchronize(h->side));
        f->coarse_pending = 0;
        *it = 1;
        return PPS_OK;
    }
    // Look-ahead: the next cycle's first fine sweep (it reads only the new
    // right-hand side) is queued behind each iteration's convergence test,
    // before the host reads that test, so the device does not idle while
    // the host decides; once the test stops the iteration the sweep returns
    // at once (S_STOP), and the device and the host take the same decision.
    const T *fb = mixed ? (const T *)(const void *)b32 : (const T *)(const void *)V.r;
    const bool ahead = can_queue_first(H, cy);
    bool queued = false;
    for (k = 1; k <= maxit; ++k) {
        // z = V(r)
        const int ndot = vcycle<T>(H, fb, cy, &h->prof, s, cg ? part : nullptr,
                                   f->coarse_pending ? h->sev[1] : nullptr, queued);
        f->coarse_pending = 0;
        queued = false;
        const T *z = H.lv[0].x;   // the sweeps swap buffers: read after the cycle
        if (cg) {
            // (r, z), sum z, sum r: from the last sweep (with r as the cycle
            // saw it: its FP32 copy in mixed mode), else a pass of their own
            int nbz = ndot;
            if (!nbz) {
                k_dot_rz<T><<<nb, 256, 0, s>>>(n, V.r, z, part);
                nbz = nb;
            }
            k_scal_beta<<<1, 1024, 0, s>>>(part, nbz, n, f->singular, S);
            // d_k's slot holds d_(k-R): applied first when still pending
            if (lazy && k - p0 >= R) flush(k - 1);
            T *dold = dslot(k - 1), *dnew = dslot(k);
            double *xu = lazy ? nullptr : V.x;
            if (f->exact)
                k_dq_tiled<T, true><<<tg.blocks, dim3(64, 16, 1), 0, s>>>(O, tg, z, dold, dnew, V.q, xu, S,
                                                                          part);
            else
                k_dq_tiled<T, false><<<tg.blocks, dim3(64, 16, 1), 0, s>>>(O, tg, z, dold, dnew, V.q, xu, S,
                                                                           part);
            if (!lazy) p0 = k;   // d_(k-1) applied in k_dq_tiled
            k_scal_alpha<<<1, 1024, 0, s>>>(part, tg.blocks, S, k % R);
            k_update<float><<<nb, 256, 0, s>>>(n, V.r, V.q, b32, S, part);
            k_scal_rr<<<1, 1024, 0, s>>>(part, nb, 0, S, h->mbox, ++h->mseq, nullptr, N, tol_abs);
        } else {
            k_zsum<T><<<nb, 256, 0, s>>>(n, z, part);
            k_add_z<T><<<nb, 256, 0, s>>>(n, V.x, z, part, nb, f->singular);
            if (f->exact) k_outer_resid<float, true><<<zg.grid, zb, 0, s>>>(O, zg.zc, V.b, V.x, V.r, b32, part);
            else k_outer_resid<float, false><<<zg.grid, zb, 0, s>>>(O, zg.zc, V.b, V.x, V.r, b32, part);
            k_scal_rr<<<1, 1024, 0, s>>>(part, zg.blocks, 0, S, h->mbox, ++h->mseq, nullptr, N, tol_abs);
        }
        if (ahead && k < maxit) {
            queue_first_sweep(H, fb, cy, s, S + S_STOP);
            queued = true;
        }
        PPS_HIP(hipGetLastError());
        PPS_TRY(wait_mailbox(h, s));
        t2 = std::sqrt(h->mbox->rr / N);
        const bool stop = h->mbox->pad != 0.0;
        if (record) h->hist.push_back(t2);
        *last = t2;
        if (!(t2 == t2) || t2 > 1e300) { *it = k; return PPS_ERR_DIVERGED; }
        if (stop) {
            // the queued sweep returned without writing: undo its swap of
            // the finest level's iterate buffers
            if (queued) std::swap(H.lv[0].x, H.lv[0].xt);
            queued = false;
            break;
        }
    }
    // the pending directions; the last one (alpha S[S_ALPHA]) left to the
    // caller when asked
    const int kl = k > maxit ? maxit : k;   // the last iteration run
    if (cg && lazy) flush(pend ? kl - 1 : kl);
    if (cg && pend) *pend = dslot(kl);
    else if (cg && !lazy) k_axpy_d<T><<<nb, 256, 0, s>>>(n, V.x, dslot(kl), S);
    PPS_HIP(hipGetLastError());
    *it = k;
    return PPS_OK;
}

// the outer scalars S (after the four partial-sum slots)
static double *part_S(FastState *f) { return f->red.as<double>() + 4 * kRedBlocks; }

// The final check (k_check_tiled) of the fast solve; dd: the pending
// direction of the last x += alpha d (null for none).
template <class TD>
static void launch_check(FastState *f, const Dims &d, const TGrid &tg, const TD *dd, const double *S,
                         double *dp, double *part, bool proj, const ProjK &P, const double *wx,
                         const double *wy, const double *wz, hipStream_t s) {
    const dim3 blk(64, 16, 1);
    const Vecs V = vecs(f);
    if (proj) {
        auto k = f->exact ? k_check_tiled<true, true, TD> : k_check_tiled<false, true, TD>;
        k<<<tg.blocks, blk, 0, s>>>(opk(f), packk(d), tg, V.b, V.x, dd, S, dp, part, P, wx, wy, wz);
    } else {
        auto k = f->exact ? k_check_tiled<true, false, TD> : k_check_tiled<false, false, TD>;
        k<<<tg.blocks, blk, 0, s>>>(opk(f), packk(d), tg, V.b, V.x, dd, S, dp, part, P, wx, wy, wz);
    }
}

// relative: maxError bounds ||r||_2 / ||b||_2 (b the folded right-hand
// side) instead of sqrt(sum r^2 / N)
int fast_solve(pps_handle *h, const double *dA, double *dp, double maxError, int maxit, int norm,
               int *it, double *tres2, hipStream_t s, bool relative, const ProjK *proj,
               const RowK *rows, const FacesK *faces) {
    if ((rows || faces) && norm != 2) return PPS_ERR_BAD_ARG;
    if (proj && norm != 2) return PPS_ERR_BAD_ARG;
    PPS_TRY(fast_ensure(h));
    FastState *f = h->fast;
    const bool mixed = h->opt.method == PPS_MG_MIXED;
    h->hist.clear();
    PPS_HIP(hipEventRecord(h->ev[0], s));
    PPS_TRY(fast_setup(h, dA, dp, mixed, s, rows, true, faces));
    PPS_HIP(hipEventRecord(h->ev[1], s));
    const Dims &d = f->d;
    const double N = (double)d.cells();
    if (relative) {
        PPS_HIP(hipMemcpyAsync(h->pinned, f->red.as<double>() + 4 * kRedBlocks + S_BB, sizeof(double),
                               hipMemcpyDeviceToHost, s));
        PPS_HIP(hipStreamSynchronize(s));
        maxError *= std::sqrt(h->pinned[0] / N);
    }
    int total = 0, k = 0;
    double t2 = 0.0, last = 0.0;
    int rc = PPS_OK;
    // norm 2: the last iteration's x += alpha d is fused into the final check
    const void *pend = nullptr;
    // iterate on the recursive residual, then confirm with the true residual
    // b - M x (norm 2: the FP64 operator, fused with the write-back of p and,
    // given proj, with project_velocity into separate outputs, which a
    // restart simply rewrites; other norms: the reference's expression on
    // the caller's A); restart from the true residual if the recursion
    // drifted
    for (int attempt = 0; attempt < 4 && total < maxit; ++attempt) {
        const int budget = maxit - total;
        // weighted rows: the recursion sees W r, and |r| <= |W r| / min w
        const double tol_rec = h->have_widths ? maxError * h->wmin : maxError;
        const void **pp = norm == 2 ? &pend : nullptr;
        if (mixed) rc = outer_iterate<float>(h, f->h32, true, tol_rec, budget, &k, &last, true, s, pp);
        else rc = outer_iterate<double>(h, f->h64, false, tol_rec, budget, &k, &last, true, s, pp);
        if (rc != PPS_OK && rc != PPS_ERR_NOT_CONVERGED) break;
        const bool exhausted = k > budget;
        total += exhausted ? budget : k;
        if (norm == 2) {
            const TGrid tg = tgrid(d.nx, d.ny, d.nz);
            double *part = f->red.as<double>();
            double *S = part + 4 * kRedBlocks;
            const double *w = h->widths.as<double>();
            const double *wx = h->have_widths ? w : nullptr;
            const double *wy = wx ? w + d.nx : nullptr, *wz = wx ? w + d.nx + d.ny : nullptr;
            const ProjK P = proj ? *proj : ProjK{};
            if (mixed) launch_check<float>(f, d, tg, (const float *)pend, S, dp, part, proj != nullptr, P, wx,
                                           wy, wz, s);
            else launch_check<double>(f, d, tg, (const double *)pend, S, dp, part, proj != nullptr, P, wx, wy,
                                      wz, s);
            k_scal_rr<<<1, 1024, 0, s>>>(part, tg.blocks, 0, S);
            PPS_HIP(hipGetLastError());
            PPS_HIP(hipMemcpyAsync(h->pinned, S + S_RR, sizeof(double), hipMemcpyDeviceToHost, s));
            PPS_HIP(hipStreamSynchronize(s));
            t2 = std::sqrt(h->pinned[0] / N);
        } else {
            k_unpack<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), vecs(f).x, dp);
            PPS_HIP(hipGetLastError());
            double res = 0.0;
            PPS_TRY(compute_residual_norm(h, d, dA, dp, norm, &res, s));
            t2 = res / N;
        }
        if (t2 < maxError) break;
        if (exhausted) break;
        // a restart iterates on from x: apply the pending update first
        if (pend) {
            const long long n = d.cells();
            if (mixed) k_axpy_d<float><<<nblk(n), 256, 0, s>>>(n, vecs(f).x, (const float *)pend, part_S(f));
            else k_axpy_d<double><<<nblk(n), 256, 0, s>>>(n, vecs(f).x, (const double *)pend, part_S(f));
            PPS_HIP(hipGetLastError());
            pend = nullptr;
        }
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

// M^T lam = g.  With cell widths the stored operator is S = W M (rows
// weighted, symmetric), so M^T = S W^-1: the hierarchy solves S mu = g and
// lam = W mu; the residual of S mu = g is that of M^T lam = g.  The
// singular case projects g mean-free either way (the null space of S, and
// of M, is the constants).
int fast_adjoint(pps_handle *h, const double *dA, const double *dg, double *dlam, double rtol,
                 int maxit, int *it, hipStream_t s, const FacesK *faces) {
    PPS_TRY(fast_ensure(h));
    FastState *f = h->fast;
    const bool mixed = h->opt.method == PPS_MG_MIXED;
    const Dims &d = f->d;
    const long long n = d.cells();
    // rebuild only if the operator changed; the fold of wall ghosts into
    // the right-hand side is irrelevant here (the adjoint ghosts are zero)
    // (a face-value system is known by the addresses of its values and of
    // its weights)
    const double *op = faces ? faces->cf : dA;
    const bool same_w = faces ? (f->built_w[0] == faces->wx && f->built_w[1] == faces->wy &&
                                 f->built_w[2] == faces->wz)
                              : f->built_w[0] == nullptr;
    if (f->built_for != op || !same_w || f->built_mixed != (mixed ? 1 : 0)) {
        PPS_TRY(fast_setup(h, dA, dlam, mixed, s, nullptr, false, faces));
    }
    Vecs V = vecs(f);
    k_interior<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), dg, V.b);
    // project the right-hand side mean-free when the operator is singular
    if (f->singular) {
        double *part = f->red.as<double>();
        const int nb = nblk(n);
        k_zsum<double><<<nb, 256, 0, s>>>(n, V.b, part);
        k_sub_mean<<<nb, 256, 0, s>>>(n, V.b, part, nb);
    }
    k_fill<<<nblk(n), 256, 0, s>>>(V.x, n, 0.0);
    // from x = 0 the residual is g itself: r, its FP32 copy and r.r = g.g
    // (the absolute tolerance) without a stencil pass; the outer iteration
    // starts from them
    double *part = f->red.as<double>();
    double *S = part + 4 * kRedBlocks;
    const int nbr = nblk(n);
    float *b32 = mixed ? f->b32.as<float>() : nullptr;
    k_resid_zero<float><<<nbr, 256, 0, s>>>(n, V.b, V.r, b32, part);
    k_scal_rr<<<1, 1024, 0, s>>>(part, nbr, 1, S, h->mbox, ++h->mseq);
    PPS_HIP(hipGetLastError());
    PPS_TRY(wait_mailbox(h, s));
    const double gg = h->mbox->bb;
    const double tol = rtol * std::sqrt(gg / (double)n);
    f->have_r = 1;
    h->hist.clear();
    if (gg == 0.0) {
        // g is zero (after the mean-free projection): lam = 0 exactly
        PPS_HIP(hipMemsetAsync(dlam, 0, sizeof(double) * d.padded(), s));
        PPS_HIP(hipStreamSynchronize(s));
        *it = 0;
        return PPS_OK;
    }
    int k = 0;
    double last = 0.0;
    int rc;
    if (mixed) rc = outer_iterate<float>(h, f->h32, true, tol, maxit, &k, &last, true, s);
    else rc = outer_iterate<double>(h, f->h64, false, tol, maxit, &k, &last, true, s);
    if (rc != PPS_OK) return rc;
    if (h->have_widths) {
        const double *w = h->widths.as<double>();
        k_weight<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), V.x, w, w + d.nx, w + d.nx + d.ny);
    }
    // padded output, ghosts: wrapped (periodic) or zero (walls)
    PPS_HIP(hipMemsetAsync(dlam, 0, sizeof(double) * d.padded(), s));
    k_unpack<<<cgrid3(d.nx, d.ny, d.nz), kB3, 0, s>>>(packk(d), V.x, dlam);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    *it = k;
    PPS_TRY(h->prof.collect());
    return last < tol ? PPS_OK : PPS_ERR_NOT_CONVERGED;
}

// ---------------------------------------------------------------------
// No-allocation solve.  Every device buffer of the fast solve (outer
// vectors, FP64 operator, reduction partials, the V-cycle hierarchy) is
// carved from one fixed array, `scratch`, instead of hipMalloc.  Each GPU
// holds its own copy of the array (kScratchBytes) from the library's load
// until the process ends; its address on each device is looked up once,
// by pps_init with that device current, and kept in scratch_at.
// ---------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static void *scratch_at[kMaxDevices];
static std::mutex scratch_mu;

int fast_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return PPS_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_at[device]) return PPS_OK;
    void *ptr = nullptr;
    PPS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return PPS_ERR_HIP;
    scratch_at[device] = ptr;
    return PPS_OK;
}

// Bytes the no-allocation solve carves for a grid, in the order and with
// the alignment fast_ensure / fast_setup / hier_alloc take them.
static size_t noalloc_bytes(const Dims &d, const double lengths[3], bool mixed) {
    std::vector<int> dims;
    if (plan_levels(d, lengths, dims) != PPS_OK) return 0;
    const size_t n = (size_t)d.cells();
    size_t tot = 0;
    auto take = [&](size_t b) { tot = ((tot + 255) & ~(size_t)255) + b; };
    take(sizeof(double) * 4 * n);
    take(sizeof(double) * 5 * n);
    take(sizeof(double) * (4 * (size_t)kRedBlocks + 64));
    take(64);
    const size_t S = mixed ? sizeof(float) : sizeof(double);
    const int nl = (int)dims.size() / 3;
    size_t nb = 0;
    for (int l = 0; l < nl; ++l) {
        const size_t nlc = (size_t)dims[3 * l] * dims[3 * l + 1] * dims[3 * l + 2];
        take(S * ((l > 0 || mixed) ? 7 : 1) * nlc);
        nb = nlc;
    }
    take(S * nb * nb);
    take(sizeof(double) * nb * nb);
    if (mixed) take(sizeof(float) * n);
    take(S * 2 * n);   // the two direction slots (outer_iterate)
    return tot;
}

}  // namespace pps

using namespace pps;

// pps_solve_device for the fast methods (norm 2) without any device
// allocation: the solver's buffers live in the scratch array.  Calls on one
// GPU share that array; calls that overlap in time on one GPU race.
extern "C" int pps_solve_device_noalloc(pps_handle *h, const double *dA, double *dp,
                                        double maxError, double beta, int maxit, int norm,
                                        int *it, double *tres2, void *stream) {
    (void)beta;   // Gauss-Seidel smoothing: the fast methods take no SOR weight
    if (!h || !dA || !dp || !it || !tres2 || maxit < 1) return PPS_ERR_BAD_ARG;
    const int method = h->opt.method;
    if ((method != PPS_MG_FP64 && method != PPS_MG_MIXED) || norm != 2 || h->opt.check_setup)
        return PPS_ERR_UNSUPPORTED;
    if (h->device < 0 || h->device >= kMaxDevices || !h->pinned) return PPS_ERR_BAD_ARG;
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        base = scratch_at[h->device];
    }
    if (!base) return PPS_ERR_HIP;   // pps_init looks the address up
    const Dims d = dims_of(h->grid);
    const bool mixed = method == PPS_MG_MIXED;
    const double lengths[3] = {h->grid.xlength, h->grid.ylength, h->grid.zlength};
    const size_t need = noalloc_bytes(d, lengths, mixed);
    if (need == 0) return PPS_ERR_BAD_SIZE;
    if (need > kScratchBytes) return PPS_ERR_SCRATCH;
    PPS_HIP(hipSetDevice(h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // a solver state whose buffers are pieces of the scratch array, used
    // for this call only (nothing of an earlier call is trusted: another
    // handle may have written the array since)
    Arena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = kScratchBytes;
    FastState local;
    local.d = d;
    local.arena = &arena;
    FastState *keep = h->fast;
    h->fast = &local;
    const int rc = fast_solve(h, dA, dp, maxError, maxit, 2, it, tres2, s);
    h->fast = keep;
    if (rc != PPS_OK) return rc;
    return *it > maxit ? PPS_ERR_NOT_CONVERGED : PPS_OK;
}

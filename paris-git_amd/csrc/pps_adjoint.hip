// pps_adjoint.hip -- fused face kernel of the adjoint gradient.
//
// The operator M p = b has, for the face f between cells L and R with face
// value c_f ("1/rho at the face", c_f = 2/(rho_L+rho_R), modules.f90:2737-
// 2742) and geometric weight w_f = dt/(dx dxh), the entries
// M_LL += w_f c_f, M_RR += w_f c_f, M_LR = M_RL -= w_f c_f.  For any
// objective J(p) with adjoint M lam = dJ/dp, dJ/dc_f = -lam^T (dM/dc_f) p
//   = -w_f (lam_L - lam_R) (p_L - p_R).
// The gradient is that of the map k_faces_to_A (A = w c, no masks): a
// face with c_f = 0 still has a gradient; only wall faces (the last face of
// a non-periodic axis, which couples to no cell) get 0.  A caller whose
// system also carries masks (SetupPoisson's umask, ...) multiplies by them.
#include "pps_internal.h"
#include "pps_xcd.h"

namespace pps {

struct FaceK {
    int nx, ny, nz, ng, X, Y;
    int px, py, pz;
};

// Workgroups of 64 x 16 threads own (i, j) columns and march through zc
// planes (a one-dimensional grid of gx x gy tiles x z chunks, numbered XCD
// by XCD so that neighbouring tiles share an L2).  Each plane of p and lam
// is fetched once: the +z values are requested two planes ahead and
// carried, so a step waits on no load of its own but the +x / +y edge
// neighbours; the +x neighbour comes from the next lane of the wavefront (a
// shuffle; the last lane of a row segment and the last cell of a row load
// it), the +y neighbour from the next thread row through LDS (the last
// thread row and the last row of the grid load it).  The gradient planes
// are written once, with non-temporal stores.  Threads past the grid work
// on a clamped cell and store nothing.
// Algorithmic bytes: p, lam read (16 B), three gradient planes written
// (24 B) per cell.
constexpr int kFgRows = 16;
__global__ void __launch_bounds__(64 * kFgRows)
k_face_grad(FaceK d, int gx, int gy, int zc, const double *__restrict__ p, const double *__restrict__ lam,
            const double *__restrict__ wx, const double *__restrict__ wy,
            const double *__restrict__ wz, double *__restrict__ g) {
    __shared__ double sp[2][kFgRows][64], sl[2][kFgRows][64];
    const int bid = xcd_order(blockIdx.x, gridDim.x);
    const int bx = bid % gx, by = (bid / gx) % gy, bz = bid / (gx * gy);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = bx * 64 + tx, j = by * kFgRows + ty;
    const int k0 = bz * zc;
    const int k1 = min(k0 + zc, d.nz);
    const bool in = i < d.nx && j < d.ny;
    const int ic = min(i, d.nx - 1), jc = min(j, d.ny - 1);
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long Py = d.X, Pz = (long long)d.X * d.Y;
    const long long col = (ic + d.ng) + Py * (jc + d.ng);
    // +x / +y neighbour offsets (wrapped on the last cell of a periodic
    // axis; unused on a wall face)
    const bool ox = ic < d.nx - 1 || d.px, oy = jc < d.ny - 1 || d.py;
    const long long ex = ic < d.nx - 1 ? 1 : -(long long)(d.nx - 1);
    const long long ey = jc < d.ny - 1 ? Py : -(long long)(d.ny - 1) * Py;
    const double gxw = ox ? -wx[ic] : 0.0, gyw = oy ? -wy[jc] : 0.0;
    // neighbours this thread loads itself instead of taking them from
    // another thread
    const bool xl = tx == 63 || ic == d.nx - 1;
    const bool yl = ty == kFgRows - 1 || jc == d.ny - 1;
    const long long cc = ic + (long long)d.nx * jc, sz = (long long)d.nx * d.ny;
    // padded index of plane kk of this column: wrapped past the last plane
    // of a periodic z axis, else clamped (the value is then not used)
    auto plane = [&](int kk) -> long long {
        const int kw = kk < d.nz ? kk : (d.pz ? kk % d.nz : d.nz - 1);
        return col + Pz * (kw + d.ng);
    };
    long long q = plane(k0), q1 = plane(k0 + 1);
    double pc = p[q], lc = lam[q];
    double p1 = p[q1], l1 = lam[q1];
    for (int k = k0; k < k1; ++k) {
        const int buf = k & 1;
        // requests: plane k+2, and the edge neighbours of plane k
        const long long q2 = plane(k + 2);
        const double p2 = p[q2], l2 = lam[q2];
        double pxe = 0.0, lxe = 0.0, pye = 0.0, lye = 0.0;
        if (xl) {
            pxe = p[q + ex];
            lxe = lam[q + ex];
        }
        if (yl) {
            pye = p[q + ey];
            lye = lam[q + ey];
        }
        sp[buf][ty][tx] = pc;
        sl[buf][ty][tx] = lc;
        const double pxs = __shfl_down(pc, 1, 64), lxs = __shfl_down(lc, 1, 64);
        if (!xl) {
            pxe = pxs;
            lxe = lxs;
        }
        __syncthreads();
        if (!yl) {
            pye = sp[buf][ty + 1][tx];
            lye = sl[buf][ty + 1][tx];
        }
        if (in) {
            // +z neighbour: the next plane, or the first one (periodic z)
            const bool oz = k < d.nz - 1 || d.pz;
            const long long c = cc + sz * k;
            __builtin_nontemporal_store(gxw * (lc - lxe) * (pc - pxe), g + c);
            __builtin_nontemporal_store(gyw * (lc - lye) * (pc - pye), g + n + c);
            __builtin_nontemporal_store(oz ? -wz[k] * (lc - l1) * (pc - p1) : 0.0, g + 2 * n + c);
        }
        pc = p1;
        lc = l1;
        p1 = p2;
        l1 = l2;
        q = q1;
        q1 = q2;
    }
}

// face values -> 8-slot system (the map k_face_grad differentiates)
__global__ void __launch_bounds__(256)
k_faces_to_A(FaceK d, const double *__restrict__ cf, const double *__restrict__ wx,
             const double *__restrict__ wy, const double *__restrict__ wz,
             const double *__restrict__ b, double *__restrict__ A) {
    // the same products and sum as the fast solve's rows from face values
    // (RowsC in pps_fast.hip), which form this system without writing it
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long long n = (long long)d.nx * d.ny * d.nz;
    const long long c = i + (long long)d.nx * (j + (long long)d.ny * k);
    const long long sy = d.nx, sz = (long long)d.nx * d.ny;
    const double *cx = cf, *cy = cf + n, *cz = cf + 2 * n;
    double a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0, a6 = 0.0;
    if (i > 0) a1 = wx[i - 1] * cx[c - 1];
    else if (d.px) a1 = wx[d.nx - 1] * cx[c + d.nx - 1];
    if (i < d.nx - 1 || d.px) a2 = wx[i] * cx[c];
    if (j > 0) a3 = wy[j - 1] * cy[c - sy];
    else if (d.py) a3 = wy[d.ny - 1] * cy[c + (long long)(d.ny - 1) * sy];
    if (j < d.ny - 1 || d.py) a4 = wy[j] * cy[c];
    if (k > 0) a5 = wz[k - 1] * cz[c - sz];
    else if (d.pz) a5 = wz[d.nz - 1] * cz[c + (long long)(d.nz - 1) * sz];
    if (k < d.nz - 1 || d.pz) a6 = wz[k] * cz[c];
    A[c] = a1; A[n + c] = a2; A[2 * n + c] = a3; A[3 * n + c] = a4;
    A[4 * n + c] = a5; A[5 * n + c] = a6;
    A[6 * n + c] = ((((a1 + a2) + a3) + a4) + a5) + a6;
    A[7 * n + c] = b[c];
}

}  // namespace pps

using namespace pps;

extern "C" int pps_faces_to_A_device(const pps_grid *grid, const double *cface, const double *wx,
                                     const double *wy, const double *wz, const double *b,
                                     double *A, void *stream) {
    if (!grid || !cface || !wx || !wy || !wz || !b || !A) return PPS_ERR_BAD_ARG;
    const Dims D = dims_of(*grid);
    hipStream_t s = static_cast<hipStream_t>(stream);
    FaceK d;
    d.nx = D.nx; d.ny = D.ny; d.nz = D.nz; d.ng = D.ng; d.X = D.X(); d.Y = D.Y();
    d.px = D.per[0]; d.py = D.per[1]; d.pz = D.per[2];
    const dim3 grid3((unsigned)((d.nx + 63) / 64), (unsigned)((d.ny + 3) / 4), (unsigned)d.nz);
    k_faces_to_A<<<grid3, dim3(64, 4, 1), 0, s>>>(d, cface, wx, wy, wz, b, A);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

extern "C" int pps_face_gradient_device(const pps_grid *grid, const double *dp, const double *dlam,
                                        const double *wx, const double *wy, const double *wz,
                                        double *dgrad, void *stream) {
    if (!grid || !dp || !dlam || !wx || !wy || !wz || !dgrad) return PPS_ERR_BAD_ARG;
    const Dims D = dims_of(*grid);
    if (D.nx < 1 || D.ny < 1 || D.nz < 1 || D.ng < 1) return PPS_ERR_BAD_ARG;
    hipStream_t s = static_cast<hipStream_t>(stream);
    FaceK d;
    d.nx = D.nx; d.ny = D.ny; d.nz = D.nz; d.ng = D.ng; d.X = D.X(); d.Y = D.Y();
    d.px = D.per[0]; d.py = D.per[1]; d.pz = D.per[2];
    // about 4 workgroups of 1024 threads per CU over (x, y) tiles x z chunks
    const int gx = (d.nx + 63) / 64, gy = (d.ny + kFgRows - 1) / kFgRows;
    long long chunks = 1024 / ((long long)gx * gy);
    if (chunks < 1) chunks = 1;
    if (chunks > d.nz) chunks = d.nz;
    const int zc = (int)((d.nz + chunks - 1) / chunks);
    chunks = (d.nz + zc - 1) / zc;
    k_face_grad<<<(unsigned)(gx * gy * chunks), dim3(64, kFgRows, 1), 0, s>>>(d, gx, gy, zc, dp, dlam, wx, wy,
                                                                           wz, dgrad);
    PPS_HIP(hipGetLastError());
    PPS_HIP(hipStreamSynchronize(s));
    return PPS_OK;
}

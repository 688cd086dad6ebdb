// pps_row.h -- one row of the PARIS pressure system from the fields:
// SetupPoisson (modules.f90:2722-2753), Poisson_BCs (modules.f90:2755-2892)
// and the solid-cell rule of check_and_debug_Poisson (modules.f90:2905-2945).
// Shared by the assembly kernel, which writes the 8-slot A, and by the
// solve that packs its operator straight from the fields.  The expressions
// keep the reference's operand order and are evaluated without fp
// contraction, so both give the same row bit for bit.
#pragma once

#include "pps_internal.h"

namespace pps {

// Boundary codes of the six faces (x-, y-, z-, x+, y+, z+: bdry_cond) and
// BoundaryPressure (-x, +x, -y, +y, -z, +z: modules.f90:878-883).
struct BcK {
    int b[6];
    double bp[6];
};

// Poisson_BCs (modules.f90:2755-2892) applied to the row of cell (i, j, k),
// block by block in the reference's statement order, so that where one cell
// lies on several faces the later block overwrites what an earlier one set,
// as the whole-plane assignments of the reference do.  Codes 3 / 4: the
// outward coupling is dropped (flux condition).  Code 5: P_0 - P_inner / 3 =
// 2/3 P_b, corners adding the conditions of both faces (pbc: which of x-, x+,
// y-, y+ carry code 5).  Code 6 (x+ only): P_b and du/dn at the face (du: the
// cell's (u(i) - u(i-1)) / dx(i)).  The z+ block of code 5 writes its
// y-coupling corner entries into the ks plane (modules.f90:2884-2887): the
// four corner cells of the first plane take them last.  FLUX_ONLY: codes 0-4
// only (the callers that refuse codes 5 and 6), which leaves the flux
// condition of codes 3 / 4 in the same block order.
template <bool FLUX_ONLY = false>
__device__ __forceinline__ void poisson_bcs(const BcK &bc, int nx, int ny, int nz, int i, int j, int k,
                                            double du, double a[8]) {
    const int ie = nx - 1, je = ny - 1, ke = nz - 1;
    if constexpr (FLUX_ONLY) {
        if (i == 0 && bc.b[0] == 3) { a[6] = a[6] - a[0]; a[0] = 0.0; }
        if (i == ie && (bc.b[3] == 4 || bc.b[3] == 3)) { a[6] = a[6] - a[1]; a[1] = 0.0; }
        if (j == 0 && bc.b[1] == 3) { a[6] = a[6] - a[2]; a[2] = 0.0; }
        if (j == je && bc.b[4] == 3) { a[6] = a[6] - a[3]; a[3] = 0.0; }
        if (k == 0 && bc.b[2] == 3) { a[6] = a[6] - a[4]; a[4] = 0.0; }
        if (k == ke && bc.b[5] == 3) { a[6] = a[6] - a[5]; a[5] = 0.0; }
        return;
    }
    const double c23 = 2.0 / 3.0, c13 = 1.0 / 3.0;
    const double pbc0 = bc.b[0] == 5 ? 1.0 : 0.0, pbc1 = bc.b[3] == 5 ? 1.0 : 0.0;
    const double pbc2 = bc.b[1] == 5 ? 1.0 : 0.0, pbc3 = bc.b[4] == 5 ? 1.0 : 0.0;
    // x- (2768-2781)
    if (i == 0) {
        if (bc.b[0] == 3) { a[6] = a[6] - a[0]; a[0] = 0.0; }
        else if (bc.b[0] == 5) {
            a[7] = c23 * bc.bp[0];
            a[6] = 1.0;
            for (int s = 0; s < 6; ++s) a[s] = 0.0;
            a[1] = c13;
        }
    }
    // x+ (2784-2801)
    if (i == ie) {
        if (bc.b[3] == 4 || bc.b[3] == 3) { a[6] = a[6] - a[1]; a[1] = 0.0; }
        else if (bc.b[3] == 5) {
            a[7] = c23 * bc.bp[1];
            a[6] = 1.0;
            for (int s = 1; s < 6; ++s) a[s] = 0.0;
            a[0] = c13;
        } else if (bc.b[3] == 6) {
            a[6] = a[6] - a[0] * 2.0 / 3.0 + a[1] * 5.0 / 3.0;
            a[7] = a[7] + 8.0 / 3.0 * a[1] * bc.bp[1] + du;
            a[0] = a[0] * 1.0 / 3.0;
            a[1] = 0.0;
        }
    }
    // y- (2804-2822), y+ (2824-2843)
    for (int side = 0; side < 2; ++side) {
        if (j != (side ? je : 0)) continue;
        const int code = bc.b[side ? 4 : 1];
        const int out = side ? 3 : 2;
        if (code == 3) { a[6] = a[6] - a[out]; a[out] = 0.0; }
        else if (code == 5) {
            const double P = bc.bp[side ? 3 : 2];
            a[7] = c23 * P;
            if (i == 0) a[7] = c23 * (P + bc.bp[0] * pbc0);
            if (i == ie) a[7] = c23 * (P + bc.bp[1] * pbc1);
            a[6] = 1.0;
            if (i == 0) a[6] = 1.0 + pbc0;
            if (i == ie) a[6] = 1.0 + pbc1;
            for (int s = 0; s < 6; ++s) a[s] = 0.0;
            if (i == 0) a[1] = c13 * pbc0;
            if (i == ie) a[0] = c13 * pbc1;
            a[side ? 2 : 3] = c13;
        }
    }
    // z- (2845-2867), z+ (2869-2891)
    for (int side = 0; side < 2; ++side) {
        const int code = bc.b[side ? 5 : 2];
        if (k == (side ? ke : 0)) {
            const int out = side ? 5 : 4;
            if (code == 3) { a[6] = a[6] - a[out]; a[out] = 0.0; }
            else if (code == 5) {
                const double P = bc.bp[side ? 5 : 4];
                a[7] = c23 * P;
                if (i == 0 && j == 0) a[7] = c23 * ((P + bc.bp[0] * pbc0) + bc.bp[2] * pbc2);
                if (i == ie && j == 0) a[7] = c23 * ((P + bc.bp[1] * pbc1) + bc.bp[2] * pbc2);
                if (i == 0 && j == je) a[7] = c23 * ((P + bc.bp[0] * pbc0) + bc.bp[3] * pbc3);
                if (i == ie && j == je) a[7] = c23 * ((P + bc.bp[1] * pbc1) + bc.bp[3] * pbc3);
                a[6] = 1.0;
                if (i == 0 && j == 0) a[6] = (1.0 + pbc0) + pbc2;
                if (i == ie && j == 0) a[6] = (1.0 + pbc1) + pbc2;
                if (i == 0 && j == je) a[6] = (1.0 + pbc0) + pbc3;
                if (i == ie && j == je) a[6] = (1.0 + pbc1) + pbc3;
                for (int s = 0; s < 6; ++s) a[s] = 0.0;
                if (i == 0 && (j == 0 || j == je)) a[1] = c13 * pbc0;
                if (i == ie && (j == 0 || j == je)) a[0] = c13 * pbc1;
            }
        }
        // the y-coupling corner entries of both blocks land in the ks plane
        if (code == 5 && k == 0) {
            if ((i == 0 || i == ie) && j == 0) a[3] = c13 * pbc2;
            if ((i == 0 || i == ie) && j == je) a[2] = c13 * pbc3;
        }
        if (code == 5 && k == (side ? ke : 0)) a[side ? 4 : 5] = c13;
    }
}

// Inputs of SetupPoisson: face velocities, face masks (SetPressureBC) and
// density, padded like p; per axis {dx, dxh} with ghosts (n + 2 ng values
// each, padded index = Fortran index - 1); dt and the volume source.
struct RowK {
    const double *u, *v, *w, *um, *vm, *wm, *rho;
    const double *ax, *ay, *az;
    double dt, vs;
    BcK bc;
};

// The coupling of a cell to a neighbour across one face
// (modules.f90:2737-2742): 2 dt mask / (d (dh (r_nb + r_c))).
__device__ __forceinline__ double face_coupling(double dt, double mask, double d, double dh,
                                                double r_nb, double r_c) {
#pragma clang fp contract(off)
    return 2.0 * dt * mask / (d * dh * (r_nb + r_c));
}

// Row of interior cell (i, j, k) (padded index q; padded strides sy, sz) into
// a[0..7]; returns 1 when the solid rule meets a row that is not a closed
// solid cell (the caller's error flag), else 0.  FLUX_ONLY: boundary codes
// 0-4 only (poisson_bcs<true>).  raw (not null): the +x, +y, +z couplings
// as SetupPoisson forms them, before Poisson_BCs.
template <bool FLUX_ONLY = false>
__device__ __forceinline__ int setup_row(const RowK &R, int nx, int ny, int nz, int ng, long long sy,
                                         long long sz, int i, int j, int k, long long q, double a[8],
                                         double *raw = nullptr) {
#pragma clang fp contract(off)
    const int NX = nx + 2 * ng, NY = ny + 2 * ng, NZ = nz + 2 * ng;
    const double *dx = R.ax, *dxh = R.ax + NX, *dy = R.ay, *dyh = R.ay + NY, *dz = R.az, *dzh = R.az + NZ;
    const int I = i + ng, J = j + ng, K = k + ng;
    const double *rho = R.rho, *um = R.um, *vm = R.vm, *wm = R.wm;
    const double r0 = rho[q], dt = R.dt;
    a[0] = face_coupling(dt, um[q - 1], dx[I], dxh[I - 1], rho[q - 1], r0);
    a[1] = face_coupling(dt, um[q], dx[I], dxh[I], rho[q + 1], r0);
    a[2] = face_coupling(dt, vm[q - sy], dy[J], dyh[J - 1], rho[q - sy], r0);
    a[3] = face_coupling(dt, vm[q], dy[J], dyh[J], rho[q + sy], r0);
    a[4] = face_coupling(dt, wm[q - sz], dz[K], dzh[K - 1], rho[q - sz], r0);
    a[5] = face_coupling(dt, wm[q], dz[K], dzh[K], rho[q + sz], r0);
    a[6] = ((((a[0] + a[1]) + a[2]) + a[3]) + a[4]) + a[5];
    if (raw) { raw[0] = a[1]; raw[1] = a[3]; raw[2] = a[5]; }
    const double dudx = (R.u[q] - R.u[q - 1]) / dx[I];
    a[7] = -(((R.vs + dudx) + (R.v[q] - R.v[q - sy]) / dy[J]) + (R.w[q] - R.w[q - sz]) / dz[K]);
    poisson_bcs<FLUX_ONLY>(R.bc, nx, ny, nz, i, j, k, dudx, a);
    int bad = 0;
    if (a[6] < 1e-50) {
        const bool solid = um[q - 1] < 0.5 && um[q] < 0.5 && vm[q - sy] < 0.5 && vm[q] < 0.5 &&
                           wm[q - sz] < 0.5 && wm[q] < 0.5;
        const double mx = fmax(fmax(fmax(a[0], a[1]), fmax(a[2], a[3])), fmax(a[4], a[5]));
        const double mn = fmin(fmin(fmin(a[0], a[1]), fmin(a[2], a[3])), fmin(a[4], a[5]));
        if (!solid || a[7] > 1e-50 || mx > 1e-50 || mn < 0.0) bad = 1;
        a[6] = 1.0;
    }
    return bad;
}

}  // namespace pps

// cc_kernels.h -- kernels of the cell-centred (approximate) projection,
// incflo::ApplyCCProjection (src/projection/incflo_apply_cc_projection.cpp:199-487):
// the method-of-lines face velocity from cell velocities and the cell update
// from the MAC fluxes.
//
// Face velocity (MOL::ExtrapVelToFaces, AMReX-Hydro, restated): the normal
// component extrapolated to the face from both neighbouring cells with
// limited slopes, u_L = u_{i-1} + s_{i-1}/2 and u_R = u_i - s_i/2, upwinded:
// u_L when (u_L + u_R)/2 > 0, u_R when < 0, and 0 for u_L < 0 < u_R or a
// vanishing mean.  Slopes: Colella's fourth-order limited slope
//   s = sign(dc) min(2 min(|dl|, |dr|) [dl dr >= 0], |4/3 dc - (m_{i+1} + m_{i-1})/6|)
// with dl, dr the one-sided and dc the centred differences and m the
// monotonised-central slopes of the neighbours; the monotonised-central slope
// in the second cell from a non-periodic face and zero slope in the cell on
// it.  A non-periodic domain face takes the ghost value (walls 0, inflow
// faces the inflow velocity), a pressure (Dirichlet) face the value of the
// adjacent cell.
#pragma once

#include <hip/hip_runtime.h>

#include "box_kernels.h"
#include "mac_kernels.h"

namespace mgx {

// component c of the ghosted canonical velocity at cell p (one ghost layer;
// periodic directions wrapped into the valid range)
__device__ __forceinline__ double ccq(const KGeom& g, const double* ug, int c, int i, int j, int k) {
    if (g.per[0]) i = (i % g.n[0] + g.n[0]) % g.n[0];
    if (g.per[1]) j = (j % g.n[1] + g.n[1]) % g.n[1];
    if (g.per[2]) k = (k % g.n[2] + g.n[2]) % g.n[2];
    const long ug3 = (long)g.sg[0] * g.sg[1] * g.sg[2];
    return ug[c * ug3 + (i + 1) + (long)g.sg[0] * ((j + 1) + (long)g.sg[1] * (k + 1))];
}

__device__ __forceinline__ double mc_slope(double dl, double dr) {
    const double dc = 0.5 * (dl + dr);
    const double lim = dl * dr >= 0.0 ? 2.0 * fmin(fabs(dl), fabs(dr)) : 0.0;
    return copysign(fmin(lim, fabs(dc)), dc);
}

// limited slope of component d along d at cell p
__device__ double cc_slope(const KGeom& g, const double* ug, int d, const int p[3]) {
    const int n = g.n[d];
    const int c = p[d];
    auto q = [&](int off) {
        int r[3] = {p[0], p[1], p[2]};
        r[d] += off;
        return ccq(g, ug, d, r[0], r[1], r[2]);
    };
    if (!g.per[d] && (c == 0 || c == n - 1)) return 0.0;
    const double qm2 = (!g.per[d] && c < 2) ? 0.0 : q(-2);
    const double qm1 = q(-1), q0 = q(0), qp1 = q(1);
    const double qp2 = (!g.per[d] && c > n - 3) ? 0.0 : q(2);
    const double dl = q0 - qm1, dr = qp1 - q0;
    if (!g.per[d] && (c == 1 || c == n - 2)) return mc_slope(dl, dr);
    const double mm = mc_slope(qm1 - qm2, q0 - qm1);
    const double mp = mc_slope(qp1 - q0, qp2 - qp1);
    const double dc = 0.5 * (dl + dr);
    const double lim = dl * dr >= 0.0 ? 2.0 * fmin(fabs(dl), fabs(dr)) : 0.0;
    const double d4 = 4.0 / 3.0 * dc - (mp + mm) / 6.0;
    return copysign(fmin(lim, fabs(d4)), dc);
}

// face velocities of direction d (canonical faces, n_d + 1 along d)
static __global__ void __launch_bounds__(256) mol_faces_kernel(KGeom g, int d, int bclo_dirichlet,
                                                              int bchi_dirichlet,
                                                              const double* __restrict__ ug,
                                                              double* __restrict__ u) {
    const int f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1), f2 = g.n[2] + (d == 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)f0 * f1 * f2) return;
    int p[3] = {(int)(t % f0), (int)((t / f0) % f1), (int)(t / ((long)f0 * f1))};
    const int n = g.n[d];
    const int f = p[d];
    double v;
    if (!g.per[d] && (f == 0 || f == n)) {
        const bool hi = f == n;
        int q[3] = {p[0], p[1], p[2]};
        if ((hi ? bchi_dirichlet : bclo_dirichlet) != 0) q[d] = hi ? n - 1 : 0;  // interior cell
        else q[d] = hi ? n : -1;                                                  // ghost cell
        v = ccq(g, ug, d, q[0], q[1], q[2]);
    } else {
        int L[3] = {p[0], p[1], p[2]}, R[3] = {p[0], p[1], p[2]};
        L[d] = f - 1;
        R[d] = f;
        if (g.per[d]) {
            if (L[d] < 0) L[d] += n;
            if (R[d] >= n) R[d] -= n;
        }
        const double uL = ccq(g, ug, d, L[0], L[1], L[2]) + 0.5 * cc_slope(g, ug, d, L);
        const double uR = ccq(g, ug, d, R[0], R[1], R[2]) - 0.5 * cc_slope(g, ug, d, R);
        const double avg = 0.5 * (uL + uR);
        v = avg > 0.0 ? uL : (avg < 0.0 ? uR : 0.0);
        if (uL < 0.0 && uR > 0.0) v = 0.0;
    }
    const MGeomK mk{{g.n[0], g.n[1], g.n[2]}, {g.per[0], g.per[1], g.per[2]}, {0, 0, 0}, {0, 0, 0}};
    u[fidx(mk, d, p[0], p[1], p[2])] = v;
}

struct CCUpdateArgs {
    KGeom g;
    const BoxInfo* boxes;
    const Tile* tiles;
    const double* ug;          // the projected field before the update (canonical, ghosted)
    const double* F[3];        // MAC fluxes -beta grad phi (canonical faces)
    const double* phi;         // canonical cells
    double* vel; int vel_ng;
    const double* vel_old; int vold_ng;  // incremental: added back
    double* gp; int gp_ng;
    double* p; int p_ng;
    const double* rho; int rho_ng;
    double const_rho;          // > 0: constant density ro_0
    double scale;
    int incremental;
};

// u = u' + gphi (+ u_old), p_cc = phi (+=), gp = -gphi rho / s (-=), with
// gphi the mean of the two MAC fluxes of each direction (average_mac_to_ccvel)
static __global__ void __launch_bounds__(256) cc_update_kernel(CCUpdateArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = a.boxes[bi];
    const KGeom& g = a.g;
    const MGeomK mk{{g.n[0], g.n[1], g.n[2]}, {g.per[0], g.per[1], g.per[2]}, {0, 0, 0}, {0, 0, 0}};
    const double r = a.const_rho > 0.0
                         ? a.const_rho
                         : a.rho[field_offset(b, a.rho_ng, 1, false) + box_index(b, a.rho_ng, false, i, j, k, 0)];
    const double ros = r / a.scale;
    const long ov = field_offset(b, a.vel_ng, 3, false), og = field_offset(b, a.gp_ng, 3, false);
    for (int c = 0; c < 3; ++c) {
        int q[3] = {i, j, k};
        const double flo = a.F[c][fidx(mk, c, q[0], q[1], q[2])];
        q[c] += 1;
        const double fhi = a.F[c][fidx(mk, c, q[0], q[1], q[2])];
        const double gphi = 0.5 * (flo + fhi);
        double u = ccq(g, a.ug, c, i, j, k) + gphi;
        if (a.incremental && a.vel_old)
            u += a.vel_old[field_offset(b, a.vold_ng, 3, false) + box_index(b, a.vold_ng, false, i, j, k, c)];
        a.vel[ov + box_index(b, a.vel_ng, false, i, j, k, c)] = u;
        double& gpc = a.gp[og + box_index(b, a.gp_ng, false, i, j, k, c)];
        gpc = a.incremental ? gpc - gphi * ros : -gphi * ros;
    }
    const double ph = a.phi[i + (long)g.n[0] * (j + (long)g.n[1] * k)];
    double& pc = a.p[field_offset(b, a.p_ng, 1, false) + box_index(b, a.p_ng, false, i, j, k, 0)];
    pc = a.incremental ? pc + ph : ph;
}

}  // namespace mgx

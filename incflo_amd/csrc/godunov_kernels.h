// godunov_kernels.h -- Godunov PPM advection on the GPU (periodic, single
// level, no EB): the velocity extrapolation to faces at t^{n+1/2} and the
// edge states / convective term of the velocity in convective form.
//
// Method (the same steps as oracle/orc_godunov.c):
//   1. PPM (Colella & Woodward 1984) traced face values Im/Ip per cell, per
//      direction d and component n, over |u| dt (cell velocity q_d for the
//      extrapolation, the face velocities otherwise);
//   2. advective face velocities: the Burgers Riemann solution of the normal
//      component's traced values (extrapolation), or the given u_mac;
//   3. upwinded edge states E[d][n];
//   4. corner-transport-upwind intermediate states T[d][e][n] (d-faces,
//      corrected along e with dt/3), upwinded by the d-face velocity;
//   5. final states: both transverse corrections (dt/2) + dt/2 force, then
//      the Burgers rule (normal velocity) or upwinding by u_mac (states);
//   6. conv = -div(u_mac s) + s div(u_mac).
// Arrays are canonical periodic: cells x fastest; d-faces have n_d entries,
// face i being the low face of cell i.  One thread per cell / face.
#pragma once

#include <hip/hip_runtime.h>

#include "box_kernels.h"
#include "nodal_kernels.h"

// separate multiplies and adds, as in the FP64 oracle (no contraction to FMA)
#pragma clang fp contract(off)

namespace mgx {

constexpr double kSmallVel = 1e-10;

struct GdGeom {
    int n[3];
    double dx[3];
};

struct GdArrays {
    const double* q[3];      // cell components
    const double* f[3];      // forces (nullptr: none)
    const double* adv[3];    // given face velocities (edge-state mode) or nullptr
    double* Im[3][3];        // [d][n]
    double* Ip[3][3];
    double* ad[3];           // advective face velocities
    double* E[3][3];         // upwinded edge states [d][n]
    double* T[3][3][3];      // intermediate states [d][e][n], d != e
    double* out[3][3];       // final states [d][n] (extrapolation: [d][d] only)
    double dt;
    int extrap;              // 1: ExtrapVelToFaces, 0: edge states with u_mac
    int nc;                  // components in q (3 for the velocity, 1..3 for scalars)
    int cons[3];             // conservative transverse corrections for component n
};

__device__ __forceinline__ int gd_wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ long gd_at(const GdGeom& g, int i, int j, int k) {
    return (long)gd_wrap(i, g.n[0]) + (long)g.n[0] * ((long)gd_wrap(j, g.n[1]) + (long)g.n[1] * gd_wrap(k, g.n[2]));
}

__device__ __forceinline__ long gd_sh(const GdGeom& g, int i, int j, int k, int d, int s) {
    return gd_at(g, i + (d == 0) * s, j + (d == 1) * s, k + (d == 2) * s);
}

__device__ __forceinline__ double gd_vl(double a, double b, double c) {
    const double d = 0.5 * (c - a);
    if ((c - b) * (b - a) <= 0.0) return 0.0;
    const double m = fmin(fabs(d), fmin(2.0 * fabs(b - a), 2.0 * fabs(c - b)));
    return d >= 0.0 ? m : -m;
}

__device__ __forceinline__ void gd_parabola(double sm2, double sm1, double s0, double sp1, double sp2,
                                            double& sL, double& sR) {
    double r = 0.5 * (s0 + sp1) - (gd_vl(s0, sp1, sp2) - gd_vl(sm1, s0, sp1)) / 6.0;
    double l = 0.5 * (sm1 + s0) - (gd_vl(sm1, s0, sp1) - gd_vl(sm2, sm1, s0)) / 6.0;
    r = fmax(fmin(s0, sp1), fmin(r, fmax(s0, sp1)));
    l = fmax(fmin(sm1, s0), fmin(l, fmax(sm1, s0)));
    if ((r - s0) * (s0 - l) <= 0.0) {
        l = s0;
        r = s0;
    } else if (fabs(r - s0) >= 2.0 * fabs(l - s0)) {
        r = 3.0 * s0 - 2.0 * l;
    } else if (fabs(l - s0) >= 2.0 * fabs(r - s0)) {
        l = 3.0 * s0 - 2.0 * r;
    }
    sL = l;
    sR = r;
}

__device__ __forceinline__ double gd_upwind(double lo, double hi, double a) {
    if (a > kSmallVel) return lo;
    if (a < -kSmallVel) return hi;
    return 0.5 * (lo + hi);
}

__device__ __forceinline__ double gd_burgers(double lo, double hi) {
    const double st = (lo + hi >= 0.0) ? lo : hi;
    const bool zero = (lo <= 0.0 && hi >= 0.0) || fabs(lo + hi) < kSmallVel;
    return zero ? 0.0 : st;
}

// 1. PPM traced values for every (d, n)
static __global__ void __launch_bounds__(256) gd_ppm_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k) || k >= g.n[2]) return;
    const long c = gd_at(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double uL, uR;
        if (a.extrap) {
            uL = uR = a.q[d][c];
        } else {
            uL = a.adv[d][c];
            uR = a.adv[d][gd_sh(g, i, j, k, d, 1)];
        }
        const long m2 = gd_sh(g, i, j, k, d, -2), m1 = gd_sh(g, i, j, k, d, -1);
        const long p1 = gd_sh(g, i, j, k, d, 1), p2 = gd_sh(g, i, j, k, d, 2);
        for (int n = 0; n < a.nc; ++n) {
            const double* s = a.q[n];
            const double s0 = s[c];
            double sL, sR;
            gd_parabola(s[m2], s[m1], s0, s[p1], s[p2], sL, sR);
            const double s6 = 6.0 * s0 - 3.0 * (sL + sR);
            double ip = s0, im = s0;
            if (uR > kSmallVel) {
                const double sg = uR * a.dt / g.dx[d];
                ip = sR - 0.5 * sg * ((sR - sL) - (1.0 - 2.0 * sg / 3.0) * s6);
            }
            if (uL < -kSmallVel) {
                const double sg = -uL * a.dt / g.dx[d];
                im = sL + 0.5 * sg * ((sR - sL) + (1.0 - 2.0 * sg / 3.0) * s6);
            }
            a.Ip[d][n][c] = ip;
            a.Im[d][n][c] = im;
        }
    }
}

// 2 + 3. advective face velocities and upwinded edge states
static __global__ void __launch_bounds__(256) gd_edge_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k) || k >= g.n[2]) return;
    const long f = gd_at(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long lc = gd_sh(g, i, j, k, d, -1);
        const double av = a.extrap ? gd_burgers(a.Ip[d][d][lc], a.Im[d][d][f]) : a.adv[d][f];
        a.ad[d][f] = av;
        for (int n = 0; n < a.nc; ++n) a.E[d][n][f] = gd_upwind(a.Ip[d][n][lc], a.Im[d][n][f], av);
    }
}

// 4. intermediate transverse states (reads ad and E of the whole grid)
static __global__ void __launch_bounds__(256) gd_trans_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k) || k >= g.n[2]) return;
    const long f = gd_at(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long lc = gd_sh(g, i, j, k, d, -1);
        const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
        for (int e = 0; e < 3; ++e) {
            if (e == d) continue;
            const double c = a.dt / (3.0 * g.dx[e]);
            const long lcp = gd_sh(g, li, lj, lk, e, 1), hcp = gd_sh(g, i, j, k, e, 1);
            const double* ae = a.ad[e];
            for (int n = 0; n < a.nc; ++n) {
                const double* Ee = a.E[e][n];
                double lo, hi;
                if (a.cons[n]) {
                    // flux differences minus q times the velocity difference
                    const double* q = a.q[n];
                    lo = a.Ip[d][n][lc] - c * (ae[lcp] * Ee[lcp] - ae[lc] * Ee[lc]) +
                         c * q[lc] * (ae[lcp] - ae[lc]);
                    hi = a.Im[d][n][f] - c * (ae[hcp] * Ee[hcp] - ae[f] * Ee[f]) +
                         c * q[f] * (ae[hcp] - ae[f]);
                } else {
                    lo = a.Ip[d][n][lc] - c * 0.5 * (ae[lcp] + ae[lc]) * (Ee[lcp] - Ee[lc]);
                    hi = a.Im[d][n][f] - c * 0.5 * (ae[hcp] + ae[f]) * (Ee[hcp] - Ee[f]);
                }
                a.T[d][e][n][f] = gd_upwind(lo, hi, a.ad[d][f]);
            }
        }
    }
}

// 5. final face states
static __global__ void __launch_bounds__(256) gd_final_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k) || k >= g.n[2]) return;
    const long f = gd_at(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
        const long lc = gd_sh(g, i, j, k, d, -1);
        const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
        const double c1 = a.dt / (2.0 * g.dx[e1]), c2 = a.dt / (2.0 * g.dx[e2]);
        const long l1 = gd_sh(g, li, lj, lk, e1, 1), l2 = gd_sh(g, li, lj, lk, e2, 1);
        const long h1 = gd_sh(g, i, j, k, e1, 1), h2 = gd_sh(g, i, j, k, e2, 1);
        const double* a1 = a.ad[e1];
        const double* a2 = a.ad[e2];
        for (int n = 0; n < a.nc; ++n) {
            if (a.extrap && n != d) continue;
            const double* t1 = a.T[e1][e2][n];
            const double* t2 = a.T[e2][e1][n];
            double lo, hi;
            if (a.cons[n]) {
                const double* q = a.q[n];
                lo = a.Ip[d][n][lc] - c1 * (a1[l1] * t1[l1] - a1[lc] * t1[lc]) +
                     c1 * q[lc] * (a1[l1] - a1[lc]) - c2 * (a2[l2] * t2[l2] - a2[lc] * t2[lc]) +
                     c2 * q[lc] * (a2[l2] - a2[lc]);
                hi = a.Im[d][n][f] - c1 * (a1[h1] * t1[h1] - a1[f] * t1[f]) +
                     c1 * q[f] * (a1[h1] - a1[f]) - c2 * (a2[h2] * t2[h2] - a2[f] * t2[f]) +
                     c2 * q[f] * (a2[h2] - a2[f]);
            } else {
                lo = a.Ip[d][n][lc] - c1 * 0.5 * (a1[l1] + a1[lc]) * (t1[l1] - t1[lc]) -
                     c2 * 0.5 * (a2[l2] + a2[lc]) * (t2[l2] - t2[lc]);
                hi = a.Im[d][n][f] - c1 * 0.5 * (a1[h1] + a1[f]) * (t1[h1] - t1[f]) -
                     c2 * 0.5 * (a2[h2] + a2[f]) * (t2[h2] - t2[f]);
            }
            if (a.f[n]) {
                lo += 0.5 * a.dt * a.f[n][lc];
                hi += 0.5 * a.dt * a.f[n][f];
            }
            a.out[d][n][f] = a.extrap ? gd_burgers(lo, hi) : gd_upwind(lo, hi, a.ad[d][f]);
        }
    }
}

// 6. conv[n] = -div(u_mac S[.][n]) + q_n div(u_mac) (the last term only for
// components in convective form); optional fluxes u_mac * S.  Outputs of
// component n go to slot n0 + n of ntot (conv: [slot][N], flux: [d][slot][N]).
static __global__ void __launch_bounds__(256) gd_conv_kernel(GdGeom g, GdArrays a, double* conv,
                                                             double* flux, int n0, int ntot) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k) || k >= g.n[2]) return;
    const long N = (long)g.n[0] * g.n[1] * g.n[2];
    const long c = gd_at(g, i, j, k);
    double divu = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) divu += (a.adv[d][gd_sh(g, i, j, k, d, 1)] - a.adv[d][c]) / g.dx[d];
    for (int n = 0; n < a.nc; ++n) {
        double div = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* st = a.out[d][n];
            const long cp = gd_sh(g, i, j, k, d, 1);
            div += (a.adv[d][cp] * st[cp] - a.adv[d][c] * st[c]) / g.dx[d];
            if (flux) flux[(long)(d * ntot + n0 + n) * N + c] = a.adv[d][c] * st[c];
        }
        conv[(long)(n0 + n) * N + c] = a.cons[n] ? -div : -div + a.q[n][c] * divu;
    }
}

// ---- boxed <-> canonical (periodic) ----------------------------------------

// canonical cells (ncomp, no ghosts) <- boxed cells (valid region)
static __global__ void __launch_bounds__(256) gd_cells_in_kernel(GdGeom g, const BoxInfo* boxes,
                                                                 const Tile* tiles, const double* boxed,
                                                                 int ng, int ncomp, double* canon) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long N = (long)g.n[0] * g.n[1] * g.n[2];
    const long o = field_offset(b, ng, ncomp, false);
    const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
    for (int q = 0; q < ncomp; ++q) canon[q * N + c] = boxed[o + box_index(b, ng, false, i, j, k, q)];
}

// boxed cells (valid region) <- canonical cells
static __global__ void __launch_bounds__(256) gd_cells_out_kernel(GdGeom g, const BoxInfo* boxes,
                                                                  const Tile* tiles, double* boxed,
                                                                  int ng, int ncomp, const double* canon) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long N = (long)g.n[0] * g.n[1] * g.n[2];
    const long o = field_offset(b, ng, ncomp, false);
    const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
    for (int q = 0; q < ncomp; ++q) boxed[o + box_index(b, ng, false, i, j, k, q)] = canon[q * N + c];
}

// faces along d (every box face incl. the high one) <-> periodic canonical faces.
// Boxed face fields with ncomp components are component-major per box.
static __global__ void __launch_bounds__(256) gd_faces_kernel(GdGeom g, const BoxInfo* boxes,
                                                              const Tile* tiles, int d, double* boxed,
                                                              int ng, int ncomp, double* canon,
                                                              long cstride, int to_box) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_FACE0 + d, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long vol = (long)(b.len[0] + (d == 0) + 2 * ng) * (b.len[1] + (d == 1) + 2 * ng) *
                     (b.len[2] + (d == 2) + 2 * ng);
    const long o = (long)ncomp * face_offset(b, ng, d) + box_index_face(b, ng, d, i, j, k);
    const long c = gd_at(g, i, j, k);
    for (int q = 0; q < ncomp; ++q) {
        if (to_box) boxed[o + q * vol] = canon[q * cstride + c];
        else if ((d == 0 ? i : (d == 1 ? j : k)) < g.n[d]) canon[q * cstride + c] = boxed[o + q * vol];
    }
}

}  // namespace mgx

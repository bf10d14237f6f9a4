// godunov_kernels.h -- Godunov PPM advection on the GPU (single level, no EB):
// the velocity extrapolation to faces at t^{n+1/2} and the edge states /
// fluxes / advection terms of the velocity, density and tracers, on periodic
// or bounded domains.
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
//      conservative components take the transverse terms as flux differences
//      minus the cell value times the velocity difference;
//   6. conv = -div(u_mac s) (+ s div(u_mac) for convective components).
// Domain faces of non-periodic directions follow the component's BC type
// (amrex::BCType values): next to an ext_dir / hoextrap face the PPM
// interface values are the ghost (boundary) value and a one-sided value from
// it; every state on a domain face comes from the interior-side trace
// (ext_dir: the ghost value; foextrap / hoextrap / reflect_even: the trace;
// reflect_odd: 0); in the extrapolation an outflow (foextrap) face's normal
// velocity may not point into the domain.
// Arrays: cell states are ghosted (2 layers: (n0+4)(n1+4)(n2+4), x fastest);
// traces are interior cells; face arrays of direction d have n_d + 1 entries
// along d (periodic: face n_d repeats face 0).  One thread per cell / face.
#pragma once

#include <hip/hip_runtime.h>

#include "box_kernels.h"
#include "nodal_kernels.h"

// separate multiplies and adds, as in the FP64 oracle (no contraction to FMA)
#pragma clang fp contract(off)

namespace mgx {

constexpr double kSmallVel = 1e-10;
constexpr int kGh = 2;  // ghost layers of the cell states

enum { GBC_RODD = -1, GBC_INT = 0, GBC_REVEN = 1, GBC_FOE = 2, GBC_EXT = 3, GBC_HOE = 4 };

struct GdGeom {
    int n[3];
    int per[3];
    double dx[3];
};

struct GdArrays {
    const double* q[3];      // ghosted cell components
    const double* f[3];      // forces (interior cells; nullptr: none)
    const double* adv[3];    // given face velocities (edge-state mode) or nullptr
    double* Im[3][3];        // [d][n], interior cells
    double* Ip[3][3];
    double* ad[3];           // advective face velocities
    double* E[3][3];         // upwinded edge states [d][n]
    double* T[3][3][3];      // intermediate states [d][e][n], d != e
    double* out[3][3];       // final states [d][n] (extrapolation: [d][d] only)
    double dt;
    int extrap;              // 1: ExtrapVelToFaces, 0: edge states with u_mac
    int nc;                  // components in q (3 for the velocity, 1..3 for scalars)
    int cons[3];             // conservative transverse corrections for component n
    int bc[3][6];            // BC type per component: lo x, y, z, hi x, y, z
};

__device__ __forceinline__ int gd_wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// interior cell (periodic directions wrap)
__device__ __forceinline__ long gd_cell(const GdGeom& g, int i, int j, int k) {
    if (g.per[0]) i = gd_wrap(i, g.n[0]);
    if (g.per[1]) j = gd_wrap(j, g.n[1]);
    if (g.per[2]) k = gd_wrap(k, g.n[2]);
    return (long)i + (long)g.n[0] * ((long)j + (long)g.n[1] * k);
}

// ghosted cell, i in [-kGh, n + kGh)
__device__ __forceinline__ long gd_gc(const GdGeom& g, int i, int j, int k) {
    return (long)(i + kGh) + (long)(g.n[0] + 2 * kGh) * ((long)(j + kGh) + (long)(g.n[1] + 2 * kGh) * (k + kGh));
}

// face of direction d, index along d in [0, n_d]
__device__ __forceinline__ long gd_face(const GdGeom& g, int d, int i, int j, int k) {
    return (long)i + (long)(g.n[0] + (d == 0)) * ((long)j + (long)(g.n[1] + (d == 1)) * k);
}

__device__ __forceinline__ double gd_vl(double a, double b, double c) {
    const double d = 0.5 * (c - a);
    if ((c - b) * (b - a) <= 0.0) return 0.0;
    const double m = fmin(fabs(d), fmin(2.0 * fabs(b - a), 2.0 * fabs(c - b)));
    return d >= 0.0 ? m : -m;
}

__device__ __forceinline__ double gd_one_sided(double sg, double s0, double s1, double s2) {
    double e = -0.2 * sg + 0.75 * s0 + 0.5 * s1 - 0.05 * s2;
    e = fmax(e, fmin(s0, s1));
    e = fmin(e, fmax(s0, s1));
    return e;
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

// state on a domain face from the interior-side value v and the ghost value gv
__device__ __forceinline__ double gd_bc_face(int bc, int side, double v, double gv, bool clip_in) {
    if (bc == GBC_EXT) return gv;
    if (bc == GBC_RODD) return 0.0;
    if (bc == GBC_FOE && clip_in) return side ? fmax(v, 0.0) : fmin(v, 0.0);
    return v;
}

// is face index id along d a domain face (non-periodic)?  side: 0 low, 1 high
__device__ __forceinline__ bool gd_bface(const GdGeom& g, int d, int id, int& side) {
    if (g.per[d]) return false;
    if (id == 0) { side = 0; return true; }
    if (id == g.n[d]) { side = 1; return true; }
    return false;
}

// ghost cell just outside domain face (i,j,k) of direction d
__device__ __forceinline__ long gd_ghost(const GdGeom& g, int d, int side, int i, int j, int k) {
    int c[3] = {i, j, k};
    c[d] = side ? g.n[d] : -1;
    return gd_gc(g, c[0], c[1], c[2]);
}

// 1. PPM traced values for every (d, n), interior cells
static __global__ void __launch_bounds__(256) gd_ppm_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k) || k >= g.n[2]) return;
    const long c = gd_cell(g, i, j, k);
    const long cg = gd_gc(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int p[3] = {i, j, k};
        const int id = p[d], n = g.n[d];
        const long st = d == 0 ? 1 : (d == 1 ? (long)(g.n[0] + 2 * kGh)
                                             : (long)(g.n[0] + 2 * kGh) * (g.n[1] + 2 * kGh));
        double uL, uR;
        if (a.extrap) {
            uL = uR = a.q[d][cg];
        } else {
            int h[3] = {i, j, k};
            h[d] += 1;
            uL = a.adv[d][gd_face(g, d, i, j, k)];
            uR = a.adv[d][gd_face(g, d, h[0], h[1], h[2])];
        }
        for (int m = 0; m < a.nc; ++m) {
            const double* s = a.q[m];
            const double sm2 = s[cg - 2 * st], sm1 = s[cg - st], s0 = s[cg], sp1 = s[cg + st],
                         sp2 = s[cg + 2 * st];
            double r = 0.5 * (s0 + sp1) - (gd_vl(s0, sp1, sp2) - gd_vl(sm1, s0, sp1)) / 6.0;
            double l = 0.5 * (sm1 + s0) - (gd_vl(sm1, s0, sp1) - gd_vl(sm2, sm1, s0)) / 6.0;
            r = fmax(fmin(s0, sp1), fmin(r, fmax(s0, sp1)));
            l = fmax(fmin(sm1, s0), fmin(l, fmax(sm1, s0)));
            if (!g.per[d]) {
                const int bl = a.bc[m][d], bh = a.bc[m][3 + d];
                if ((bl == GBC_EXT || bl == GBC_HOE) && id <= 1) {
                    const long b0 = cg - (long)id * st;  // first cell
                    const double e = gd_one_sided(s[b0 - st], s[b0], s[b0 + st], s[b0 + 2 * st]);
                    if (id == 0) { l = s[b0 - st]; r = e; }
                    else { l = e; }
                }
                if ((bh == GBC_EXT || bh == GBC_HOE) && id >= n - 2) {
                    const long b0 = cg + (long)(n - 1 - id) * st;  // last cell
                    const double e = gd_one_sided(s[b0 + st], s[b0], s[b0 - st], s[b0 - 2 * st]);
                    if (id == n - 1) { r = s[b0 + st]; l = e; }
                    else { r = e; }
                }
            }
            if ((r - s0) * (s0 - l) <= 0.0) {
                l = s0;
                r = s0;
            } else if (fabs(r - s0) >= 2.0 * fabs(l - s0)) {
                r = 3.0 * s0 - 2.0 * l;
            } else if (fabs(l - s0) >= 2.0 * fabs(r - s0)) {
                l = 3.0 * s0 - 2.0 * r;
            }
            const double s6 = 6.0 * s0 - 3.0 * (l + r);
            double ip = s0, im = s0;
            if (uR > kSmallVel) {
                const double sg = uR * a.dt / g.dx[d];
                ip = r - 0.5 * sg * ((r - l) - (1.0 - 2.0 * sg / 3.0) * s6);
            }
            if (uL < -kSmallVel) {
                const double sg = -uL * a.dt / g.dx[d];
                im = l + 0.5 * sg * ((r - l) + (1.0 - 2.0 * sg / 3.0) * s6);
            }
            a.Ip[d][m][c] = ip;
            a.Im[d][m][c] = im;
        }
    }
}

// face (i,j,k) of direction d inside the face grid?
__device__ __forceinline__ bool gd_on_faces(const GdGeom& g, int d, int i, int j, int k) {
    return i < g.n[0] + (d == 0) && j < g.n[1] + (d == 1) && k < g.n[2] + (d == 2);
}

// 2 + 3. advective face velocities and upwinded edge states
static __global__ void __launch_bounds__(256) gd_edge_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0] + 1, g.n[1] + 1, i, j, k)) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!gd_on_faces(g, d, i, j, k)) continue;
        const long f = gd_face(g, d, i, j, k);
        const int p[3] = {i, j, k};
        int side = 0;
        const bool bnd = gd_bface(g, d, p[d], side);
        const long lc = gd_cell(g, i - (d == 0), j - (d == 1), k - (d == 2));
        const long hc = gd_cell(g, i, j, k);
        double av;
        if (!a.extrap) {
            av = a.adv[d][f];
        } else if (bnd) {
            const double v = side ? a.Ip[d][d][lc] : a.Im[d][d][hc];
            const double b = gd_bc_face(a.bc[d][3 * side + d], side, v,
                                        a.q[d][gd_ghost(g, d, side, i, j, k)], true);
            av = gd_burgers(b, b);
        } else {
            av = gd_burgers(a.Ip[d][d][lc], a.Im[d][d][hc]);
        }
        a.ad[d][f] = av;
        for (int m = 0; m < a.nc; ++m) {
            double e;
            if (bnd) {
                const double v = side ? a.Ip[d][m][lc] : a.Im[d][m][hc];
                e = gd_bc_face(a.bc[m][3 * side + d], side, v, a.q[m][gd_ghost(g, d, side, i, j, k)],
                               false);
            } else {
                e = gd_upwind(a.Ip[d][m][lc], a.Im[d][m][hc], av);
            }
            a.E[d][m][f] = e;
        }
    }
}

// 4. intermediate transverse states (reads ad and E of the whole grid)
static __global__ void __launch_bounds__(256) gd_trans_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0] + 1, g.n[1] + 1, i, j, k)) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!gd_on_faces(g, d, i, j, k)) continue;
        const long f = gd_face(g, d, i, j, k);
        const int p[3] = {i, j, k};
        int side = -1;
        const bool bnd = gd_bface(g, d, p[d], side);
        for (int e = 0; e < 3; ++e) {
            if (e == d) continue;
            const double c = a.dt / (3.0 * g.dx[e]);
            const double* ae = a.ad[e];
            for (int m = 0; m < a.nc; ++m) {
                const double* Ee = a.E[e][m];
                double val[2] = {0.0, 0.0};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    if (bnd && s == side) continue;  // the ghost cell is not traced
                    int cc[3] = {i, j, k};
                    if (s == 0) cc[d] -= 1;
                    if (g.per[d]) cc[d] = gd_wrap(cc[d], g.n[d]);
                    const long ci = gd_cell(g, cc[0], cc[1], cc[2]);
                    int up[3] = {cc[0], cc[1], cc[2]};
                    up[e] += 1;
                    const long fl = gd_face(g, e, cc[0], cc[1], cc[2]);
                    const long fh = gd_face(g, e, up[0], up[1], up[2]);
                    const double tr = s == 0 ? a.Ip[d][m][ci] : a.Im[d][m][ci];
                    if (a.cons[m])
                        val[s] = tr - c * (ae[fh] * Ee[fh] - ae[fl] * Ee[fl]) +
                                 c * a.q[m][gd_gc(g, cc[0], cc[1], cc[2])] * (ae[fh] - ae[fl]);
                    else
                        val[s] = tr - c * 0.5 * (ae[fh] + ae[fl]) * (Ee[fh] - Ee[fl]);
                }
                a.T[d][e][m][f] =
                    bnd ? gd_bc_face(a.bc[m][3 * side + d], side, val[1 - side],
                                     a.q[m][gd_ghost(g, d, side, i, j, k)], false)
                        : gd_upwind(val[0], val[1], a.ad[d][f]);
            }
        }
    }
}

// 5. final face states
static __global__ void __launch_bounds__(256) gd_final_kernel(GdGeom g, GdArrays a) {
    int i, j, k;
    if (!grid_point(g.n[0] + 1, g.n[1] + 1, i, j, k)) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!gd_on_faces(g, d, i, j, k)) continue;
        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
        const long f = gd_face(g, d, i, j, k);
        const int p[3] = {i, j, k};
        int side = -1;
        const bool bnd = gd_bface(g, d, p[d], side);
        const double c1 = a.dt / (2.0 * g.dx[e1]), c2 = a.dt / (2.0 * g.dx[e2]);
        const double* a1 = a.ad[e1];
        const double* a2 = a.ad[e2];
        for (int m = 0; m < a.nc; ++m) {
            if (a.extrap && m != d) continue;
            const double* t1 = a.T[e1][e2][m];
            const double* t2 = a.T[e2][e1][m];
            double val[2] = {0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (bnd && s == side) continue;
                int cc[3] = {i, j, k};
                if (s == 0) cc[d] -= 1;
                if (g.per[d]) cc[d] = gd_wrap(cc[d], g.n[d]);
                const long ci = gd_cell(g, cc[0], cc[1], cc[2]);
                int u1[3] = {cc[0], cc[1], cc[2]}, u2[3] = {cc[0], cc[1], cc[2]};
                u1[e1] += 1;
                u2[e2] += 1;
                const long l1 = gd_face(g, e1, cc[0], cc[1], cc[2]), h1 = gd_face(g, e1, u1[0], u1[1], u1[2]);
                const long l2 = gd_face(g, e2, cc[0], cc[1], cc[2]), h2 = gd_face(g, e2, u2[0], u2[1], u2[2]);
                double v = s == 0 ? a.Ip[d][m][ci] : a.Im[d][m][ci];
                if (a.cons[m]) {
                    const double qc = a.q[m][gd_gc(g, cc[0], cc[1], cc[2])];
                    v = v - c1 * (a1[h1] * t1[h1] - a1[l1] * t1[l1]) + c1 * qc * (a1[h1] - a1[l1]) -
                        c2 * (a2[h2] * t2[h2] - a2[l2] * t2[l2]) + c2 * qc * (a2[h2] - a2[l2]);
                } else {
                    v = v - c1 * 0.5 * (a1[h1] + a1[l1]) * (t1[h1] - t1[l1]) -
                        c2 * 0.5 * (a2[h2] + a2[l2]) * (t2[h2] - t2[l2]);
                }
                if (a.f[m]) v += 0.5 * a.dt * a.f[m][ci];
                val[s] = v;
            }
            double o;
            if (bnd) {
                const double b = gd_bc_face(a.bc[m][3 * side + d], side, val[1 - side],
                                            a.q[m][gd_ghost(g, d, side, i, j, k)], a.extrap != 0);
                o = a.extrap ? gd_burgers(b, b) : b;
            } else {
                o = a.extrap ? gd_burgers(val[0], val[1]) : gd_upwind(val[0], val[1], a.ad[d][f]);
            }
            a.out[d][m][f] = o;
        }
    }
}

// 6. conv[n] = -div(u_mac S[.][n]) + q_n div(u_mac) (the last term only for
// components in convective form); optional fluxes u_mac * S.  Outputs of
// component n go to slot n0 + n of ntot (conv: [slot][N]; flux: direction d
// at fl[d] + slot * nface(d)).
struct GdFluxOut {
    double* f[3];
};

static __global__ void __launch_bounds__(256) gd_conv_kernel(GdGeom g, GdArrays a, double* conv,
                                                             GdFluxOut flux, int n0) {
    int i, j, k;
    if (!grid_point(g.n[0] + 1, g.n[1] + 1, i, j, k)) return;
    // fluxes on every face (each thread: its faces of the three directions)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!flux.f[d] || !gd_on_faces(g, d, i, j, k)) continue;
        const long f = gd_face(g, d, i, j, k);
        const long nf = (long)(g.n[0] + (d == 0)) * (g.n[1] + (d == 1)) * (g.n[2] + (d == 2));
        for (int m = 0; m < a.nc; ++m) flux.f[d][(long)(n0 + m) * nf + f] = a.adv[d][f] * a.out[d][m][f];
    }
    if (i >= g.n[0] || j >= g.n[1] || k >= g.n[2]) return;
    const long N = (long)g.n[0] * g.n[1] * g.n[2];
    const long c = gd_cell(g, i, j, k);
    long fl[3], fh[3];
    double divu = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        fl[d] = gd_face(g, d, i, j, k);
        fh[d] = gd_face(g, d, i + (d == 0), j + (d == 1), k + (d == 2));
        divu += (a.adv[d][fh[d]] - a.adv[d][fl[d]]) / g.dx[d];
    }
    for (int m = 0; m < a.nc; ++m) {
        double div = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* st = a.out[d][m];
            div += (a.adv[d][fh[d]] * st[fh[d]] - a.adv[d][fl[d]] * st[fl[d]]) / g.dx[d];
        }
        conv[(long)(n0 + m) * N + c] =
            a.cons[m] ? -div : -div + a.q[m][gd_gc(g, i, j, k)] * divu;
    }
}

// ---- boxed <-> canonical ------------------------------------------------

// ghosted canonical cells (ncomp; interior only) <- boxed cells (valid region)
static __global__ void __launch_bounds__(256) gd_cells_in_kernel(GdGeom g, const BoxInfo* boxes,
                                                                 const Tile* tiles, const double* boxed,
                                                                 int ng, int ncomp, double* canon) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long NG = (long)(g.n[0] + 2 * kGh) * (g.n[1] + 2 * kGh) * (g.n[2] + 2 * kGh);
    const long o = field_offset(b, ng, ncomp, false);
    const long c = gd_gc(g, i, j, k);
    for (int q = 0; q < ncomp; ++q) canon[q * NG + c] = boxed[o + box_index(b, ng, false, i, j, k, q)];
}

// ghost layers of the ghosted canonical cells: periodic directions wrap (from
// the interior written before), others read the caller's ghost cells from the
// box holding the nearest valid cell (cells outside the domain in two or more
// non-periodic directions are never read by the stages and are set to 0)
static __global__ void __launch_bounds__(256) gd_ghosts_kernel(GdGeom g, const BoxInfo* boxes,
                                                               int nboxes, const double* boxed,
                                                               int ng, int ncomp, double* canon) {
    const int GX = g.n[0] + 2 * kGh, GY = g.n[1] + 2 * kGh;
    int ii, jj, kk;
    if (!grid_point(GX, GY, ii, jj, kk)) return;
    const int p[3] = {ii - kGh, jj - kGh, kk - kGh};
    int nout = 0, nonper = 0;
    int src[3];
    for (int d = 0; d < 3; ++d) {
        src[d] = p[d];
        if (p[d] < 0 || p[d] >= g.n[d]) {
            ++nout;
            if (g.per[d]) src[d] = gd_wrap(p[d], g.n[d]);
            else ++nonper;
        }
    }
    if (nout == 0) return;
    const long NG = (long)GX * GY * (g.n[2] + 2 * kGh);
    const long dst = gd_gc(g, p[0], p[1], p[2]);
    if (nonper == 0) {
        const long s = gd_gc(g, src[0], src[1], src[2]);
        for (int q = 0; q < ncomp; ++q) canon[q * NG + dst] = canon[q * NG + s];
        return;
    }
    if (nonper > 1) {
        for (int q = 0; q < ncomp; ++q) canon[q * NG + dst] = 0.0;
        return;
    }
    // one non-periodic direction out of range: the box holding the clamped cell
    int cl[3];
    for (int d = 0; d < 3; ++d) cl[d] = min(max(src[d], 0), g.n[d] - 1);
    for (int bi = 0; bi < nboxes; ++bi) {
        const BoxInfo& b = boxes[bi];
        if (cl[0] < b.lo[0] || cl[0] >= b.lo[0] + b.len[0] || cl[1] < b.lo[1] ||
            cl[1] >= b.lo[1] + b.len[1] || cl[2] < b.lo[2] || cl[2] >= b.lo[2] + b.len[2])
            continue;
        const long o = field_offset(b, ng, ncomp, false);
        for (int q = 0; q < ncomp; ++q)
            canon[q * NG + dst] = boxed[o + box_index(b, ng, false, src[0], src[1], src[2], q)];
        return;
    }
    for (int q = 0; q < ncomp; ++q) canon[q * NG + dst] = 0.0;
}

// interior canonical cells (ncomp, [q][N]) <- boxed cells (valid region)
static __global__ void __launch_bounds__(256) gd_plain_in_kernel(GdGeom g, const BoxInfo* boxes,
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

// faces along d (every box face incl. the high one) <-> canonical faces
// (n_d + 1 along d).  Boxed face fields with ncomp components are
// component-major per box; canonical component q at canon + q * cstride.
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
    const long c = gd_face(g, d, i, j, k);
    for (int q = 0; q < ncomp; ++q) {
        if (to_box) boxed[o + q * vol] = canon[q * cstride + c];
        else canon[q * cstride + c] = boxed[o + q * vol];
    }
}

}  // namespace mgx

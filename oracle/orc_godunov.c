/* orc_godunov.c -- FP64 CPU restatement of the Godunov (PPM) advection pieces
 * of incflo's predictor: ExtrapVelToFaces (face-normal velocities at
 * t^{n+1/2}) and the edge states / fluxes / convective term
 * (ComputeFluxesOnBoxFromState + ComputeDivergence + ComputeConvectiveTerm),
 * driven from src/convection/incflo_compute_MAC_projected_velocities.cpp:168-178
 * and src/convection/incflo_compute_advection_term.cpp:47-808.  TEST
 * INFRASTRUCTURE: the checker of the HIP kernels; never linked into the
 * product library.
 *
 * The reference's Godunov lives in AMReX-Hydro, which is not vendored, so
 * this restates the published method (single level, no EB):
 *   PPM (Colella & Woodward 1984): interface values from van Leer-limited
 *     slopes, bounded by the neighbours, CW84 monotonicity fix; the parabola
 *     is traced over the domain of dependence of dt with sigma = |u| dt / h:
 *       u > 0: Ip = sR - sigma/2 [(sR - sL) - (1 - 2 sigma/3) s6],  s6 = 6 s - 3 (sL + sR)
 *       u < 0: Im = sL + sigma/2 [(sR - sL) + (1 - 2 sigma/3) s6]
 *     (the other side takes the cell value);
 *   corner-transport-upwind transverse corrections (Colella 1990, Saltzman
 *     1994): intermediate states at d-faces corrected by one transverse
 *     direction e with dt/3 * avg(u_e) * (edge_e(+) - edge_e(-)) / h_e,
 *     upwinded; final states corrected by both other directions with dt/2,
 *     plus dt/2 * force (forces are not used in the transverse terms);
 *     conservative components (density, conservative tracers) take the
 *     transverse terms as flux differences minus the cell value times the
 *     velocity difference, -c [(a s)_+ - (a s)_-] + c q [a_+ - a_-];
 *   normal velocity: Burgers Riemann problem (0 in the transonic case lo<=0<=hi
 *     or when |lo+hi| < small_vel); advected states: upwind by the face
 *     velocity, average when |u| < small_vel;
 *   conv = -div(u_mac s) (+ s div(u_mac) for convective components; fluxes
 *     u_mac*s, not area weighted; src/convection/incflo_compute_advection_term.cpp:587,646-661).
 * Domain boundaries (non-periodic directions), per component and side, with
 * amrex::BCType values (ext_dir 3, foextrap 2, hoextrap 4, reflect_even 1,
 * reflect_odd -1):
 *   - the cell state comes with 2 ghost layers filled by the caller (an
 *     ext_dir ghost holds the boundary value on the face);
 *   - PPM next to an ext_dir / hoextrap face: the boundary-side interface
 *     value of the first cell is the ghost value, and the interface between
 *     the first two cells is the one-sided fourth-order value
 *     -0.2 s_g + 0.75 s_0 + 0.5 s_1 - 0.05 s_2 bounded by s_0 and s_1 (before
 *     the CW84 fix);
 *   - every state on a boundary face (edge, intermediate, final) is set from
 *     the interior-side trace: ext_dir -> the ghost value, foextrap /
 *     hoextrap / reflect_even -> the interior trace, reflect_odd -> 0; in
 *     ExtrapVelToFaces the normal velocity on a foextrap (outflow) face may
 *     not point into the domain (clipped to 0);
 *   - given MAC velocities are used as they are on boundary faces.
 * Arrays: cells (nz,ny,nx) x fastest; ghosted cells (nz+4,ny+4,nx+4); the
 * faces of direction d have n_d + 1 entries along d (periodic: face n_d
 * repeats face 0). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define SMALL_VEL 1e-10
#define GH 2

enum { BC_RODD = -1, BC_INT = 0, BC_REVEN = 1, BC_FOE = 2, BC_EXT = 3, BC_HOE = 4 };

typedef struct {
    int n[3];
    int per[3];
    double dx[3];
} gg_t;

static gg_t gg(const orc_geom* g) {
    gg_t p;
    for (int d = 0; d < 3; ++d) {
        p.n[d] = g->n[d];
        p.per[d] = g->periodic[d];
        p.dx[d] = g->dx[d];
    }
    return p;
}

static inline int wr(int i, int n) { return ((i % n) + n) % n; }

/* interior cell index (periodic directions wrap) */
static inline long cat(const gg_t* p, int i, int j, int k) {
    if (p->per[0]) i = wr(i, p->n[0]);
    if (p->per[1]) j = wr(j, p->n[1]);
    if (p->per[2]) k = wr(k, p->n[2]);
    return (long)i + (long)p->n[0] * ((long)j + (long)p->n[1] * k);
}

/* ghosted cell index, i in [-GH, n+GH) */
static inline long gat(const gg_t* p, int i, int j, int k) {
    return (long)(i + GH) + (long)(p->n[0] + 2 * GH) * ((long)(j + GH) + (long)(p->n[1] + 2 * GH) * (k + GH));
}

/* face index of direction d, face (i,j,k) with the index along d in [0, n_d] */
static inline long fat(const gg_t* p, int d, int i, int j, int k) {
    return (long)i + (long)(p->n[0] + (d == 0)) * ((long)j + (long)(p->n[1] + (d == 1)) * k);
}

static inline long nface(const gg_t* p, int d) {
    return (long)(p->n[0] + (d == 0)) * (p->n[1] + (d == 1)) * (p->n[2] + (d == 2));
}

static inline long ncell(const gg_t* p) { return (long)p->n[0] * p->n[1] * p->n[2]; }

static double vl_slope(double a, double b, double c) {
    const double d = 0.5 * (c - a);
    if ((c - b) * (b - a) <= 0.0) return 0.0;
    const double m = fmin(fabs(d), fmin(2.0 * fabs(b - a), 2.0 * fabs(c - b)));
    return d >= 0.0 ? m : -m;
}

static double one_sided_edge(double sg, double s0, double s1, double s2) {
    double e = -0.2 * sg + 0.75 * s0 + 0.5 * s1 - 0.05 * s2;
    e = fmax(e, fmin(s0, s1));
    e = fmin(e, fmax(s0, s1));
    return e;
}

/* PPM interface values of cell c (index id along d, n cells) from the ghosted
 * column s[-2..n+1] around it; bl / bh: BC types of the low / high side */
static void ppm_parabola(const double* col, int id, int n, int bl, int bh, double* sL,
                         double* sR) {
    const double sm2 = col[id - 2], sm1 = col[id - 1], s0 = col[id], sp1 = col[id + 1],
                 sp2 = col[id + 2];
    double r = 0.5 * (s0 + sp1) - (vl_slope(s0, sp1, sp2) - vl_slope(sm1, s0, sp1)) / 6.0;
    double l = 0.5 * (sm1 + s0) - (vl_slope(sm1, s0, sp1) - vl_slope(sm2, sm1, s0)) / 6.0;
    r = fmax(fmin(s0, sp1), fmin(r, fmax(s0, sp1)));
    l = fmax(fmin(sm1, s0), fmin(l, fmax(sm1, s0)));
    if (bl == BC_EXT || bl == BC_HOE) {
        const double e = one_sided_edge(col[-1], col[0], col[1], col[2]);
        if (id == 0) { l = col[-1]; r = e; }
        else if (id == 1) { l = e; }
    }
    if (bh == BC_EXT || bh == BC_HOE) {
        const double e = one_sided_edge(col[n], col[n - 1], col[n - 2], col[n - 3]);
        if (id == n - 1) { r = col[n]; l = e; }
        else if (id == n - 2) { r = e; }
    }
    if ((r - s0) * (s0 - l) <= 0.0) {
        l = s0;
        r = s0;
    } else if (fabs(r - s0) >= 2.0 * fabs(l - s0)) {
        r = 3.0 * s0 - 2.0 * l;
    } else if (fabs(l - s0) >= 2.0 * fabs(r - s0)) {
        l = 3.0 * s0 - 2.0 * r;
    }
    *sL = l;
    *sR = r;
}

static void ppm_trace_vals(double sL, double sR, double s0, double uL, double uR, double dt,
                           double h, double* Im, double* Ip) {
    const double s6 = 6.0 * s0 - 3.0 * (sL + sR);
    if (uR > SMALL_VEL) {
        const double sg = uR * dt / h;
        *Ip = sR - 0.5 * sg * ((sR - sL) - (1.0 - 2.0 * sg / 3.0) * s6);
    } else {
        *Ip = s0;
    }
    if (uL < -SMALL_VEL) {
        const double sg = -uL * dt / h;
        *Im = sL + 0.5 * sg * ((sR - sL) + (1.0 - 2.0 * sg / 3.0) * s6);
    } else {
        *Im = s0;
    }
}

/* traces Im/Ip along d of every interior cell of the ghosted component sg;
 * velocities: ucell (ghosted, the cell's own) or uface (faces of d) */
static void ppm_traces(const gg_t* p, const double* sg, int d, int bl, int bh,
                       const double* ucell_g, const double* uface, double dt, double* Im,
                       double* Ip) {
    const int n = p->n[d];
    double colbuf[4096 + 8];
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                const int id = d == 0 ? i : (d == 1 ? j : k);
                /* the column along d, ghosts included, at offset GH */
                double* col = colbuf + GH;
                for (int q = -GH; q < n + GH; ++q) {
                    int c[3] = {i, j, k};
                    c[d] = q;
                    col[q] = sg[gat(p, c[0], c[1], c[2])];
                }
                double sL, sR;
                ppm_parabola(col, id, n, p->per[d] ? BC_INT : bl, p->per[d] ? BC_INT : bh, &sL, &sR);
                double uL, uR;
                if (ucell_g) {
                    uL = uR = ucell_g[gat(p, i, j, k)];
                } else {
                    int c[3] = {i, j, k};
                    uL = uface[fat(p, d, c[0], c[1], c[2])];
                    c[d] += 1;
                    uR = uface[fat(p, d, c[0], c[1], c[2])];
                }
                const long ci = cat(p, i, j, k);
                ppm_trace_vals(sL, sR, col[id], uL, uR, dt, p->dx[d], &Im[ci], &Ip[ci]);
            }
}

static inline double upwind(double lo, double hi, double a) {
    if (a > SMALL_VEL) return lo;
    if (a < -SMALL_VEL) return hi;
    return 0.5 * (lo + hi);
}

static inline double burgers(double lo, double hi) {
    const double st = (lo + hi >= 0.0) ? lo : hi;
    const int zero = (lo <= 0.0 && hi >= 0.0) || fabs(lo + hi) < SMALL_VEL;
    return zero ? 0.0 : st;
}

/* state on a boundary face from the interior-side value v and the ghost
 * value gv; clip_in: the normal velocity on an outflow face may not point
 * into the domain (low side: <= 0, high side: >= 0) */
static inline double bc_face(int bc, int side, double v, double gv, int clip_in) {
    switch (bc) {
        case BC_EXT: return gv;
        case BC_RODD: return 0.0;
        case BC_FOE:
            if (clip_in) return side ? fmax(v, 0.0) : fmin(v, 0.0);
            return v;
        default: return v;
    }
}

/* face (i,j,k) of direction d: is it a boundary face (non-periodic d, index 0 or n)?
 * side: 0 low, 1 high */
static inline int bface(const gg_t* p, int d, int id, int* side) {
    if (p->per[d]) return 0;
    if (id == 0) { *side = 0; return 1; }
    if (id == p->n[d]) { *side = 1; return 1; }
    return 0;
}

/* ghost cell just outside boundary face (i,j,k) of direction d */
static inline long ghost_of(const gg_t* p, int d, int side, int i, int j, int k) {
    int c[3] = {i, j, k};
    c[d] = side ? p->n[d] : -1;
    return gat(p, c[0], c[1], c[2]);
}

static double* dalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

/* Godunov face states of ncomp components.
 * extrap = 1: ExtrapVelToFaces (ncomp 3, q = velocity); out[d] = normal
 *   velocity on d-faces (nface(d) doubles each).
 * extrap = 0: edge states with the given MAC velocities umac[d] (faces);
 *   out + off(d, n) receives component n on d-faces. */
static void gd_general(const gg_t* p, int extrap, int ncomp, const double* qg, const int* bc,
                       const int* iconserv, const double* force, const double* const umac[3],
                       double dt, double* const out[3], int out_per_comp) {
    const long N = ncell(p);
    const long NG = (long)(p->n[0] + 2 * GH) * (p->n[1] + 2 * GH) * (p->n[2] + 2 * GH);
    double *Im[3][3], *Ip[3][3], *ad[3], *E[3][3], *T[3][3][3];
    memset(T, 0, sizeof(T));
    const int nc = ncomp;  /* <= 3 per call */
    for (int d = 0; d < 3; ++d) {
        ad[d] = dalloc(nface(p, d));
        for (int n = 0; n < nc; ++n) {
            Im[d][n] = dalloc(N);
            Ip[d][n] = dalloc(N);
            E[d][n] = dalloc(nface(p, d));
            for (int e = 0; e < 3; ++e)
                if (e != d) T[d][e][n] = dalloc(nface(p, d));
        }
    }
    /* 1. traces */
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < nc; ++n)
            ppm_traces(p, qg + (long)n * NG, d, bc[6 * n + d], bc[6 * n + 3 + d],
                       extrap ? qg + (long)d * NG : NULL, extrap ? NULL : umac[d], dt, Im[d][n],
                       Ip[d][n]);
    /* 2. advective face velocities */
    for (int d = 0; d < 3; ++d) {
        const int fd[3] = {p->n[0] + (d == 0), p->n[1] + (d == 1), p->n[2] + (d == 2)};
        for (int k = 0; k < fd[2]; ++k)
            for (int j = 0; j < fd[1]; ++j)
                for (int i = 0; i < fd[0]; ++i) {
                    const long f = fat(p, d, i, j, k);
                    if (!extrap) { ad[d][f] = umac[d][f]; continue; }
                    const int id = d == 0 ? i : (d == 1 ? j : k);
                    int side;
                    const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
                    if (bface(p, d, id, &side)) {
                        const double v = side ? Ip[d][d][cat(p, li, lj, lk)] : Im[d][d][cat(p, i, j, k)];
                        const double b = bc_face(bc[6 * d + 3 * side + d], side, v,
                                                 qg[(long)d * NG + ghost_of(p, d, side, i, j, k)], 1);
                        ad[d][f] = burgers(b, b);
                    } else {
                        ad[d][f] = burgers(Ip[d][d][cat(p, li, lj, lk)], Im[d][d][cat(p, i, j, k)]);
                    }
                }
    }
    /* 3. upwinded edge states */
    for (int d = 0; d < 3; ++d) {
        const int fd[3] = {p->n[0] + (d == 0), p->n[1] + (d == 1), p->n[2] + (d == 2)};
        for (int n = 0; n < nc; ++n)
            for (int k = 0; k < fd[2]; ++k)
                for (int j = 0; j < fd[1]; ++j)
                    for (int i = 0; i < fd[0]; ++i) {
                        const long f = fat(p, d, i, j, k);
                        const int id = d == 0 ? i : (d == 1 ? j : k);
                        const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
                        int side;
                        if (bface(p, d, id, &side)) {
                            const double v = side ? Ip[d][n][cat(p, li, lj, lk)] : Im[d][n][cat(p, i, j, k)];
                            E[d][n][f] = bc_face(bc[6 * n + 3 * side + d], side, v,
                                                 qg[(long)n * NG + ghost_of(p, d, side, i, j, k)], 0);
                        } else {
                            E[d][n][f] = upwind(Ip[d][n][cat(p, li, lj, lk)], Im[d][n][cat(p, i, j, k)],
                                                ad[d][f]);
                        }
                    }
    }
    /* 4. intermediate transverse states T[d][e][n] on d-faces, corrected along e */
    for (int n = 0; n < nc; ++n) {
        const int cons = iconserv ? iconserv[n] : 0;
        const double* qn = qg + (long)n * NG;
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) {
                if (e == d) continue;
                const double c = dt / (3.0 * p->dx[e]);
                const int fd[3] = {p->n[0] + (d == 0), p->n[1] + (d == 1), p->n[2] + (d == 2)};
                for (int k = 0; k < fd[2]; ++k)
                    for (int j = 0; j < fd[1]; ++j)
                        for (int i = 0; i < fd[0]; ++i) {
                            const long f = fat(p, d, i, j, k);
                            const int id = d == 0 ? i : (d == 1 ? j : k);
                            int side = -1;
                            const int bnd = bface(p, d, id, &side);
                            double val[2];
                            /* s = 0: low-side cell, s = 1: high-side cell */
                            for (int s = 0; s < 2; ++s) {
                                if (bnd && s == side) continue;  /* ghost cell: not traced */
                                int cc[3] = {i, j, k};
                                if (s == 0) cc[d] -= 1;
                                if (p->per[d]) cc[d] = wr(cc[d], p->n[d]);
                                const long ci = cat(p, cc[0], cc[1], cc[2]);
                                int up[3] = {cc[0], cc[1], cc[2]};
                                up[e] += 1;
                                const long fl = fat(p, e, cc[0], cc[1], cc[2]);
                                const long fh = fat(p, e, up[0], up[1], up[2]);
                                const double* Ee = E[e][n];
                                const double* ae = ad[e];
                                const double tr = s == 0 ? Ip[d][n][ci] : Im[d][n][ci];
                                if (cons)
                                    val[s] = tr - c * (ae[fh] * Ee[fh] - ae[fl] * Ee[fl]) +
                                             c * qn[gat(p, cc[0], cc[1], cc[2])] * (ae[fh] - ae[fl]);
                                else
                                    val[s] = tr - c * 0.5 * (ae[fh] + ae[fl]) * (Ee[fh] - Ee[fl]);
                            }
                            if (bnd) {
                                const double v = val[1 - side];
                                T[d][e][n][f] = bc_face(bc[6 * n + 3 * side + d], side, v,
                                                        qn[ghost_of(p, d, side, i, j, k)], 0);
                            } else {
                                T[d][e][n][f] = upwind(val[0], val[1], ad[d][f]);
                            }
                        }
            }
    }
    /* 5. final states on d-faces */
    for (int d = 0; d < 3; ++d) {
        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
        const double c1 = dt / (2.0 * p->dx[e1]), c2 = dt / (2.0 * p->dx[e2]);
        const int fd[3] = {p->n[0] + (d == 0), p->n[1] + (d == 1), p->n[2] + (d == 2)};
        for (int n = 0; n < nc; ++n) {
            if (extrap && n != d) continue;
            const int cons = iconserv ? iconserv[n] : 0;
            const double* qn = qg + (long)n * NG;
            const double* fn = force ? force + (long)n * N : NULL;
            double* o = extrap ? out[d] : out[d] + (long)n * (out_per_comp ? nface(p, d) : 0);
            const double* t1 = T[e1][e2][n];
            const double* t2 = T[e2][e1][n];
            for (int k = 0; k < fd[2]; ++k)
                for (int j = 0; j < fd[1]; ++j)
                    for (int i = 0; i < fd[0]; ++i) {
                        const long f = fat(p, d, i, j, k);
                        const int id = d == 0 ? i : (d == 1 ? j : k);
                        int side = -1;
                        const int bnd = bface(p, d, id, &side);
                        double val[2];
                        for (int s = 0; s < 2; ++s) {
                            if (bnd && s == side) continue;
                            int cc[3] = {i, j, k};
                            if (s == 0) cc[d] -= 1;
                            if (p->per[d]) cc[d] = wr(cc[d], p->n[d]);
                            const long ci = cat(p, cc[0], cc[1], cc[2]);
                            int u1[3] = {cc[0], cc[1], cc[2]}, u2[3] = {cc[0], cc[1], cc[2]};
                            u1[e1] += 1;
                            u2[e2] += 1;
                            const long l1 = fat(p, e1, cc[0], cc[1], cc[2]), h1 = fat(p, e1, u1[0], u1[1], u1[2]);
                            const long l2 = fat(p, e2, cc[0], cc[1], cc[2]), h2 = fat(p, e2, u2[0], u2[1], u2[2]);
                            const double* a1 = ad[e1];
                            const double* a2 = ad[e2];
                            double v = s == 0 ? Ip[d][n][ci] : Im[d][n][ci];
                            if (cons) {
                                const double qc = qn[gat(p, cc[0], cc[1], cc[2])];
                                v = v - c1 * (a1[h1] * t1[h1] - a1[l1] * t1[l1]) + c1 * qc * (a1[h1] - a1[l1]) -
                                    c2 * (a2[h2] * t2[h2] - a2[l2] * t2[l2]) + c2 * qc * (a2[h2] - a2[l2]);
                            } else {
                                v = v - c1 * 0.5 * (a1[h1] + a1[l1]) * (t1[h1] - t1[l1]) -
                                    c2 * 0.5 * (a2[h2] + a2[l2]) * (t2[h2] - t2[l2]);
                            }
                            if (fn) v += 0.5 * dt * fn[ci];
                            val[s] = v;
                        }
                        if (bnd) {
                            const double b = bc_face(bc[6 * n + 3 * side + d], side, val[1 - side],
                                                     qn[ghost_of(p, d, side, i, j, k)], extrap);
                            o[f] = extrap ? burgers(b, b) : b;
                        } else {
                            o[f] = extrap ? burgers(val[0], val[1]) : upwind(val[0], val[1], ad[d][f]);
                        }
                    }
        }
    }
    for (int d = 0; d < 3; ++d) {
        free(ad[d]);
        for (int n = 0; n < nc; ++n) {
            free(Im[d][n]);
            free(Ip[d][n]);
            free(E[d][n]);
            for (int e = 0; e < 3; ++e) free(T[d][e][n]);
        }
    }
}

/* ---- general entry points (ghosted cell states, n+1 faces) ------------- */

/* ExtrapVelToFaces: vel_g (3 ghosted comps), force (3 interior comps or
 * NULL), bc[3][6]; umac/vmac/wmac receive the normal velocities */
void orc_gd_extrap_vel(const orc_geom* g, const double* vel_g, const int* bc, const double* force,
                       double dt, double* umac, double* vmac, double* wmac) {
    gg_t p = gg(g);
    double* out[3] = {umac, vmac, wmac};
    gd_general(&p, 1, 3, vel_g, bc, NULL, force, NULL, dt, out, 0);
}

/* edge states of ncomp components: states[d] + n * nface(d) */
void orc_gd_edge_states(const orc_geom* g, int ncomp, const double* s_g, const int* bc,
                        const int* iconserv, const double* force, const double* umac,
                        const double* vmac, const double* wmac, double dt, double* sx, double* sy,
                        double* sz) {
    gg_t p = gg(g);
    const double* um[3] = {umac, vmac, wmac};
    const long NG = (long)(p.n[0] + 2 * GH) * (p.n[1] + 2 * GH) * (p.n[2] + 2 * GH);
    const long N = ncell(&p);
    for (int n0 = 0; n0 < ncomp; n0 += 3) {
        const int nc = ncomp - n0 < 3 ? ncomp - n0 : 3;
        double* out[3] = {sx + (long)n0 * nface(&p, 0), sy + (long)n0 * nface(&p, 1),
                          sz + (long)n0 * nface(&p, 2)};
        gd_general(&p, 0, nc, s_g + (long)n0 * NG, bc + 6 * n0, iconserv ? iconserv + n0 : NULL,
                   force ? force + (long)n0 * N : NULL, um, dt, out, 1);
    }
}

/* conv[n] = -div(u_mac s_face[n]) (+ s_n div(u_mac) when iconserv[n] == 0);
 * s: interior cells (ncomp), states as from orc_gd_edge_states */
void orc_gd_conv(const orc_geom* g, int ncomp, const double* s, const int* iconserv,
                 const double* umac, const double* vmac, const double* wmac, const double* sx,
                 const double* sy, const double* sz, double* conv) {
    gg_t p = gg(g);
    const long N = ncell(&p);
    const double* a[3] = {umac, vmac, wmac};
    const double* st[3] = {sx, sy, sz};
    for (int k = 0; k < p.n[2]; ++k)
        for (int j = 0; j < p.n[1]; ++j)
            for (int i = 0; i < p.n[0]; ++i) {
                const long c = cat(&p, i, j, k);
                double divu = 0.0;
                long fl[3], fh[3];
                for (int d = 0; d < 3; ++d) {
                    fl[d] = fat(&p, d, i, j, k);
                    fh[d] = fat(&p, d, i + (d == 0), j + (d == 1), k + (d == 2));
                    divu += (a[d][fh[d]] - a[d][fl[d]]) / p.dx[d];
                }
                for (int n = 0; n < ncomp; ++n) {
                    double div = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double* sd = st[d] + (long)n * nface(&p, d);
                        div += (a[d][fh[d]] * sd[fh[d]] - a[d][fl[d]] * sd[fl[d]]) / p.dx[d];
                    }
                    const int cons = iconserv ? iconserv[n] : 0;
                    conv[(long)n * N + c] = -div + (cons ? 0.0 : s[(long)n * N + c] * divu);
                }
            }
}

/* ---- periodic canonical API (face arrays with n entries per direction) -- */

/* ghosted copy of ncomp periodic components */
static double* ghosted_periodic(const gg_t* p, const double* const* s, int ncomp) {
    const long NG = (long)(p->n[0] + 2 * GH) * (p->n[1] + 2 * GH) * (p->n[2] + 2 * GH);
    double* o = dalloc(NG * ncomp);
    for (int n = 0; n < ncomp; ++n)
        for (int k = -GH; k < p->n[2] + GH; ++k)
            for (int j = -GH; j < p->n[1] + GH; ++j)
                for (int i = -GH; i < p->n[0] + GH; ++i)
                    o[(long)n * NG + gat(p, i, j, k)] =
                        s[n][(long)wr(i, p->n[0]) + (long)p->n[0] * (wr(j, p->n[1]) + (long)p->n[1] * wr(k, p->n[2]))];
    return o;
}

/* n-entry periodic faces <-> n+1-entry faces */
static double* faces_plus(const gg_t* p, int d, const double* f) {
    double* o = dalloc(nface(p, d));
    const int fd[3] = {p->n[0] + (d == 0), p->n[1] + (d == 1), p->n[2] + (d == 2)};
    for (int k = 0; k < fd[2]; ++k)
        for (int j = 0; j < fd[1]; ++j)
            for (int i = 0; i < fd[0]; ++i)
                o[fat(p, d, i, j, k)] = f[cat(p, i, j, k)];
    return o;
}

static void faces_minus(const gg_t* p, int d, const double* f, double* o) {
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) o[cat(p, i, j, k)] = f[fat(p, d, i, j, k)];
}

static gg_t periodic_of(const orc_geom* g) {
    gg_t p = gg(g);
    p.per[0] = p.per[1] = p.per[2] = 1;
    return p;
}

void orc_ppm_predict(const orc_geom* g, const double* s, int d, const double* ucell,
                     const double* uface, double dt, double* Im, double* Ip) {
    gg_t p = periodic_of(g);
    const double* sp[1] = {s};
    double* sg = ghosted_periodic(&p, sp, 1);
    double* ug = NULL;
    double* uf = NULL;
    if (ucell) {
        const double* up[1] = {ucell};
        ug = ghosted_periodic(&p, up, 1);
    } else {
        uf = faces_plus(&p, d, uface);
    }
    ppm_traces(&p, sg, d, BC_INT, BC_INT, ug, uf, dt, Im, Ip);
    free(sg);
    free(ug);
    free(uf);
}

void orc_godunov_extrap_vel(const orc_geom* g, const double* u, const double* v, const double* w,
                            const double* fu, const double* fv, const double* fw, double dt,
                            double* umac, double* vmac, double* wmac) {
    gg_t p = periodic_of(g);
    const long N = ncell(&p);
    const double* q[3] = {u, v, w};
    double* qg = ghosted_periodic(&p, q, 3);
    double* f = NULL;
    if (fu) {
        f = dalloc(3 * N);
        memcpy(f, fu, N * sizeof(double));
        memcpy(f + N, fv, N * sizeof(double));
        memcpy(f + 2 * N, fw, N * sizeof(double));
    }
    int bc[18] = {0};
    double* o[3];
    for (int d = 0; d < 3; ++d) o[d] = dalloc(nface(&p, d));
    gd_general(&p, 1, 3, qg, bc, NULL, f, NULL, dt, o, 0);
    double* dst[3] = {umac, vmac, wmac};
    for (int d = 0; d < 3; ++d) {
        faces_minus(&p, d, o[d], dst[d]);
        free(o[d]);
    }
    free(qg);
    free(f);
}

/* edge states of ncomp periodic components: states[(d * ncomp + n) * N] */
static void periodic_states(const orc_geom* g, const double* const* q, int ncomp,
                            const int* iconserv, const double* force, const double* const um_in[3],
                            double dt, double* states) {
    gg_t p = periodic_of(g);
    const long N = ncell(&p);
    double* qg = ghosted_periodic(&p, q, ncomp);
    double* um[3];
    double* st[3];
    for (int d = 0; d < 3; ++d) {
        um[d] = faces_plus(&p, d, um_in[d]);
        st[d] = dalloc(nface(&p, d) * ncomp);
    }
    int* bc = (int*)calloc((size_t)6 * ncomp, sizeof(int));
    orc_geom gp = *g;
    gp.periodic[0] = gp.periodic[1] = gp.periodic[2] = 1;
    orc_gd_edge_states(&gp, ncomp, qg, bc, iconserv, force, um[0], um[1], um[2], dt, st[0], st[1],
                       st[2]);
    for (int d = 0; d < 3; ++d) {
        for (int n = 0; n < ncomp; ++n)
            faces_minus(&p, d, st[d] + (long)n * nface(&p, d), states + (long)(d * ncomp + n) * N);
        free(um[d]);
        free(st[d]);
    }
    free(bc);
    free(qg);
}

void orc_godunov_edge_states(const orc_geom* g, const double* u, const double* v, const double* w,
                             const double* fu, const double* fv, const double* fw,
                             const double* umac, const double* vmac, const double* wmac,
                             double dt, double* states /* [3 faces][3 comps][N] */) {
    const long N = (long)g->n[0] * g->n[1] * g->n[2];
    const double* q[3] = {u, v, w};
    const double* a[3] = {umac, vmac, wmac};
    double* f = NULL;
    if (fu) {
        f = dalloc(3 * N);
        memcpy(f, fu, N * sizeof(double));
        memcpy(f + N, fv, N * sizeof(double));
        memcpy(f + 2 * N, fw, N * sizeof(double));
    }
    periodic_states(g, q, 3, NULL, f, a, dt, states);
    free(f);
}

void orc_godunov_scalar_states(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                               const double* force, const double* umac, const double* vmac,
                               const double* wmac, double dt, double* states) {
    const long N = (long)g->n[0] * g->n[1] * g->n[2];
    const double** q = (const double**)malloc(sizeof(double*) * (size_t)ncomp);
    for (int n = 0; n < ncomp; ++n) q[n] = s + (long)n * N;
    const double* a[3] = {umac, vmac, wmac};
    periodic_states(g, q, ncomp, iconserv, force, a, dt, states);
    free(q);
}

static void periodic_conv(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                          const double* const um_in[3], const double* states, double* conv) {
    gg_t p = periodic_of(g);
    const long N = ncell(&p);
    double* um[3];
    double* st[3];
    for (int d = 0; d < 3; ++d) {
        um[d] = faces_plus(&p, d, um_in[d]);
        st[d] = dalloc(nface(&p, d) * ncomp);
        for (int n = 0; n < ncomp; ++n) {
            double* t = faces_plus(&p, d, states + (long)(d * ncomp + n) * N);
            memcpy(st[d] + (long)n * nface(&p, d), t, nface(&p, d) * sizeof(double));
            free(t);
        }
    }
    orc_geom gp = *g;
    gp.periodic[0] = gp.periodic[1] = gp.periodic[2] = 1;
    orc_gd_conv(&gp, ncomp, s, iconserv, um[0], um[1], um[2], st[0], st[1], st[2], conv);
    for (int d = 0; d < 3; ++d) {
        free(um[d]);
        free(st[d]);
    }
}

void orc_scalar_advection(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                          const double* umac, const double* vmac, const double* wmac,
                          const double* states, double* conv) {
    const double* a[3] = {umac, vmac, wmac};
    periodic_conv(g, s, ncomp, iconserv, a, states, conv);
}

/* conv[n] = -div(u_mac s_n) + s_n div(u_mac), s_n cell values, from face states */
void orc_convective_term(const orc_geom* g, const double* u, const double* v, const double* w,
                         const double* umac, const double* vmac, const double* wmac,
                         const double* states, double* conv /* [3][N] */) {
    const long N = (long)g->n[0] * g->n[1] * g->n[2];
    double* s = dalloc(3 * N);
    memcpy(s, u, N * sizeof(double));
    memcpy(s + N, v, N * sizeof(double));
    memcpy(s + 2 * N, w, N * sizeof(double));
    const double* a[3] = {umac, vmac, wmac};
    periodic_conv(g, s, 3, NULL, a, states, conv);
    free(s);
}

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
 * this restates the published method (single level, periodic, no EB):
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
 *   normal velocity: Burgers Riemann problem (0 in the transonic case lo<=0<=hi
 *     or when |lo+hi| < small_vel); advected states: upwind by the face
 *     velocity, average when |u| < small_vel.
 *   convective term: conv = -div(u_mac s) + s div(u_mac)  (fluxes u_mac*s on
 *     faces, not area weighted; src/convection/incflo_compute_advection_term.cpp:587,646-661).
 * Periodic canonical arrays: cells (nz,ny,nx) x fastest; d-faces have n
 * entries per direction, face i being the low face of cell i. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define SMALL_VEL 1e-10

typedef struct {
    int n[3];
    double dx[3];
} pg_t;

static pg_t pg(const orc_geom* g) {
    pg_t p;
    for (int d = 0; d < 3; ++d) {
        p.n[d] = g->n[d];
        p.dx[d] = g->dx[d];
    }
    return p;
}

static inline int wr(int i, int n) { return ((i % n) + n) % n; }

static inline long at(const pg_t* p, int i, int j, int k) {
    return (long)wr(i, p->n[0]) + (long)p->n[0] * ((long)wr(j, p->n[1]) + (long)p->n[1] * wr(k, p->n[2]));
}

/* index of the cell shifted by s along d */
static inline long sh(const pg_t* p, int i, int j, int k, int d, int s) {
    return at(p, i + (d == 0) * s, j + (d == 1) * s, k + (d == 2) * s);
}

static double vl_slope(double a, double b, double c) {
    const double d = 0.5 * (c - a);
    if ((c - b) * (b - a) <= 0.0) return 0.0;
    const double m = fmin(fabs(d), fmin(2.0 * fabs(b - a), 2.0 * fabs(c - b)));
    return d >= 0.0 ? m : -m;
}

static void ppm_parabola(double sm2, double sm1, double s0, double sp1, double sp2, double* sL,
                         double* sR) {
    double r = 0.5 * (s0 + sp1) - (vl_slope(s0, sp1, sp2) - vl_slope(sm1, s0, sp1)) / 6.0;
    double l = 0.5 * (sm1 + s0) - (vl_slope(sm1, s0, sp1) - vl_slope(sm2, sm1, s0)) / 6.0;
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
    *sL = l;
    *sR = r;
}

/* traced right (Ip) and left (Im) face values of cell (i,j,k) along d;
 * uR / uL: the velocities deciding each side (cell velocity for both, or the
 * right / left face velocities) */
static void ppm_trace(const pg_t* p, const double* s, int d, int i, int j, int k, double uL,
                      double uR, double dt, double* Im, double* Ip) {
    const double sm2 = s[sh(p, i, j, k, d, -2)], sm1 = s[sh(p, i, j, k, d, -1)];
    const double s0 = s[at(p, i, j, k)];
    const double sp1 = s[sh(p, i, j, k, d, 1)], sp2 = s[sh(p, i, j, k, d, 2)];
    double sL, sR;
    ppm_parabola(sm2, sm1, s0, sp1, sp2, &sL, &sR);
    const double s6 = 6.0 * s0 - 3.0 * (sL + sR);
    if (uR > SMALL_VEL) {
        const double sg = uR * dt / p->dx[d];
        *Ip = sR - 0.5 * sg * ((sR - sL) - (1.0 - 2.0 * sg / 3.0) * s6);
    } else {
        *Ip = s0;
    }
    if (uL < -SMALL_VEL) {
        const double sg = -uL * dt / p->dx[d];
        *Im = sL + 0.5 * sg * ((sR - sL) + (1.0 - 2.0 * sg / 3.0) * s6);
    } else {
        *Im = s0;
    }
}

void orc_ppm_predict(const orc_geom* g, const double* s, int d, const double* ucell,
                     const double* uface, double dt, double* Im, double* Ip) {
    pg_t p = pg(g);
    for (int k = 0; k < p.n[2]; ++k)
        for (int j = 0; j < p.n[1]; ++j)
            for (int i = 0; i < p.n[0]; ++i) {
                const long c = at(&p, i, j, k);
                double uL, uR;
                if (ucell) {
                    uL = uR = ucell[c];
                } else {
                    uL = uface[c];
                    uR = uface[sh(&p, i, j, k, d, 1)];
                }
                ppm_trace(&p, s, d, i, j, k, uL, uR, dt, &Im[c], &Ip[c]);
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

/* edge[f] at d-faces: upwind(Ip(cell f-1), Im(cell f); a[f]) */
static void edges(const pg_t* p, int d, const double* Im, const double* Ip, const double* a,
                  double* e) {
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                const long f = at(p, i, j, k);
                e[f] = upwind(Ip[sh(p, i, j, k, d, -1)], Im[f], a[f]);
            }
}

/* intermediate states at d-faces of component s corrected along e (dt/3),
 * upwinded by ad (d-face advective velocity): needs Im/Ip along d, the
 * e-edges of the same component and the e-face advective velocity ae */
static void trans1(const pg_t* p, int d, int e, const double* Imd, const double* Ipd,
                   const double* edge_e, const double* ae, const double* ad, double dt,
                   double* out) {
    const double c = dt / (3.0 * p->dx[e]);
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                const long f = at(p, i, j, k);
                const long lc = sh(p, i, j, k, d, -1);  /* cell on the low side */
                const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
                const long lcp = sh(p, li, lj, lk, e, 1), hcp = sh(p, i, j, k, e, 1);
                const double lo = Ipd[lc] - c * 0.5 * (ae[lcp] + ae[lc]) * (edge_e[lcp] - edge_e[lc]);
                const double hi = Imd[f] - c * 0.5 * (ae[hcp] + ae[f]) * (edge_e[hcp] - edge_e[f]);
                out[f] = upwind(lo, hi, ad[f]);
            }
}

/* final low/high states at d-faces with the two transverse corrections (dt/2)
 * and dt/2 * force; t1 = (d, e1) intermediate e1-face states of the component
 * corrected along e2 ... i.e. t_e1 lives on e1-faces, t_e2 on e2-faces */
static void final_states(const pg_t* p, int d, int e1, int e2, const double* Imd,
                         const double* Ipd, const double* t_e1, const double* a_e1,
                         const double* t_e2, const double* a_e2, const double* force, double dt,
                         long f, int i, int j, int k, double* lo, double* hi) {
    const long lc = sh(p, i, j, k, d, -1);
    const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
    const double c1 = dt / (2.0 * p->dx[e1]), c2 = dt / (2.0 * p->dx[e2]);
    const long l1 = sh(p, li, lj, lk, e1, 1), l2 = sh(p, li, lj, lk, e2, 1);
    const long h1 = sh(p, i, j, k, e1, 1), h2 = sh(p, i, j, k, e2, 1);
    double l = Ipd[lc] - c1 * 0.5 * (a_e1[l1] + a_e1[lc]) * (t_e1[l1] - t_e1[lc]) -
               c2 * 0.5 * (a_e2[l2] + a_e2[lc]) * (t_e2[l2] - t_e2[lc]);
    double h = Imd[f] - c1 * 0.5 * (a_e1[h1] + a_e1[f]) * (t_e1[h1] - t_e1[f]) -
               c2 * 0.5 * (a_e2[h2] + a_e2[f]) * (t_e2[h2] - t_e2[f]);
    if (force) {
        l += 0.5 * dt * force[lc];
        h += 0.5 * dt * force[f];
    }
    *lo = l;
    *hi = h;
}

static double* dalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

/* Godunov edge states of the 3 components q[n] on all faces:
 * adv = NULL: ExtrapVelToFaces mode -- the predictor uses cell velocities,
 *   the advective velocities come from the normal components (Burgers), and
 *   only the normal component is finalised (out[d] = normal velocity on d-faces);
 * adv given (face velocities u_mac): edge-state mode -- every component q[n]
 *   on every d-face, upwinded by adv[d] (out[3*d + n]). */
static void godunov(const orc_geom* g, const double* const q[3], const double* const force[3],
                    const double* const adv_in[3], double dt, double** out) {
    pg_t p = pg(g);
    const long N = (long)p.n[0] * p.n[1] * p.n[2];
    double *Im[3][3], *Ip[3][3], *ad[3];
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) {
            Im[d][n] = dalloc(N);
            Ip[d][n] = dalloc(N);
            if (adv_in)
                orc_ppm_predict(g, q[n], d, NULL, adv_in[d], dt, Im[d][n], Ip[d][n]);
            else
                orc_ppm_predict(g, q[n], d, q[d], NULL, dt, Im[d][n], Ip[d][n]);
        }
    for (int d = 0; d < 3; ++d) {
        ad[d] = dalloc(N);
        if (adv_in) {
            memcpy(ad[d], adv_in[d], sizeof(double) * N);
        } else {
            for (int k = 0; k < p.n[2]; ++k)
                for (int j = 0; j < p.n[1]; ++j)
                    for (int i = 0; i < p.n[0]; ++i) {
                        const long f = at(&p, i, j, k);
                        ad[d][f] = burgers(Ip[d][d][sh(&p, i, j, k, d, -1)], Im[d][d][f]);
                    }
        }
    }
    /* upwinded edge states of every component on every face direction */
    double* E[3][3];
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) {
            E[d][n] = dalloc(N);
            edges(&p, d, Im[d][n], Ip[d][n], ad[d], E[d][n]);
        }
    /* intermediate transverse states T[d][e][n]: component n on d-faces,
     * corrected along e != d */
    double* T[3][3][3];
    memset(T, 0, sizeof(T));
    for (int n = 0; n < 3; ++n)
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) {
                if (e == d) continue;
                T[d][e][n] = dalloc(N);
                trans1(&p, d, e, Im[d][n], Ip[d][n], E[e][n], ad[e], ad[d], dt, T[d][e][n]);
            }
    /* final states: component n on d-faces uses T[e1][e2][n] (on e1-faces) and T[e2][e1][n] */
    for (int d = 0; d < 3; ++d) {
        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
        for (int n = 0; n < 3; ++n) {
            if (!adv_in && n != d) continue;
            double* o = adv_in ? out[3 * d + n] : out[d];
            for (int k = 0; k < p.n[2]; ++k)
                for (int j = 0; j < p.n[1]; ++j)
                    for (int i = 0; i < p.n[0]; ++i) {
                        const long f = at(&p, i, j, k);
                        double lo, hi;
                        final_states(&p, d, e1, e2, Im[d][n], Ip[d][n], T[e1][e2][n], ad[e1],
                                     T[e2][e1][n], ad[e2], force ? force[n] : NULL, dt, f, i, j,
                                     k, &lo, &hi);
                        o[f] = adv_in ? upwind(lo, hi, ad[d][f]) : burgers(lo, hi);
                    }
        }
    }
    for (int d = 0; d < 3; ++d) {
        free(ad[d]);
        for (int n = 0; n < 3; ++n) {
            free(Im[d][n]);
            free(Ip[d][n]);
            free(E[d][n]);
            for (int e = 0; e < 3; ++e) free(T[d][e][n]);
        }
    }
}

void orc_godunov_extrap_vel(const orc_geom* g, const double* u, const double* v, const double* w,
                            const double* fu, const double* fv, const double* fw, double dt,
                            double* umac, double* vmac, double* wmac) {
    const double* q[3] = {u, v, w};
    const double* f[3] = {fu, fv, fw};
    double* out[3] = {umac, vmac, wmac};
    godunov(g, q, fu ? f : NULL, NULL, dt, out);
}

void orc_godunov_edge_states(const orc_geom* g, const double* u, const double* v, const double* w,
                             const double* fu, const double* fv, const double* fw,
                             const double* umac, const double* vmac, const double* wmac,
                             double dt, double* states /* [3 faces][3 comps][N] */) {
    const double* q[3] = {u, v, w};
    const double* f[3] = {fu, fv, fw};
    const double* a[3] = {umac, vmac, wmac};
    const long N = (long)g->n[0] * g->n[1] * g->n[2];
    double* out[9];
    for (int q9 = 0; q9 < 9; ++q9) out[q9] = states + (long)q9 * N;
    godunov(g, q, fu ? f : NULL, a, dt, out);
}

/* conv[n] = -div(u_mac s_n) + s_n div(u_mac), s_n cell values, from face states */
void orc_convective_term(const orc_geom* g, const double* u, const double* v, const double* w,
                         const double* umac, const double* vmac, const double* wmac,
                         const double* states, double* conv /* [3][N] */) {
    pg_t p = pg(g);
    const long N = (long)p.n[0] * p.n[1] * p.n[2];
    const double* s[3] = {u, v, w};
    const double* a[3] = {umac, vmac, wmac};
    for (int k = 0; k < p.n[2]; ++k)
        for (int j = 0; j < p.n[1]; ++j)
            for (int i = 0; i < p.n[0]; ++i) {
                const long c = at(&p, i, j, k);
                double divu = 0.0;
                for (int d = 0; d < 3; ++d)
                    divu += (a[d][sh(&p, i, j, k, d, 1)] - a[d][c]) / p.dx[d];
                for (int n = 0; n < 3; ++n) {
                    double div = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double* st = states + (long)(3 * d + n) * N;
                        const long cp = sh(&p, i, j, k, d, 1);
                        div += (a[d][cp] * st[cp] - a[d][c] * st[c]) / p.dx[d];
                    }
                    conv[(long)n * N + c] = -div + s[n][c] * divu;
                }
            }
}

/* ---- scalars: density and tracers (edge-state mode with u_mac) -----------
 * incflo advects density and tracers with the MAC velocities
 * (src/convection/incflo_compute_advection_term.cpp:446-561): edge states by
 * the same Godunov steps, then fluxes u_mac s_face and conv = -div(F)
 * (ComputeDivergence, mult = -1, :666-742), plus s div(u_mac) for the
 * components advected in convective form (ComputeConvectiveTerm, :744-760).
 * A component with iconserv = 1 (density always, tracers by
 * trac_is_conservative, :28-35) takes its transverse corrections in
 * conservative form: the difference of the transverse fluxes minus the cell
 * value times the difference of the transverse velocities,
 *     - c [(a s)_+ - (a s)_-] + c q [a_+ - a_-],
 * which vanishes for a constant state like the convective form
 * - c avg(a) (s_+ - s_-) used otherwise. */
static void trans1_cons(const pg_t* p, int d, int e, const double* q, const double* Imd,
                        const double* Ipd, const double* edge_e, const double* ae, const double* ad,
                        double dt, double* out) {
    const double c = dt / (3.0 * p->dx[e]);
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                const long f = at(p, i, j, k);
                const long lc = sh(p, i, j, k, d, -1);
                const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
                const long lcp = sh(p, li, lj, lk, e, 1), hcp = sh(p, i, j, k, e, 1);
                const double lo = Ipd[lc] - c * (ae[lcp] * edge_e[lcp] - ae[lc] * edge_e[lc]) +
                                  c * q[lc] * (ae[lcp] - ae[lc]);
                const double hi = Imd[f] - c * (ae[hcp] * edge_e[hcp] - ae[f] * edge_e[f]) +
                                  c * q[f] * (ae[hcp] - ae[f]);
                out[f] = upwind(lo, hi, ad[f]);
            }
}

static void final_states_cons(const pg_t* p, int d, int e1, int e2, const double* q,
                              const double* Imd, const double* Ipd, const double* t_e1,
                              const double* a_e1, const double* t_e2, const double* a_e2,
                              const double* force, double dt, long f, int i, int j, int k,
                              double* lo, double* hi) {
    const long lc = sh(p, i, j, k, d, -1);
    const int li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
    const double c1 = dt / (2.0 * p->dx[e1]), c2 = dt / (2.0 * p->dx[e2]);
    const long l1 = sh(p, li, lj, lk, e1, 1), l2 = sh(p, li, lj, lk, e2, 1);
    const long h1 = sh(p, i, j, k, e1, 1), h2 = sh(p, i, j, k, e2, 1);
    double l = Ipd[lc] - c1 * (a_e1[l1] * t_e1[l1] - a_e1[lc] * t_e1[lc]) +
               c1 * q[lc] * (a_e1[l1] - a_e1[lc]) -
               c2 * (a_e2[l2] * t_e2[l2] - a_e2[lc] * t_e2[lc]) + c2 * q[lc] * (a_e2[l2] - a_e2[lc]);
    double h = Imd[f] - c1 * (a_e1[h1] * t_e1[h1] - a_e1[f] * t_e1[f]) +
               c1 * q[f] * (a_e1[h1] - a_e1[f]) -
               c2 * (a_e2[h2] * t_e2[h2] - a_e2[f] * t_e2[f]) + c2 * q[f] * (a_e2[h2] - a_e2[f]);
    if (force) {
        l += 0.5 * dt * force[lc];
        h += 0.5 * dt * force[f];
    }
    *lo = l;
    *hi = h;
}

/* edge states of ncomp cell scalars s[n*N..] on every face direction:
 * states[(d * ncomp + n) * N + f]; force may be NULL (else ncomp comps) */
void orc_godunov_scalar_states(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                               const double* force, const double* umac, const double* vmac,
                               const double* wmac, double dt, double* states) {
    pg_t p = pg(g);
    const long N = (long)p.n[0] * p.n[1] * p.n[2];
    const double* ad[3] = {umac, vmac, wmac};
    double *Im[3], *Ip[3], *E[3], *T[3][3];
    for (int d = 0; d < 3; ++d) {
        Im[d] = dalloc(N);
        Ip[d] = dalloc(N);
        E[d] = dalloc(N);
        for (int e = 0; e < 3; ++e) T[d][e] = (e == d) ? NULL : dalloc(N);
    }
    for (int n = 0; n < ncomp; ++n) {
        const double* q = s + (long)n * N;
        const double* fn = force ? force + (long)n * N : NULL;
        const int cons = iconserv ? iconserv[n] : 0;
        for (int d = 0; d < 3; ++d) {
            orc_ppm_predict(g, q, d, NULL, ad[d], dt, Im[d], Ip[d]);
            edges(&p, d, Im[d], Ip[d], ad[d], E[d]);
        }
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) {
                if (e == d) continue;
                if (cons)
                    trans1_cons(&p, d, e, q, Im[d], Ip[d], E[e], ad[e], ad[d], dt, T[d][e]);
                else
                    trans1(&p, d, e, Im[d], Ip[d], E[e], ad[e], ad[d], dt, T[d][e]);
            }
        for (int d = 0; d < 3; ++d) {
            const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
            double* o = states + (long)(d * ncomp + n) * N;
            for (int k = 0; k < p.n[2]; ++k)
                for (int j = 0; j < p.n[1]; ++j)
                    for (int i = 0; i < p.n[0]; ++i) {
                        const long f = at(&p, i, j, k);
                        double lo, hi;
                        if (cons)
                            final_states_cons(&p, d, e1, e2, q, Im[d], Ip[d], T[e1][e2], ad[e1],
                                              T[e2][e1], ad[e2], fn, dt, f, i, j, k, &lo, &hi);
                        else
                            final_states(&p, d, e1, e2, Im[d], Ip[d], T[e1][e2], ad[e1],
                                         T[e2][e1], ad[e2], fn, dt, f, i, j, k, &lo, &hi);
                        o[f] = upwind(lo, hi, ad[d][f]);
                    }
        }
    }
    for (int d = 0; d < 3; ++d) {
        free(Im[d]);
        free(Ip[d]);
        free(E[d]);
        for (int e = 0; e < 3; ++e) free(T[d][e]);
    }
}

/* conv[n] = -div(u_mac s_face[n]) (+ s_n div(u_mac) when iconserv[n] == 0) */
void orc_scalar_advection(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                          const double* umac, const double* vmac, const double* wmac,
                          const double* states, double* conv) {
    pg_t p = pg(g);
    const long N = (long)p.n[0] * p.n[1] * p.n[2];
    const double* a[3] = {umac, vmac, wmac};
    for (int k = 0; k < p.n[2]; ++k)
        for (int j = 0; j < p.n[1]; ++j)
            for (int i = 0; i < p.n[0]; ++i) {
                const long c = at(&p, i, j, k);
                double divu = 0.0;
                for (int d = 0; d < 3; ++d) divu += (a[d][sh(&p, i, j, k, d, 1)] - a[d][c]) / p.dx[d];
                for (int n = 0; n < ncomp; ++n) {
                    double div = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double* st = states + (long)(d * ncomp + n) * N;
                        const long cp = sh(&p, i, j, k, d, 1);
                        div += (a[d][cp] * st[cp] - a[d][c] * st[c]) / p.dx[d];
                    }
                    const int cons = iconserv ? iconserv[n] : 0;
                    conv[(long)n * N + c] = -div + (cons ? 0.0 : s[(long)n * N + c] * divu);
                }
            }
}

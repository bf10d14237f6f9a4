/*
 * orc_mac.c -- FP64 CPU oracle for the MAC projection (cell-centred
 * variable-coefficient Poisson on face velocities).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product path).
 *
 * Reference behaviour restated:
 *  - face coefficient beta_f = dt / (0.5 (rho_L + rho_R))
 *    (average_cellcenter_to_face then invert(dt),
 *     src/convection/incflo_compute_MAC_projected_velocities.cpp:70-91);
 *    boundary faces of non-periodic directions take the adjacent cell's rho
 *    (first-order extrapolated density ghost cell).
 *  - BC mapping (src/projection/incflo_projection_bc.cpp:44-78): pressure
 *    in/outflow -> Dirichlet phi = 0 on the face; walls and mass inflow ->
 *    homogeneous Neumann (the face velocity is kept).
 *  - MacProjector::project (AMReX-Hydro, upstream): solve
 *        div(beta grad phi) = div(u_mac),  then  u_mac <- u_mac - beta grad phi,
 *    so that the discrete divergence of u_mac is the solver residual.
 * Discretisation (cells c, faces f):
 *    F_f = beta_f (phi_R - phi_L)/h  (interior / periodic faces)
 *    F_f = 0                           (Neumann boundary faces)
 *    F_lo = 2 beta_f phi_0 / h, F_hi = -2 beta_f phi_{n-1} / h (Dirichlet faces)
 *    (L phi)_c = sum_d (F_{d,c+1/2} - F_{d,c-1/2}) / h_d.
 * Multigrid: red-black Gauss-Seidel, cell-average restriction, piecewise
 * constant prolongation, coarse face beta = mean of the 4 fine faces on it,
 * BiCGStab bottom solve, same outer stopping rule as the nodal solve.
 * Face arrays: x-faces (n0+1) x n1 x n2 etc.; the periodic face n equals face 0.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define CIDX(i, j, k, g) ((long)(i) + (long)(g)->n[0] * ((long)(j) + (long)(g)->n[1] * (long)(k)))

static inline long ncell(const orc_geom* g) { return (long)g->n[0] * g->n[1] * g->n[2]; }

/* face array dims for direction d */
static inline void fdims(const orc_geom* g, int d, int fd[3])
{
    for (int e = 0; e < 3; ++e) fd[e] = g->n[e] + (e == d ? 1 : 0);
}

static inline long FIDX(const orc_geom* g, int d, int i, int j, int k)
{
    int fd[3];
    fdims(g, d, fd);
    return (long)i + (long)fd[0] * ((long)j + (long)fd[1] * (long)k);
}

long orc_mac_face_count(const orc_geom* g, int d)
{
    int fd[3];
    fdims(g, d, fd);
    return (long)fd[0] * fd[1] * fd[2];
}

int orc_mac_is_singular(const orc_geom* g)
{
    for (int d = 0; d < 3; ++d)
        if (!g->periodic[d] && (g->bclo[d] == ORC_BC_DIRICHLET || g->bchi[d] == ORC_BC_DIRICHLET))
            return 0;
    return 1;
}

/* Flux through face f (index along d = fi) of the cell pair; beta arrays per dir. */
static double face_flux(const orc_geom* g, const double* const beta[3], const double* phi, int d,
                        int i, int j, int k)
{
    /* face (i,j,k) in direction d lies between cells c-e_d and c */
    int c[3] = {i, j, k};
    const int n = g->n[d];
    const double bf = beta[d][FIDX(g, d, i, j, k)];
    const double h = g->dx[d];
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    lo[d] = c[d] - 1;
    hi[d] = c[d];
    if (g->periodic[d]) {
        if (lo[d] < 0) lo[d] += n;
        if (hi[d] >= n) hi[d] -= n;
    } else {
        if (c[d] == 0) {
            if (g->bclo[d] != ORC_BC_DIRICHLET) return 0.0;
            return 2.0 * bf * phi[CIDX(hi[0], hi[1], hi[2], g)] / h;
        }
        if (c[d] == n) {
            if (g->bchi[d] != ORC_BC_DIRICHLET) return 0.0;
            return -2.0 * bf * phi[CIDX(lo[0], lo[1], lo[2], g)] / h;
        }
    }
    return bf * (phi[CIDX(hi[0], hi[1], hi[2], g)] - phi[CIDX(lo[0], lo[1], lo[2], g)]) / h;
}

static double apply_cell(const orc_geom* g, const double* const beta[3], const double* phi, int i,
                         int j, int k)
{
    double acc = 0.0;
    int c[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        int p[3] = {i, j, k};
        p[d] = c[d] + 1;
        acc += (face_flux(g, beta, phi, d, p[0], p[1], p[2]) - face_flux(g, beta, phi, d, i, j, k)) /
               g->dx[d];
    }
    return acc;
}

void orc_mac_apply(const orc_geom* g, const double* bx, const double* by, const double* bz,
                   const double* phi, double* out)
{
    const double* beta[3] = {bx, by, bz};
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) out[CIDX(i, j, k, g)] = apply_cell(g, beta, phi, i, j, k);
}

void orc_mac_residual(const orc_geom* g, const double* bx, const double* by, const double* bz,
                      const double* b, const double* x, double* r)
{
    const double* beta[3] = {bx, by, bz};
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                long id = CIDX(i, j, k, g);
                r[id] = b[id] - apply_cell(g, beta, x, i, j, k);
            }
}

/* diagonal of L at cell (i,j,k) */
static double diag_cell(const orc_geom* g, const double* const beta[3], int i, int j, int k)
{
    double dg = 0.0;
    int c[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        const double ih2 = 1.0 / (g->dx[d] * g->dx[d]);
        int p[3] = {i, j, k};
        /* lo face */
        if (g->periodic[d] || c[d] > 0) dg -= beta[d][FIDX(g, d, p[0], p[1], p[2])] * ih2;
        else if (g->bclo[d] == ORC_BC_DIRICHLET) dg -= 2.0 * beta[d][FIDX(g, d, p[0], p[1], p[2])] * ih2;
        /* hi face */
        p[d] = c[d] + 1;
        if (g->periodic[d] || c[d] + 1 < g->n[d]) dg -= beta[d][FIDX(g, d, p[0], p[1], p[2])] * ih2;
        else if (g->bchi[d] == ORC_BC_DIRICHLET) dg -= 2.0 * beta[d][FIDX(g, d, p[0], p[1], p[2])] * ih2;
    }
    return dg;
}

/* One red-black Gauss-Seidel sweep (red: (i+j+k) even first). */
void orc_mac_gsrb(const orc_geom* g, const double* bx, const double* by, const double* bz,
                  const double* b, double* x)
{
    const double* beta[3] = {bx, by, bz};
    for (int color = 0; color < 2; ++color) {
#pragma omp parallel for collapse(2)
        for (int k = 0; k < g->n[2]; ++k)
            for (int j = 0; j < g->n[1]; ++j)
                for (int i = 0; i < g->n[0]; ++i) {
                    if (((i + j + k) & 1) != color) continue;
                    long id = CIDX(i, j, k, g);
                    const double dg = diag_cell(g, beta, i, j, k);
                    if (dg == 0.0) continue;
                    const double Lx = apply_cell(g, beta, x, i, j, k);
                    if (dg != 0.0) x[id] += (b[id] - Lx) / dg; /* zero rows: covered cells */
                }
    }
}

void orc_mac_restrict(const orc_geom* gf, const double* rf, double* rc)
{
    const int N0 = gf->n[0] / 2, N1 = gf->n[1] / 2, N2 = gf->n[2] / 2;
#pragma omp parallel for collapse(2)
    for (int K = 0; K < N2; ++K)
        for (int J = 0; J < N1; ++J)
            for (int I = 0; I < N0; ++I) {
                double s = 0.0;
                for (int q = 0; q < 8; ++q)
                    s += rf[CIDX(2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + ((q >> 2) & 1), gf)];
                rc[(long)I + (long)N0 * ((long)J + (long)N1 * K)] = 0.125 * s;
            }
}

void orc_mac_prolong_add(const orc_geom* gf, const double* xc, double* xf)
{
    const int N0 = gf->n[0] / 2, N1 = gf->n[1] / 2;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < gf->n[2]; ++k)
        for (int j = 0; j < gf->n[1]; ++j)
            for (int i = 0; i < gf->n[0]; ++i)
                xf[CIDX(i, j, k, gf)] += xc[(long)(i / 2) + (long)N0 * ((long)(j / 2) + (long)N1 * (k / 2))];
}

void orc_mac_beta_coarsen(const orc_geom* gf, const double* bf, int d, double* bc)
{
    orc_geom gc = *gf;
    for (int e = 0; e < 3; ++e) gc.n[e] = gf->n[e] / 2;
    int fd[3];
    fdims(&gc, d, fd);
#pragma omp parallel for collapse(2)
    for (int K = 0; K < fd[2]; ++K)
        for (int J = 0; J < fd[1]; ++J)
            for (int I = 0; I < fd[0]; ++I) {
                int C[3] = {I, J, K};
                double s = 0.0;
                for (int q = 0; q < 4; ++q) {
                    int f[3];
                    int t = 0;
                    for (int e = 0; e < 3; ++e) {
                        if (e == d) f[e] = 2 * C[e];
                        else { f[e] = 2 * C[e] + ((q >> t) & 1); ++t; }
                    }
                    s += bf[FIDX(gf, d, f[0], f[1], f[2])];
                }
                bc[FIDX(&gc, d, I, J, K)] = 0.25 * s;
            }
}

/* beta_f = dt / (0.5 (rho_L + rho_R)); boundary faces use the adjacent cell. */
void orc_mac_beta_from_density(const orc_geom* g, const double* rho, double dt, double* bx,
                               double* by, double* bz)
{
    double* beta[3] = {bx, by, bz};
    for (int d = 0; d < 3; ++d) {
        int fd[3];
        fdims(g, d, fd);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < fd[2]; ++k)
            for (int j = 0; j < fd[1]; ++j)
                for (int i = 0; i < fd[0]; ++i) {
                    int L[3] = {i, j, k}, R[3] = {i, j, k};
                    L[d] -= 1;
                    const int n = g->n[d];
                    if (g->periodic[d]) {
                        if (L[d] < 0) L[d] += n;
                        if (R[d] >= n) R[d] -= n;
                    } else {
                        if (L[d] < 0) L[d] = 0;
                        if (R[d] >= n) R[d] = n - 1;
                    }
                    const double rf = 0.5 * (rho[CIDX(L[0], L[1], L[2], g)] + rho[CIDX(R[0], R[1], R[2], g)]);
                    beta[d][FIDX(g, d, i, j, k)] = dt / rf;
                }
    }
}

void orc_mac_divergence(const orc_geom* g, const double* u, const double* v, const double* w,
                        double* div)
{
    const double* U[3] = {u, v, w};
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                double s = 0.0;
                for (int d = 0; d < 3; ++d) {
                    int p[3] = {i, j, k};
                    p[d] += 1;
                    s += (U[d][FIDX(g, d, p[0], p[1], p[2])] - U[d][FIDX(g, d, i, j, k)]) / g->dx[d];
                }
                div[CIDX(i, j, k, g)] = s;
            }
}

/* Face fluxes -beta grad phi (MacProjector::getFluxes sign convention). */
void orc_mac_fluxes(const orc_geom* g, const double* bx, const double* by, const double* bz,
                    const double* phi, double* fx, double* fy, double* fz)
{
    const double* beta[3] = {bx, by, bz};
    double* F[3] = {fx, fy, fz};
    for (int d = 0; d < 3; ++d) {
        int fd[3];
        fdims(g, d, fd);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < fd[2]; ++k)
            for (int j = 0; j < fd[1]; ++j)
                for (int i = 0; i < fd[0]; ++i)
                    F[d][FIDX(g, d, i, j, k)] = -face_flux(g, beta, phi, d, i, j, k);
    }
}

/* ---- BiCGStab on a MAC level ----------------------------------------- */
int orc_mac_bicgstab(const orc_geom* g, const double* bx, const double* by, const double* bz,
                     const double* b, double* x, double rtol, int maxiter, int* iters_out)
{
    const long n = ncell(g);
    double* r = (double*)calloc(n, sizeof(double));
    double* rh = (double*)calloc(n, sizeof(double));
    double* p = (double*)calloc(n, sizeof(double));
    double* v = (double*)calloc(n, sizeof(double));
    double* s = (double*)calloc(n, sizeof(double));
    double* t = (double*)calloc(n, sizeof(double));
    orc_mac_apply(g, bx, by, bz, x, t);
    for (long i = 0; i < n; ++i) { r[i] = b[i] - t[i]; rh[i] = r[i]; }
    const double rnorm0 = orc_norm_inf(r, n);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0, ret = 0;
    if (rnorm0 == 0.0) goto done;
    for (it = 1; it <= maxiter; ++it) {
        double rho_new = orc_dot(rh, r, n);
        if (rho_new == 0.0) { ret = 1; break; }
        double beta = (rho_new / rho) * (alpha / omega);
        for (long i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        orc_mac_apply(g, bx, by, bz, p, v);
        double rhv = orc_dot(rh, v, n);
        if (rhv == 0.0) { ret = 1; break; }
        alpha = rho_new / rhv;
        for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; s[i] = r[i] - alpha * v[i]; }
        if (orc_norm_inf(s, n) <= rtol * rnorm0) break;
        orc_mac_apply(g, bx, by, bz, s, t);
        double tt = orc_dot(t, t, n);
        if (tt == 0.0) { ret = 1; break; }
        omega = orc_dot(t, s, n) / tt;
        for (long i = 0; i < n; ++i) { x[i] += omega * s[i]; r[i] = s[i] - omega * t[i]; }
        if (orc_norm_inf(r, n) <= rtol * rnorm0) break;
        if (omega == 0.0) { ret = 1; break; }
        rho = rho_new;
    }
    if (it > maxiter) it = maxiter;
done:
    if (iters_out) *iters_out = it;
    free(r); free(rh); free(p); free(v); free(s); free(t);
    return ret;
}

/* ---- MLMG ------------------------------------------------------------- */
typedef struct {
    orc_geom g;
    double* beta[3];
    double* b;
    double* x;
    double* r;
    long n;
} mlev_t;

static void sub_mean(double* a, long n)
{
    double m = orc_mean(a, n);
    for (long i = 0; i < n; ++i) a[i] -= m;
}

static void mac_vcycle(mlev_t* L, int lev, int nlev, const orc_mg_opts* o, int singular, int* bot)
{
    mlev_t* l = &L[lev];
    if (lev == nlev - 1) {
        memset(l->x, 0, sizeof(double) * l->n);
        if (singular) sub_mean(l->b, l->n);
        int its = 0;
        orc_mac_bicgstab(&l->g, l->beta[0], l->beta[1], l->beta[2], l->b, l->x, o->bottom_rtol,
                         o->bottom_max_iter, &its);
        if (singular) sub_mean(l->x, l->n);
        *bot += its;
        return;
    }
    memset(l->x, 0, sizeof(double) * l->n);
    for (int s = 0; s < o->nu1; ++s) orc_mac_gsrb(&l->g, l->beta[0], l->beta[1], l->beta[2], l->b, l->x);
    orc_mac_residual(&l->g, l->beta[0], l->beta[1], l->beta[2], l->b, l->x, l->r);
    orc_mac_restrict(&l->g, l->r, L[lev + 1].b);
    mac_vcycle(L, lev + 1, nlev, o, singular, bot);
    orc_mac_prolong_add(&l->g, L[lev + 1].x, l->x);
    for (int s = 0; s < o->nu2; ++s) orc_mac_gsrb(&l->g, l->beta[0], l->beta[1], l->beta[2], l->b, l->x);
}

int orc_mac_solve(const orc_geom* g, const double* bx, const double* by, const double* bz,
                  const double* b_in, double* x, double rtol, double atol, const orc_mg_opts* o,
                  orc_solve_info* info)
{
    const int nlev = orc_nodal_nlevels(g, o->max_coarsening_level);
    const int singular = orc_mac_is_singular(g);
    mlev_t* L = (mlev_t*)calloc(nlev, sizeof(mlev_t));
    const double* b0[3] = {bx, by, bz};
    for (int l = 0; l < nlev; ++l) {
        if (l == 0) L[l].g = *g;
        else {
            L[l].g = L[l - 1].g;
            for (int d = 0; d < 3; ++d) { L[l].g.n[d] /= 2; L[l].g.dx[d] *= 2.0; }
        }
        L[l].n = ncell(&L[l].g);
        for (int d = 0; d < 3; ++d) {
            long nf = orc_mac_face_count(&L[l].g, d);
            L[l].beta[d] = (double*)malloc(sizeof(double) * nf);
            if (l == 0) memcpy(L[l].beta[d], b0[d], sizeof(double) * nf);
            else orc_mac_beta_coarsen(&L[l - 1].g, L[l - 1].beta[d], d, L[l].beta[d]);
        }
        L[l].b = (double*)calloc(L[l].n, sizeof(double));
        L[l].x = (double*)calloc(L[l].n, sizeof(double));
        L[l].r = (double*)calloc(L[l].n, sizeof(double));
    }
    const long n = L[0].n;
    double* b = (double*)malloc(sizeof(double) * n);
    double* res = (double*)malloc(sizeof(double) * n);
    memcpy(b, b_in, sizeof(double) * n);
    if (singular) sub_mean(b, n);
    const double bnorm = orc_norm_inf(b, n);
    orc_mac_residual(g, bx, by, bz, b, x, res);
    const double rnorm0 = orc_norm_inf(res, n);
    const double tol = fmax(atol, rtol * bnorm);
    double rnorm = rnorm0;
    int it = 0, converged = (rnorm0 <= tol), bot = 0;
    while (!converged && it < o->max_iter) {
        ++it;
        memcpy(L[0].b, res, sizeof(double) * n);
        mac_vcycle(L, 0, nlev, o, singular, &bot);
        for (long i = 0; i < n; ++i) x[i] += L[0].x[i];
        orc_mac_residual(g, bx, by, bz, b, x, res);
        rnorm = orc_norm_inf(res, n);
        if (o->verbose) fprintf(stderr, "orc MAC MLMG: iter %d |r| %.6e\n", it, rnorm);
        converged = (rnorm <= tol);
    }
    if (singular) sub_mean(x, n);
    if (info) {
        info->iters = it;
        info->converged = converged;
        info->bnorm = bnorm;
        info->resnorm0 = rnorm0;
        info->resnorm = rnorm;
        info->nlevels = nlev;
        info->bottom_iters = bot;
    }
    for (int l = 0; l < nlev; ++l) {
        for (int d = 0; d < 3; ++d) free(L[l].beta[d]);
        free(L[l].b); free(L[l].x); free(L[l].r);
    }
    free(L); free(b); free(res);
    return converged ? 0 : 1;
}

/* MacProjector::project: solve, then u_mac -= beta grad phi (in place). */
int orc_mac_project(const orc_geom* g, const double* bx, const double* by, const double* bz,
                    double* u, double* v, double* w, double* phi, double rtol, double atol,
                    const orc_mg_opts* o, orc_solve_info* info)
{
    const long n = ncell(g);
    double* rhs = (double*)malloc(sizeof(double) * n);
    orc_mac_divergence(g, u, v, w, rhs);
    int ret = orc_mac_solve(g, bx, by, bz, rhs, phi, rtol, atol, o, info);
    const double* beta[3] = {bx, by, bz};
    double* U[3] = {u, v, w};
    for (int d = 0; d < 3; ++d) {
        int fd[3];
        fdims(g, d, fd);
#pragma omp parallel for collapse(2)
        for (int k = 0; k < fd[2]; ++k)
            for (int j = 0; j < fd[1]; ++j)
                for (int i = 0; i < fd[0]; ++i)
                    U[d][FIDX(g, d, i, j, k)] -= face_flux(g, beta, phi, d, i, j, k);
    }
    free(rhs);
    return ret;
}

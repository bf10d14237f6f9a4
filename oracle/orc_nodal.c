/*
 * orc_nodal.c -- FP64 CPU oracle: nodal operator, D, G, MG transfer operators,
 * BiCGStab bottom solve and the nodal MLMG V-cycle driver.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product path).
 *
 * Reference behaviour restated here:
 *  - incflo::ApplyNodalProjection core, src/projection/incflo_apply_nodal_projection.cpp:95-267
 *    (sigma = s/rho at :117, inflow ghosts :137-163, project(rtol,atol) at :215,
 *     p_nd = phi and gp = G(phi) at :245-253).
 *  - Hydro::NodalProjector / MLNodeLaplacian / MLMG are upstream AMReX (not in the
 *    reference tree); the operator restated here is the Q1 finite-element
 *    nodal Laplacian L = div(sigma grad) in FD normalisation, assembled per
 *    cell from the 8x8 element matrix
 *        E[p][m] = sum_d K[p_d][m_d] prod_{e!=d} M[p_e][m_e] / (36 h_d^2),
 *        K = [[1,-1],[-1,1]],  M = [[2,1],[1,2]],
 *    and (L phi)_n = - sum_{cells c around n} sigma_c sum_m E[p][m] phi_m.
 *  - MLMG stopping rule: ||b - L x||_inf <= max(atol, rtol*||b||_inf)
 *    (incflo passes rtol=1e-11, atol=1e-14 by default, src/incflo.H:453-454).
 *  - maxiter/bottom_maxiter defaults 100/100 for the nodal solve
 *    (Docs/sphinx_documentation/source/InputsMultigrid.rst:23-30).
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define IDX(i, j, k, nx, ny) ((long)(i) + (long)(nx) * ((long)(j) + (long)(ny) * (long)(k)))

static inline int nn_of(const orc_geom* g, int d) { return g->periodic[d] ? g->n[d] : g->n[d] + 1; }
static inline long nnodes(const orc_geom* g) { return (long)nn_of(g, 0) * nn_of(g, 1) * nn_of(g, 2); }
static inline long ncells(const orc_geom* g) { return (long)g->n[0] * g->n[1] * g->n[2]; }

void orc_default_opts(orc_mg_opts* o)
{
    o->max_iter = 100;
    o->bottom_max_iter = 100;
    o->bottom_rtol = 1.0e-4;
    o->nu1 = 2;
    o->nu2 = 2;
    o->omega = 8.0 / 9.0;
    o->max_coarsening_level = 100;
    o->verbose = 0;
}

double orc_norm_inf(const double* a, long n)
{
    double m = 0.0;
#pragma omp parallel for reduction(max : m)
    for (long i = 0; i < n; ++i) {
        double v = fabs(a[i]);
        if (v > m) m = v;
    }
    return m;
}

double orc_dot(const double* a, const double* b, long n)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

double orc_mean(const double* a, long n)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (long i = 0; i < n; ++i) s += a[i];
    return s / (double)n;
}

/* 8x8 element matrix for the current cell size. */
static void element_matrix(const orc_geom* g, double E[8][8])
{
    for (int p = 0; p < 8; ++p)
        for (int m = 0; m < 8; ++m) {
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
                int pd = (p >> d) & 1, md = (m >> d) & 1;
                double kd = (pd == md) ? 1.0 : -1.0;
                double w = 1.0;
                for (int e = 0; e < 3; ++e) {
                    if (e == d) continue;
                    int pe = (p >> e) & 1, me = (m >> e) & 1;
                    w *= (pe == me) ? 2.0 : 1.0;
                }
                s += kd * w / (36.0 * g->dx[d] * g->dx[d]);
            }
            E[p][m] = s;
        }
}

int orc_is_dirichlet_node(const orc_geom* g, int i, int j, int k)
{
    int idx[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        if (g->periodic[d]) continue;
        if (idx[d] == 0 && g->bclo[d] == ORC_BC_DIRICHLET) return 1;
        if (idx[d] == g->n[d] && g->bchi[d] == ORC_BC_DIRICHLET) return 1;
    }
    return 0;
}

int orc_nodal_is_singular(const orc_geom* g)
{
    for (int d = 0; d < 3; ++d) {
        if (g->periodic[d]) continue;
        if (g->bclo[d] == ORC_BC_DIRICHLET || g->bchi[d] == ORC_BC_DIRICHLET) return 0;
    }
    return 1;
}

/* cell index in direction d for node index i and side a (0: cell i-1, 1: cell i);
 * returns -1 when the cell lies outside a non-periodic domain. */
static inline int cell_of(const orc_geom* g, int d, int i, int a)
{
    int c = i - 1 + a;
    if (g->periodic[d]) {
        int n = g->n[d];
        return ((c % n) + n) % n;
    }
    return (c < 0 || c >= g->n[d]) ? -1 : c;
}

static inline int node_wrap(const orc_geom* g, int d, int i)
{
    if (g->periodic[d]) {
        int n = g->n[d];
        return ((i % n) + n) % n;
    }
    return i;
}

/* (L phi) at one node, element-by-element. */
static double apply_node(const orc_geom* g, const double E[8][8], const double* sigma,
                         const double* phi, int i, int j, int k)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1);
    const int n[3] = {i, j, k};
    double acc = 0.0;
    for (int cz = 0; cz < 2; ++cz)
        for (int cy = 0; cy < 2; ++cy)
            for (int cx = 0; cx < 2; ++cx) {
                int a[3] = {cx, cy, cz};
                int c[3];
                int ok = 1;
                for (int d = 0; d < 3; ++d) {
                    c[d] = cell_of(g, d, n[d], a[d]);
                    if (c[d] < 0) ok = 0;
                }
                if (!ok) continue;
                double s = sigma[IDX(c[0], c[1], c[2], g->n[0], g->n[1])];
                if (s == 0.0) continue;
                /* local corner of node n inside this cell */
                int p = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
                double e = 0.0;
                for (int m = 0; m < 8; ++m) {
                    int mi = node_wrap(g, 0, (n[0] - 1 + cx) + (m & 1));
                    int mj = node_wrap(g, 1, (n[1] - 1 + cy) + ((m >> 1) & 1));
                    int mk = node_wrap(g, 2, (n[2] - 1 + cz) + ((m >> 2) & 1));
                    e += E[p][m] * phi[IDX(mi, mj, mk, nx, ny)];
                }
                acc -= s * e;
            }
    return acc;
}

void orc_nodal_apply(const orc_geom* g, const double* sigma, const double* phi, double* out)
{
    double E[8][8];
    element_matrix(g, E);
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                out[IDX(i, j, k, nx, ny)] = apply_node(g, E, sigma, phi, i, j, k);
}

void orc_nodal_diag(const orc_geom* g, const double* sigma, double* diag)
{
    double E[8][8];
    element_matrix(g, E);
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int n[3] = {i, j, k};
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) {
                    int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
                    int c0 = cell_of(g, 0, n[0], ax), c1 = cell_of(g, 1, n[1], ay),
                        c2 = cell_of(g, 2, n[2], az);
                    if (c0 < 0 || c1 < 0 || c2 < 0) continue;
                    int p = (1 - ax) | ((1 - ay) << 1) | ((1 - az) << 2);
                    acc -= sigma[IDX(c0, c1, c2, g->n[0], g->n[1])] * E[p][p];
                }
                diag[IDX(i, j, k, nx, ny)] = acc;
            }
}

void orc_nodal_residual(const orc_geom* g, const double* sigma, const double* b,
                        const double* x, double* r)
{
    double E[8][8];
    element_matrix(g, E);
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long id = IDX(i, j, k, nx, ny);
                r[id] = orc_is_dirichlet_node(g, i, j, k)
                            ? 0.0
                            : b[id] - apply_node(g, E, sigma, x, i, j, k);
            }
}

void orc_nodal_jacobi(const orc_geom* g, const double* sigma, const double* b,
                      const double* xin, double* xout, double omega)
{
    double E[8][8];
    element_matrix(g, E);
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
    double* diag = (double*)malloc(sizeof(double) * nnodes(g));
    orc_nodal_diag(g, sigma, diag);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long id = IDX(i, j, k, nx, ny);
                if (orc_is_dirichlet_node(g, i, j, k) || diag[id] == 0.0) {
                    xout[id] = orc_is_dirichlet_node(g, i, j, k) ? 0.0 : xin[id];
                    continue;
                }
                double res = b[id] - apply_node(g, E, sigma, xin, i, j, k);
                xout[id] = xin[id] + omega * res / diag[id];
            }
    free(diag);
}

/* Full weighting R = P^T / 8: 1-D weights (1/4, 1/2, 1/4) per direction. */
void orc_nodal_restrict(const orc_geom* gf, const double* rf, double* rc)
{
    orc_geom gc = *gf;
    for (int d = 0; d < 3; ++d) { gc.n[d] = gf->n[d] / 2; gc.dx[d] = 2.0 * gf->dx[d]; }
    const int fx = nn_of(gf, 0), fy = nn_of(gf, 1), fz = nn_of(gf, 2);
    const int cx = nn_of(&gc, 0), cy = nn_of(&gc, 1), cz = nn_of(&gc, 2);
#pragma omp parallel for collapse(2)
    for (int K = 0; K < cz; ++K)
        for (int J = 0; J < cy; ++J)
            for (int I = 0; I < cx; ++I) {
                double acc = 0.0;
                if (!orc_is_dirichlet_node(&gc, I, J, K)) {
                    for (int oz = -1; oz <= 1; ++oz)
                        for (int oy = -1; oy <= 1; ++oy)
                            for (int ox = -1; ox <= 1; ++ox) {
                                int f[3] = {2 * I + ox, 2 * J + oy, 2 * K + oz};
                                int o[3] = {ox, oy, oz};
                                int nf[3] = {fx, fy, fz};
                                int ok = 1;
                                double w = 1.0;
                                for (int d = 0; d < 3; ++d) {
                                    if (gf->periodic[d]) {
                                        f[d] = ((f[d] % nf[d]) + nf[d]) % nf[d];
                                    } else if (f[d] < 0 || f[d] >= nf[d]) {
                                        ok = 0;
                                    }
                                    w *= (o[d] == 0) ? 0.5 : 0.25;
                                }
                                if (ok) acc += w * rf[IDX(f[0], f[1], f[2], fx, fy)];
                            }
                }
                rc[IDX(I, J, K, cx, cy)] = acc;
            }
}

/* Trilinear interpolation of coarse nodes, added onto fine nodes. */
void orc_nodal_prolong_add(const orc_geom* gf, const double* xc, double* xf)
{
    orc_geom gc = *gf;
    for (int d = 0; d < 3; ++d) { gc.n[d] = gf->n[d] / 2; gc.dx[d] = 2.0 * gf->dx[d]; }
    const int fx = nn_of(gf, 0), fy = nn_of(gf, 1), fz = nn_of(gf, 2);
    const int cx = nn_of(&gc, 0), cy = nn_of(&gc, 1);
    const int nc[3] = {nn_of(&gc, 0), nn_of(&gc, 1), nn_of(&gc, 2)};
#pragma omp parallel for collapse(2)
    for (int k = 0; k < fz; ++k)
        for (int j = 0; j < fy; ++j)
            for (int i = 0; i < fx; ++i) {
                if (orc_is_dirichlet_node(gf, i, j, k)) continue;
                int f[3] = {i, j, k};
                int lo[3], hi[3];
                for (int d = 0; d < 3; ++d) {
                    lo[d] = f[d] >> 1;
                    hi[d] = (f[d] & 1) ? lo[d] + 1 : lo[d];
                    if (gf->periodic[d]) hi[d] %= nc[d];
                }
                double acc = 0.0;
                for (int c = 0; c < 8; ++c) {
                    int I = (c & 1) ? hi[0] : lo[0];
                    int J = (c & 2) ? hi[1] : lo[1];
                    int K = (c & 4) ? hi[2] : lo[2];
                    acc += xc[IDX(I, J, K, cx, cy)];
                }
                xf[IDX(i, j, k, fx, fy)] += 0.125 * acc;
            }
}

void orc_sigma_coarsen(const orc_geom* gf, const double* sf, double* sc)
{
    const int Nx = gf->n[0] / 2, Ny = gf->n[1] / 2, Nz = gf->n[2] / 2;
#pragma omp parallel for collapse(2)
    for (int K = 0; K < Nz; ++K)
        for (int J = 0; J < Ny; ++J)
            for (int I = 0; I < Nx; ++I) {
                double s = 0.0;
                for (int a = 0; a < 8; ++a)
                    s += sf[IDX(2 * I + (a & 1), 2 * J + ((a >> 1) & 1), 2 * K + ((a >> 2) & 1),
                                gf->n[0], gf->n[1])];
                sc[IDX(I, J, K, Nx, Ny)] = 0.125 * s;
            }
}

/* Weighted-Jacobi factor for a level: the user's omega is meant for isotropic
 * cells, where the Jacobi-preconditioned constant-sigma symbol peaks at 3/2.
 * On anisotropic cells the peak is
 *   lmax = max(4 max_d ih2_d, 4/3 max_{d<e}(ih2_d+ih2_e), 4/9 S) / (8/9 S),
 * (the symbol's values at the corners of [0,pi]^3), so omega is scaled by
 * min(1, 3/2 / lmax) to keep |1 - omega lambda| < 1. */
double orc_nodal_omega(const orc_geom* g, double omega)
{
    double ih2[3], S = 0.0;
    for (int d = 0; d < 3; ++d) { ih2[d] = 1.0 / (g->dx[d] * g->dx[d]); S += ih2[d]; }
    double m1 = fmax(ih2[0], fmax(ih2[1], ih2[2]));
    double m2 = fmax(ih2[0] + ih2[1], fmax(ih2[0] + ih2[2], ih2[1] + ih2[2]));
    double lmax = fmax(4.0 * m1, fmax(4.0 / 3.0 * m2, 4.0 / 9.0 * S)) / (8.0 / 9.0 * S);
    return omega * fmin(1.0, 1.5 / lmax);
}

/* A level is coarsened while every cell count is even and the coarse count is >= 4. */
int orc_nodal_nlevels(const orc_geom* g, int max_coarsening_level)
{
    int n[3] = {g->n[0], g->n[1], g->n[2]};
    int nl = 1;
    while (nl - 1 < max_coarsening_level) {
        int ok = 1;
        for (int d = 0; d < 3; ++d)
            if (n[d] % 2 != 0 || n[d] / 2 < 4) ok = 0;
        if (!ok) break;
        for (int d = 0; d < 3; ++d) n[d] /= 2;
        ++nl;
    }
    return nl;
}

/* ---- D and G ---------------------------------------------------------- */

#define GIDX(i, j, k, g) \
    ((long)((i) + 1) + (long)((g)->n[0] + 2) * ((long)((j) + 1) + (long)((g)->n[1] + 2) * (long)((k) + 1)))

/* Ghost rule: all outside dims periodic -> periodic image; exactly one
 * non-periodic outside dim whose face is inflow -> that face's inflow
 * velocity; otherwise 0 (vel->setBndry(0) followed by the inflow fill,
 * src/projection/incflo_apply_nodal_projection.cpp:137-162). */
void orc_fill_vel_ghosts(const orc_geom* g, double* ug, const double* inflow_vel)
{
    const long ng3 = (long)(g->n[0] + 2) * (g->n[1] + 2) * (g->n[2] + 2);
    for (int k = -1; k <= g->n[2]; ++k)
        for (int j = -1; j <= g->n[1]; ++j)
            for (int i = -1; i <= g->n[0]; ++i) {
                int c[3] = {i, j, k};
                int outside = 0, nonper = 0, face = -1;
                int src[3];
                for (int d = 0; d < 3; ++d) {
                    src[d] = c[d];
                    if (c[d] < 0 || c[d] >= g->n[d]) {
                        outside = 1;
                        if (g->periodic[d]) {
                            src[d] = (c[d] + g->n[d]) % g->n[d];
                        } else {
                            ++nonper;
                            face = 2 * d + (c[d] < 0 ? 0 : 1);
                        }
                    }
                }
                if (!outside) continue;
                for (int comp = 0; comp < 3; ++comp) {
                    double v;
                    if (nonper == 0) {
                        v = ug[comp * ng3 + GIDX(src[0], src[1], src[2], g)];
                    } else if (nonper == 1 && inflow_vel) {
                        int d = face / 2;
                        int bc = (face & 1) ? g->bchi[d] : g->bclo[d];
                        v = (bc == ORC_BC_INFLOW) ? inflow_vel[face * 3 + comp] : 0.0;
                    } else {
                        v = 0.0;
                    }
                    ug[comp * ng3 + GIDX(i, j, k, g)] = v;
                }
            }
}

void orc_nodal_divergence(const orc_geom* g, const double* ug, double* rhs)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
    const long ng3 = (long)(g->n[0] + 2) * (g->n[1] + 2) * (g->n[2] + 2);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) {
                    int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
                    long gi = GIDX(i - 1 + ax, j - 1 + ay, k - 1 + az, g);
                    double ux = ug[gi], uy = ug[ng3 + gi], uz = ug[2 * ng3 + gi];
                    acc += (ax ? 1.0 : -1.0) * ux / (4.0 * g->dx[0]);
                    acc += (ay ? 1.0 : -1.0) * uy / (4.0 * g->dx[1]);
                    acc += (az ? 1.0 : -1.0) * uz / (4.0 * g->dx[2]);
                }
                rhs[IDX(i, j, k, nx, ny)] = acc;
            }
}

void orc_nodal_gradient(const orc_geom* g, const double* phi, double* gphi)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1);
    const long nc = ncells(g);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                double gx = 0.0, gy = 0.0, gz = 0.0;
                for (int a = 0; a < 8; ++a) {
                    int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
                    double v = phi[IDX(node_wrap(g, 0, i + ax), node_wrap(g, 1, j + ay),
                                       node_wrap(g, 2, k + az), nx, ny)];
                    gx += (ax ? v : -v);
                    gy += (ay ? v : -v);
                    gz += (az ? v : -v);
                }
                long c = IDX(i, j, k, g->n[0], g->n[1]);
                gphi[c] = gx / (4.0 * g->dx[0]);
                gphi[nc + c] = gy / (4.0 * g->dx[1]);
                gphi[2 * nc + c] = gz / (4.0 * g->dx[2]);
            }
}

/* ---- BiCGStab (van der Vorst 1992) on one level ------------------------ */

static void masked_apply(const orc_geom* g, const double* sigma, const double* x, double* y)
{
    orc_nodal_apply(g, sigma, x, y);
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                if (orc_is_dirichlet_node(g, i, j, k)) y[IDX(i, j, k, nx, ny)] = 0.0;
}

int orc_nodal_bicgstab(const orc_geom* g, const double* sigma, const double* b, double* x,
                       double rtol, int maxiter, int* iters_out)
{
    const long n = nnodes(g);
    double* r = (double*)calloc(n, sizeof(double));
    double* rh = (double*)calloc(n, sizeof(double));
    double* p = (double*)calloc(n, sizeof(double));
    double* v = (double*)calloc(n, sizeof(double));
    double* s = (double*)calloc(n, sizeof(double));
    double* t = (double*)calloc(n, sizeof(double));
    masked_apply(g, sigma, x, t);
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
        masked_apply(g, sigma, p, v);
        double rhv = orc_dot(rh, v, n);
        if (rhv == 0.0) { ret = 1; break; }
        alpha = rho_new / rhv;
        for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; s[i] = r[i] - alpha * v[i]; }
        if (orc_norm_inf(s, n) <= rtol * rnorm0) { for (long i = 0; i < n; ++i) r[i] = s[i]; break; }
        masked_apply(g, sigma, s, t);
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

/* ---- MLMG -------------------------------------------------------------- */

typedef struct {
    orc_geom g;
    double* sigma;
    double* b;
    double* x;
    double* x2;
    double* r;
    long nn;
} level_t;

static void subtract_mean(double* a, long n)
{
    double m = orc_mean(a, n);
    for (long i = 0; i < n; ++i) a[i] -= m;
}

static void vcycle(level_t* L, int lev, int nlev, const orc_mg_opts* o, int singular, int* bot_its)
{
    level_t* l = &L[lev];
    if (lev == nlev - 1) {
        memset(l->x, 0, sizeof(double) * l->nn);
        if (singular) subtract_mean(l->b, l->nn);
        int its = 0;
        orc_nodal_bicgstab(&l->g, l->sigma, l->b, l->x, o->bottom_rtol, o->bottom_max_iter, &its);
        if (singular) subtract_mean(l->x, l->nn);
        *bot_its += its;
        return;
    }
    memset(l->x, 0, sizeof(double) * l->nn);
    for (int s = 0; s < o->nu1; ++s) {
        orc_nodal_jacobi(&l->g, l->sigma, l->b, l->x, l->x2, orc_nodal_omega(&l->g, o->omega));
        double* t = l->x; l->x = l->x2; l->x2 = t;
    }
    orc_nodal_residual(&l->g, l->sigma, l->b, l->x, l->r);
    orc_nodal_restrict(&l->g, l->r, L[lev + 1].b);
    vcycle(L, lev + 1, nlev, o, singular, bot_its);
    orc_nodal_prolong_add(&l->g, L[lev + 1].x, l->x);
    for (int s = 0; s < o->nu2; ++s) {
        orc_nodal_jacobi(&l->g, l->sigma, l->b, l->x, l->x2, orc_nodal_omega(&l->g, o->omega));
        double* t = l->x; l->x = l->x2; l->x2 = t;
    }
}

int orc_nodal_solve(const orc_geom* g, const double* sigma, const double* b_in, double* x,
                    double rtol, double atol, const orc_mg_opts* o, orc_solve_info* info)
{
    const int nlev = orc_nodal_nlevels(g, o->max_coarsening_level);
    const int singular = orc_nodal_is_singular(g);
    level_t* L = (level_t*)calloc(nlev, sizeof(level_t));
    for (int l = 0; l < nlev; ++l) {
        if (l == 0) {
            L[l].g = *g;
        } else {
            L[l].g = L[l - 1].g;
            for (int d = 0; d < 3; ++d) { L[l].g.n[d] /= 2; L[l].g.dx[d] *= 2.0; }
        }
        L[l].nn = nnodes(&L[l].g);
        long nc = ncells(&L[l].g);
        L[l].sigma = (double*)malloc(sizeof(double) * nc);
        if (l == 0) memcpy(L[l].sigma, sigma, sizeof(double) * nc);
        else orc_sigma_coarsen(&L[l - 1].g, L[l - 1].sigma, L[l].sigma);
        L[l].b = (double*)calloc(L[l].nn, sizeof(double));
        L[l].x = (double*)calloc(L[l].nn, sizeof(double));
        L[l].x2 = (double*)calloc(L[l].nn, sizeof(double));
        L[l].r = (double*)calloc(L[l].nn, sizeof(double));
    }
    const long n = L[0].nn;
    double* b = (double*)malloc(sizeof(double) * n);
    double* res = (double*)malloc(sizeof(double) * n);
    memcpy(b, b_in, sizeof(double) * n);
    /* Dirichlet nodes carry no equation. */
    {
        const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i)
                    if (orc_is_dirichlet_node(g, i, j, k)) {
                        b[IDX(i, j, k, nx, ny)] = 0.0;
                        x[IDX(i, j, k, nx, ny)] = 0.0;
                    }
    }
    if (singular) subtract_mean(b, n);
    const double bnorm = orc_norm_inf(b, n);
    orc_nodal_residual(g, sigma, b, x, res);
    const double rnorm0 = orc_norm_inf(res, n);
    const double tol = fmax(atol, rtol * bnorm);
    double rnorm = rnorm0;
    int it = 0, converged = (rnorm0 <= tol), bot_its = 0;
    if (o->verbose) fprintf(stderr, "orc MLMG: nlev %d bnorm %.3e r0 %.3e tol %.3e\n", nlev, bnorm, rnorm0, tol);
    while (!converged && it < o->max_iter) {
        ++it;
        memcpy(L[0].b, res, sizeof(double) * n);
        vcycle(L, 0, nlev, o, singular, &bot_its);
        for (long i = 0; i < n; ++i) x[i] += L[0].x[i];
        orc_nodal_residual(g, sigma, b, x, res);
        rnorm = orc_norm_inf(res, n);
        if (o->verbose) fprintf(stderr, "orc MLMG: iter %d |r| %.6e rel %.3e\n", it, rnorm, rnorm / (bnorm > 0 ? bnorm : 1));
        converged = (rnorm <= tol);
    }
    if (singular) subtract_mean(x, n);
    if (info) {
        info->iters = it;
        info->converged = converged;
        info->bnorm = bnorm;
        info->resnorm0 = rnorm0;
        info->resnorm = rnorm;
        info->nlevels = nlev;
        info->bottom_iters = bot_its;
    }
    for (int l = 0; l < nlev; ++l) {
        free(L[l].sigma); free(L[l].b); free(L[l].x); free(L[l].x2); free(L[l].r);
    }
    free(L); free(b); free(res);
    return converged ? 0 : 1;
}

int orc_nodal_project(const orc_geom* g, const double* sigma, double* vel_g, double* phi,
                      double* gphi, double rtol, double atol, const orc_mg_opts* o,
                      orc_solve_info* info)
{
    const long nn = nnodes(g), nc = ncells(g);
    const long ng3 = (long)(g->n[0] + 2) * (g->n[1] + 2) * (g->n[2] + 2);
    double* rhs = (double*)malloc(sizeof(double) * nn);
    orc_nodal_divergence(g, vel_g, rhs);
    memset(phi, 0, sizeof(double) * nn);
    int ret = orc_nodal_solve(g, sigma, rhs, phi, rtol, atol, o, info);
    orc_nodal_gradient(g, phi, gphi);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                long c = IDX(i, j, k, g->n[0], g->n[1]);
                long gi = GIDX(i, j, k, g);
                for (int d = 0; d < 3; ++d) vel_g[d * ng3 + gi] -= sigma[c] * gphi[d * nc + c];
            }
    free(rhs);
    return ret;
}

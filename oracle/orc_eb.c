/* orc_eb.c -- FP64 CPU restatement of the embedded-boundary (cut-cell) pieces
 * of the nodal projection: EB geometry of an analytic body (the cylinder of
 * src/embedded_boundaries/eb_cylinder.cpp:53-63, axis along z, fluid outside),
 * the EB nodal operator, divergence and gradient, and a Krylov solve of the
 * EB projection.  TEST INFRASTRUCTURE: the checker of the HIP EB kernels.
 *
 * The reference builds EB data with AMReX's EB2 and assembles its EB nodal
 * operator inside MLNodeLaplacian (both not vendored); the discretisation
 * restated here is the Q1 finite-element one on the fluid part of each cell:
 *   E_c[a][b] = sum_d (1/h_d^2) s_a^d s_b^d  int_{fluid, unit cell} prod_{e!=d} l_a^e l_b^e
 *   (L phi)_n = - sum_{cells c of n} sigma_c sum_b E_c[n_local][b] phi_b
 *   b_n       = - sum_{c} sum_d u_{c,d} s_n^d / h_d  int_{fluid, unit cell} prod_{e!=d} l_n^e
 *   (G phi)_c = (1 / V_c) sum_b phi_b sum_d s_b^d / h_d  int_{fluid} prod_{e!=d} l_b^e
 * with l_a^e the 1-D linear shape function of corner a along e (1 - xi or xi)
 * and s_a^d = d l_a^d / d xi = -1 or +1; V_c = fluid volume fraction.  Regular
 * cells reduce to the standard operator (oracle/orc_nodal.c).  The integrals
 * come from the fluid moments m[p][q][r] = int xi^p eta^q zeta^r (p,q,r <= 2)
 * over the fluid part of the unit cell.  Here they are computed exactly
 * along y and by adaptive Gauss-Legendre quadrature along x (independent of
 * the product's boundary-integral computation). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define IDX(i, j, k, nx, ny) ((long)(i) + (long)(nx) * ((long)(j) + (long)(ny) * (long)(k)))

static inline int nn_of(const orc_geom* g, int d) { return g->periodic[d] ? g->n[d] : g->n[d] + 1; }

/* ---- geometry: cylinder along z, fluid where (x-cx)^2 + (y-cy)^2 >= r^2 ---- */

static const double GL8_X[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                                -0.1834346424956498, 0.1834346424956498, 0.5255324099163290,
                                0.7966664774136267, 0.9602898564975363};
static const double GL8_W[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                                0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                                0.2223810344533745, 0.1012285362903763};

typedef struct {
    double x0, y0, hx, hy, cx, cy, r;
} sq_t;

/* int over the fluid part of the vertical line xi = const of eta^q, q = 0..2:
 * exact (the fluid part is [0,1] minus one interval) */
static void line_moments(const sq_t* S, double xi, double f[3])
{
    const double x = S->x0 + S->hx * xi;
    const double dx = x - S->cx;
    double lo = 1.0, hi = 0.0; /* covered interval in eta (empty by default) */
    if (dx * dx < S->r * S->r) {
        const double h = sqrt(S->r * S->r - dx * dx);
        lo = (S->cy - h - S->y0) / S->hy;
        hi = (S->cy + h - S->y0) / S->hy;
        lo = fmax(lo, 0.0);
        hi = fmin(hi, 1.0);
    }
    for (int q = 0; q < 3; ++q) {
        double v = 1.0 / (q + 1);
        if (hi > lo) v -= (pow(hi, q + 1) - pow(lo, q + 1)) / (q + 1);
        f[q] = v;
    }
}

/* Gauss-Legendre (8 points) of the 9 moment integrands over xi in [a, b];
 * mode 1 / 2: the integrand has a square-root endpoint singularity (where the
 * circle's x-extent begins / ends) at a / b, removed by xi = a + (b-a) t^2 /
 * xi = b - (b-a) t^2, t in [t0, t1] */
static void gl8(const sq_t* S, double a, double b, int mode, double t0, double t1, double m[9])
{
    for (int q = 0; q < 9; ++q) m[q] = 0.0;
    for (int g = 0; g < 8; ++g) {
        const double t = 0.5 * (t0 + t1) + 0.5 * (t1 - t0) * GL8_X[g];
        double w = 0.5 * (t1 - t0) * GL8_W[g];
        double xi = t;
        if (mode == 1) {
            xi = a + (b - a) * t * t;
            w *= 2.0 * (b - a) * t;
        } else if (mode == 2) {
            xi = b - (b - a) * t * t;
            w *= 2.0 * (b - a) * t;
        }
        double f[3];
        line_moments(S, xi, f);
        double xp = 1.0;
        for (int p = 0; p < 3; ++p) {
            for (int q = 0; q < 3; ++q) m[p * 3 + q] += w * xp * f[q];
            xp *= xi;
        }
    }
}

/* adaptive bisection in t until two levels agree to 1e-15 */
static void adapt(const sq_t* S, double a, double b, int mode, double t0, double t1,
                  const double whole[9], int depth, double m[9])
{
    double l[9], r[9];
    const double c = 0.5 * (t0 + t1);
    gl8(S, a, b, mode, t0, c, l);
    gl8(S, a, b, mode, c, t1, r);
    double err = 0.0;
    for (int q = 0; q < 9; ++q) err = fmax(err, fabs(l[q] + r[q] - whole[q]));
    if ((err < 1e-15 && depth > 2) || depth > 40) {
        for (int q = 0; q < 9; ++q) m[q] = l[q] + r[q];
        return;
    }
    double ml[9], mr[9];
    adapt(S, a, b, mode, t0, c, l, depth + 1, ml);
    adapt(S, a, b, mode, c, t1, r, depth + 1, mr);
    for (int q = 0; q < 9; ++q) m[q] = ml[q] + mr[q];
}

/* 2-D fluid moments m[p][q] = int xi^p eta^q over the fluid part of the unit
 * square mapped to [x0, x0+hx] x [y0, y0+hy]: exact in eta, adaptive
 * Gauss-Legendre in xi, split where the circle's x-extent begins and ends,
 * with the square-root behaviour there removed by substitution */
static void square_moments(double x0, double y0, double hx, double hy, double cx, double cy,
                           double r, double m[3][3])
{
    sq_t S = {x0, y0, hx, hy, cx, cy, r};
    double br[4];
    int sa[4], sb[4];
    int nb = 0;
    const double e0 = (cx - r - x0) / hx, e1 = (cx + r - x0) / hx;
    br[nb++] = 0.0;
    if (e0 > 0.0 && e0 < 1.0) br[nb++] = e0;
    if (e1 > 0.0 && e1 < 1.0) br[nb++] = e1;
    br[nb++] = 1.0;
    double acc[9] = {0};
    for (int s = 0; s + 1 < nb; ++s) {
        const double a = br[s], b = br[s + 1];
        sa[s] = (a == e0);
        sb[s] = (b == e1);
        /* pieces: [a, mid] and [mid, b] when both ends are singular */
        double pieces[3] = {a, b, b};
        int np = 1;
        if (sa[s] && sb[s]) {
            pieces[1] = 0.5 * (a + b);
            pieces[2] = b;
            np = 2;
        }
        for (int q = 0; q < np; ++q) {
            const double pa = pieces[q], pb = pieces[q + 1];
            const int mode = (q == 0 && sa[s]) ? 1 : ((q == np - 1 && sb[s]) ? 2 : 0);
            const double t0 = mode ? 0.0 : pa, t1 = mode ? 1.0 : pb;
            double w[9], v[9];
            gl8(&S, pa, pb, mode, t0, t1, w);
            adapt(&S, pa, pb, mode, t0, t1, w, 0, v);
            for (int k = 0; k < 9; ++k) acc[k] += v[k];
        }
    }
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) m[p][q] = acc[p * 3 + q];
}

/* per cell: type (0 covered, 1 regular, 2 cut) and 27 moments m[p][q][r]
 * (index p + 3q + 9r) of the fluid part of the unit cell */
void orc_eb_cylinder(const orc_geom* g, const double prob_lo[3], double cx, double cy, double r,
                     int* ctype, double* mom)
{
    const double hx = g->dx[0], hy = g->dx[1];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                const long c = IDX(i, j, k, g->n[0], g->n[1]);
                double m2[3][3];
                square_moments(prob_lo[0] + i * hx, prob_lo[1] + j * hy, hx, hy, cx, cy, r, m2);
                const double v = m2[0][0];
                int t = 2;
                if (v <= 1e-14) t = 0;
                else if (v >= 1.0 - 1e-14) t = 1;
                ctype[c] = t;
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q)
                        for (int s = 0; s < 3; ++s)
                            mom[c * 27 + p + 3 * q + 9 * s] =
                                (t == 1) ? 1.0 / ((p + 1) * (q + 1) * (s + 1))
                                         : (t == 0 ? 0.0 : m2[p][q] / (s + 1));
            }
}

/* int over the fluid part of prod_{e in dims} l_{a_e}(x_e) l_{b_e}(x_e) where
 * l_0 = 1 - x, l_1 = x; dims with a_e < 0 are left out (factor 1) */
static double poly_moment(const double* mom, const int a[3], const int b[3])
{
    /* expand each factor into monomial coefficients c_e[0..2] */
    double c[3][3];
    for (int e = 0; e < 3; ++e) {
        c[e][0] = c[e][1] = c[e][2] = 0.0;
        if (a[e] < 0) {
            c[e][0] = 1.0;
            continue;
        }
        /* (l_a)(l_b): l_0 l_0 = 1 - 2x + x^2; l_0 l_1 = x - x^2; l_1 l_1 = x^2;
         * single factor (b < 0): l_0 = 1 - x, l_1 = x */
        if (b[e] < 0) {
            if (a[e] == 0) { c[e][0] = 1.0; c[e][1] = -1.0; }
            else c[e][1] = 1.0;
        } else if (a[e] == 0 && b[e] == 0) {
            c[e][0] = 1.0; c[e][1] = -2.0; c[e][2] = 1.0;
        } else if (a[e] == 1 && b[e] == 1) {
            c[e][2] = 1.0;
        } else {
            c[e][1] = 1.0; c[e][2] = -1.0;
        }
    }
    double s = 0.0;
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < 3; ++r) {
                const double w = c[0][p] * c[1][q] * c[2][r];
                if (w != 0.0) s += w * mom[p + 3 * q + 9 * r];
            }
    return s;
}

/* FD-normalised EB element matrix of one cell (sigma not included) */
void orc_eb_element(const double* mom, const double dx[3], double E[64])
{
    for (int pa = 0; pa < 8; ++pa)
        for (int pb = 0; pb < 8; ++pb) {
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
                const int ad = (pa >> d) & 1, bd = (pb >> d) & 1;
                const double sa = ad ? 1.0 : -1.0, sb = bd ? 1.0 : -1.0;
                int a[3], b[3];
                for (int e = 0; e < 3; ++e) {
                    if (e == d) { a[e] = -1; b[e] = -1; }
                    else { a[e] = (pa >> e) & 1; b[e] = (pb >> e) & 1; }
                }
                s += sa * sb * poly_moment(mom, a, b) / (dx[d] * dx[d]);
            }
            E[pa * 8 + pb] = s;
        }
}

/* divergence weights: W[a][d] = s_a^d / h_d * int_fluid prod_{e != d} l_a^e */
static void eb_grad_weights(const double* mom, const double dx[3], double W[8][3])
{
    for (int pa = 0; pa < 8; ++pa)
        for (int d = 0; d < 3; ++d) {
            int a[3], b[3] = {-1, -1, -1};
            for (int e = 0; e < 3; ++e) a[e] = (e == d) ? -1 : ((pa >> e) & 1);
            const double sa = ((pa >> d) & 1) ? 1.0 : -1.0;
            W[pa][d] = sa / dx[d] * poly_moment(mom, a, b);
        }
}

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

/* node active: touches at least one cell with fluid, and is not Dirichlet */
void orc_eb_node_mask(const orc_geom* g, const int* ctype, int* mask)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int n[3] = {i, j, k}, act = 0;
                for (int a = 0; a < 8; ++a) {
                    int c[3], ok = 1;
                    for (int d = 0; d < 3; ++d) {
                        c[d] = cell_of(g, d, n[d], (a >> d) & 1);
                        if (c[d] < 0) ok = 0;
                    }
                    if (ok && ctype[IDX(c[0], c[1], c[2], g->n[0], g->n[1])] != 0) act = 1;
                }
                mask[IDX(i, j, k, nx, ny)] = act && !orc_is_dirichlet_node(g, i, j, k);
            }
}

/* precomputed per-cell element matrices (sigma applied) */
static double* eb_elements(const orc_geom* g, const int* ctype, const double* mom,
                           const double* sigma)
{
    const long nc = (long)g->n[0] * g->n[1] * g->n[2];
    double* K = (double*)malloc(sizeof(double) * 64 * nc);
#pragma omp parallel for
    for (long c = 0; c < nc; ++c) {
        if (ctype[c] == 0) {
            memset(K + 64 * c, 0, sizeof(double) * 64);
            continue;
        }
        orc_eb_element(mom + 27 * c, g->dx, K + 64 * c);
        for (int q = 0; q < 64; ++q) K[64 * c + q] *= sigma[c];
    }
    return K;
}

static void eb_apply_K(const orc_geom* g, const double* K, const int* mask, const double* x,
                       double* y)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long id = IDX(i, j, k, nx, ny);
                if (!mask[id]) {
                    y[id] = 0.0;
                    continue;
                }
                int n[3] = {i, j, k};
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) {
                    int cd[3], c[3], ok = 1;
                    for (int d = 0; d < 3; ++d) {
                        cd[d] = (a >> d) & 1;
                        c[d] = cell_of(g, d, n[d], cd[d]);
                        if (c[d] < 0) ok = 0;
                    }
                    if (!ok) continue;
                    const double* Kc = K + 64 * IDX(c[0], c[1], c[2], g->n[0], g->n[1]);
                    const int p = (1 - cd[0]) | ((1 - cd[1]) << 1) | ((1 - cd[2]) << 2);
                    for (int m = 0; m < 8; ++m) {
                        const int mi = node_wrap(g, 0, n[0] - 1 + cd[0] + (m & 1));
                        const int mj = node_wrap(g, 1, n[1] - 1 + cd[1] + ((m >> 1) & 1));
                        const int mk = node_wrap(g, 2, n[2] - 1 + cd[2] + ((m >> 2) & 1));
                        const long mid = IDX(mi, mj, mk, nx, ny);
                        if (!mask[mid]) continue; /* inactive / Dirichlet nodes hold 0 */
                        acc -= Kc[p * 8 + m] * x[mid];
                    }
                }
                y[id] = acc;
            }
}

void orc_eb_nodal_apply(const orc_geom* g, const int* ctype, const double* mom,
                        const double* sigma, const double* phi, double* out)
{
    double* K = eb_elements(g, ctype, mom, sigma);
    int* mask = (int*)malloc(sizeof(int) * (size_t)nn_of(g, 0) * nn_of(g, 1) * nn_of(g, 2));
    orc_eb_node_mask(g, ctype, mask);
    eb_apply_K(g, K, mask, phi, out);
    free(mask);
    free(K);
}

/* EB nodal divergence of the ghosted cell velocity (ghost rule of
 * orc_fill_vel_ghosts: ghosts are read for cells outside the domain) */
void orc_eb_nodal_divergence(const orc_geom* g, const int* ctype, const double* mom,
                             const double* ug, double* rhs)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
    const int gx = g->n[0] + 2, gy = g->n[1] + 2, gz = g->n[2] + 2;
    const long gs = (long)gx * gy * gz;
    int* mask = (int*)malloc(sizeof(int) * (size_t)nx * ny * nz);
    orc_eb_node_mask(g, ctype, mask);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long id = IDX(i, j, k, nx, ny);
                if (!mask[id]) {
                    rhs[id] = 0.0;
                    continue;
                }
                int n[3] = {i, j, k};
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) {
                    int cd[3], c[3], cg[3], inside = 1;
                    for (int d = 0; d < 3; ++d) {
                        cd[d] = (a >> d) & 1;
                        c[d] = n[d] - 1 + cd[d];
                        int cw = c[d];
                        if (g->periodic[d]) cw = ((cw % g->n[d]) + g->n[d]) % g->n[d];
                        if (cw < 0 || cw >= g->n[d]) inside = 0;
                        cg[d] = (g->periodic[d] ? cw : c[d]) + 1;
                        c[d] = cw;
                    }
                    const int p = (1 - cd[0]) | ((1 - cd[1]) << 1) | ((1 - cd[2]) << 2);
                    double W[8][3];
                    if (inside) {
                        const long cc = IDX(c[0], c[1], c[2], g->n[0], g->n[1]);
                        if (ctype[cc] == 0) continue;
                        eb_grad_weights(mom + 27 * cc, g->dx, W);
                    } else {
                        /* ghost cell outside a non-periodic face: regular */
                        double full[27];
                        for (int q = 0; q < 27; ++q)
                            full[q] = 1.0 / ((q % 3 + 1) * ((q / 3) % 3 + 1) * (q / 9 + 1));
                        eb_grad_weights(full, g->dx, W);
                    }
                    const long gc = (long)cg[0] + (long)gx * ((long)cg[1] + (long)gy * cg[2]);
                    for (int d = 0; d < 3; ++d) acc -= ug[d * gs + gc] * W[p][d];
                }
                rhs[id] = acc;
            }
    free(mask);
}

/* EB cell gradient of nodal phi (0 in covered cells) */
void orc_eb_nodal_gradient(const orc_geom* g, const int* ctype, const double* mom,
                           const double* phi, double* gphi)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1);
    const long nc = (long)g->n[0] * g->n[1] * g->n[2];
#pragma omp parallel for collapse(2)
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                const long c = IDX(i, j, k, g->n[0], g->n[1]);
                if (ctype[c] == 0) {
                    for (int d = 0; d < 3; ++d) gphi[d * nc + c] = 0.0;
                    continue;
                }
                double W[8][3];
                eb_grad_weights(mom + 27 * c, g->dx, W);
                const double v = mom[27 * c];
                double gsum[3] = {0, 0, 0};
                for (int m = 0; m < 8; ++m) {
                    const int mi = node_wrap(g, 0, i + (m & 1));
                    const int mj = node_wrap(g, 1, j + ((m >> 1) & 1));
                    const int mk = node_wrap(g, 2, k + ((m >> 2) & 1));
                    const double pv = phi[IDX(mi, mj, mk, nx, ny)];
                    for (int d = 0; d < 3; ++d) gsum[d] += W[m][d] * pv;
                }
                for (int d = 0; d < 3; ++d) gphi[d * nc + c] = gsum[d] / v;
            }
}

/* Jacobi-preconditioned conjugate gradients on -L (SPD on the active nodes;
 * the mean is removed in the singular case).  Returns 0 when converged. */
int orc_eb_nodal_solve(const orc_geom* g, const int* ctype, const double* mom,
                       const double* sigma, const double* b_in, double* x, double rtol,
                       int maxiter, int* iters)
{
    const long nn = (long)nn_of(g, 0) * nn_of(g, 1) * nn_of(g, 2);
    double* K = eb_elements(g, ctype, mom, sigma);
    int* mask = (int*)malloc(sizeof(int) * nn);
    orc_eb_node_mask(g, ctype, mask);
    double *r = calloc(nn, 8), *z = calloc(nn, 8), *p = calloc(nn, 8), *q = calloc(nn, 8),
           *dg = calloc(nn, 8);
    const int singular = orc_nodal_is_singular(g);
    long nact = 0;
    double bmean = 0.0;
    for (long i = 0; i < nn; ++i)
        if (mask[i]) { nact++; bmean += b_in[i]; }
    bmean = singular ? bmean / (double)(nact ? nact : 1) : 0.0;
    {
        /* diagonal of -L, element by element */
        const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const long id = IDX(i, j, k, nx, ny);
                    if (!mask[id]) continue;
                    int n[3] = {i, j, k};
                    double acc = 0.0;
                    for (int a = 0; a < 8; ++a) {
                        int cd[3], c[3], ok = 1;
                        for (int d = 0; d < 3; ++d) {
                            cd[d] = (a >> d) & 1;
                            c[d] = cell_of(g, d, n[d], cd[d]);
                            if (c[d] < 0) ok = 0;
                        }
                        if (!ok) continue;
                        const int pp = (1 - cd[0]) | ((1 - cd[1]) << 1) | ((1 - cd[2]) << 2);
                        acc += K[64 * IDX(c[0], c[1], c[2], g->n[0], g->n[1]) + pp * 9];
                    }
                    dg[id] = acc;
                }
    }
    for (long i = 0; i < nn; ++i) {
        x[i] = mask[i] ? x[i] : 0.0;
    }
    eb_apply_K(g, K, mask, x, q);
    double bn = 0.0;
    for (long i = 0; i < nn; ++i) {
        const double bi = mask[i] ? b_in[i] - bmean : 0.0;
        r[i] = mask[i] ? -(bi - q[i]) : 0.0; /* residual of (-L) x = -b */
        bn = fmax(bn, fabs(bi));
    }
    double rz = 0.0;
    for (long i = 0; i < nn; ++i) {
        z[i] = (mask[i] && dg[i] > 0) ? r[i] / dg[i] : 0.0;
        p[i] = z[i];
        rz += r[i] * z[i];
    }
    int it = 0, conv = 0;
    for (it = 0; it < maxiter; ++it) {
        double rn = 0.0;
        for (long i = 0; i < nn; ++i) rn = fmax(rn, fabs(r[i]));
        if (rn <= rtol * bn) { conv = 1; break; }
        eb_apply_K(g, K, mask, p, q); /* q = L p */
        double pq = 0.0;
        for (long i = 0; i < nn; ++i) pq += p[i] * (-q[i]);
        if (pq == 0.0) break;
        const double alpha = rz / pq;
        for (long i = 0; i < nn; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * (-q[i]);
        }
        if (singular) {
            double rm = 0.0;
            for (long i = 0; i < nn; ++i) rm += mask[i] ? r[i] : 0.0;
            rm /= (double)(nact ? nact : 1);
            for (long i = 0; i < nn; ++i) if (mask[i]) r[i] -= rm;
        }
        double rz2 = 0.0;
        for (long i = 0; i < nn; ++i) {
            z[i] = (mask[i] && dg[i] > 0) ? r[i] / dg[i] : 0.0;
            rz2 += r[i] * z[i];
        }
        const double beta = rz2 / rz;
        rz = rz2;
        for (long i = 0; i < nn; ++i) p[i] = z[i] + beta * p[i];
    }
    if (singular) {
        double xm = 0.0;
        for (long i = 0; i < nn; ++i) xm += mask[i] ? x[i] : 0.0;
        xm /= (double)(nact ? nact : 1);
        for (long i = 0; i < nn; ++i) x[i] = mask[i] ? x[i] - xm : 0.0;
    }
    if (iters) *iters = it;
    free(r); free(z); free(p); free(q); free(dg); free(mask); free(K);
    return conv ? 0 : 1;
}

/* ---- EB inflow velocity (setEBInflowVelocity) ---------------------------
 * With a velocity u_eb on the body surface the weak divergence gains the
 * surface term: b_n += sum_cells (1/V) int_{body surface in cell} N_n (u_eb . n) dA,
 * n pointing out of the fluid (src/projection/incflo_apply_nodal_projection.cpp:196-200).
 * The surface weights B[p][d] = (1/V) int N_p n_d dA follow from the divergence
 * theorem on the fluid part of the cell: the volume integral of d N_p / d x_d
 * (the weights W above, from the fluid moments) minus the integrals over the
 * fluid parts of the cell faces, which for the z-cylinder are exact segment
 * integrals (x- and y-faces) and the 2-D fluid moments (z-faces).  This is
 * independent of the product's arc quadrature. */

/* int over the fluid part of the horizontal line eta = const of xi^q (exact) */
static void row_moments(const sq_t* S, double eta, double f[3])
{
    const double y = S->y0 + S->hy * eta;
    const double dy = y - S->cy;
    double lo = 1.0, hi = 0.0;
    if (dy * dy < S->r * S->r) {
        const double h = sqrt(S->r * S->r - dy * dy);
        lo = fmax((S->cx - h - S->x0) / S->hx, 0.0);
        hi = fmin((S->cx + h - S->x0) / S->hx, 1.0);
    }
    for (int q = 0; q < 3; ++q) {
        double v = 1.0 / (q + 1);
        if (hi > lo) v -= (pow(hi, q + 1) - pow(lo, q + 1)) / (q + 1);
        f[q] = v;
    }
}

void orc_eb_cylinder_boundary(const orc_geom* g, const double prob_lo[3], double cx, double cy,
                              double r, const int* ctype, const double* mom, double* B)
{
    const long nc = (long)g->n[0] * g->n[1] * g->n[2];
    for (long c = 0; c < nc; ++c) {
        double* b = B + 24 * c;
        for (int q = 0; q < 24; ++q) b[q] = 0.0;
        if (ctype[c] != 2) continue;
        const int i = (int)(c % g->n[0]), j = (int)((c / g->n[0]) % g->n[1]);
        sq_t S = {prob_lo[0] + i * g->dx[0], prob_lo[1] + j * g->dx[1], g->dx[0], g->dx[1], cx, cy, r};
        double W[8][3];
        eb_grad_weights(mom + 27 * c, g->dx, W);
        double fx[2][3], fy[2][3];  /* fluid moments along the x-faces (in eta) and y-faces (in xi) */
        line_moments(&S, 0.0, fx[0]);
        line_moments(&S, 1.0, fx[1]);
        row_moments(&S, 0.0, fy[0]);
        row_moments(&S, 1.0, fy[1]);
        const double* m = mom + 27 * c;
        for (int p = 0; p < 8; ++p) {
            const int bx = p & 1, by = (p >> 1) & 1, bz = (p >> 2) & 1;
            /* x: the x-face the corner lies on, int N_y over its fluid segment, int N_z = 1/2 */
            const double Ny = by ? fx[bx][1] : fx[bx][0] - fx[bx][1];
            b[3 * p + 0] = W[p][0] - (bx ? 1.0 : -1.0) / g->dx[0] * Ny * 0.5;
            const double Nx = bx ? fy[by][1] : fy[by][0] - fy[by][1];
            b[3 * p + 1] = W[p][1] - (by ? 1.0 : -1.0) / g->dx[1] * Nx * 0.5;
            /* z: the 2-D fluid moment of N_x N_y (z-extruded fluid) */
            const double a00 = m[0], a10 = m[1], a01 = m[3], a11 = m[4];
            const double nxny = bx ? (by ? a11 : a10 - a11) : (by ? a01 - a11 : a00 - a10 - a01 + a11);
            b[3 * p + 2] = W[p][2] - (bz ? 1.0 : -1.0) / g->dx[2] * nxny;
        }
    }
}

/* rhs += sum over the cut cells around each active node of B[p] . u_eb
 * (ueb: 3 x cells, x fastest) */
void orc_eb_inflow_rhs(const orc_geom* g, const int* ctype, const double* B, const double* ueb,
                       double* rhs)
{
    const int nx = nn_of(g, 0), ny = nn_of(g, 1), nz = nn_of(g, 2);
    const long nc = (long)g->n[0] * g->n[1] * g->n[2];
    int* mask = (int*)malloc(sizeof(int) * (size_t)nx * ny * nz);
    orc_eb_node_mask(g, ctype, mask);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long id = IDX(i, j, k, nx, ny);
                if (!mask[id]) continue;
                int n[3] = {i, j, k};
                double acc = 0.0;
                for (int a = 0; a < 8; ++a) {
                    int c[3], inside = 1;
                    for (int d = 0; d < 3; ++d) {
                        int cw = n[d] - 1 + ((a >> d) & 1);
                        if (g->periodic[d]) cw = ((cw % g->n[d]) + g->n[d]) % g->n[d];
                        if (cw < 0 || cw >= g->n[d]) inside = 0;
                        c[d] = cw;
                    }
                    if (!inside) continue;
                    const long cc = IDX(c[0], c[1], c[2], g->n[0], g->n[1]);
                    if (ctype[cc] != 2) continue;
                    const int p = (1 - (a & 1)) | ((1 - ((a >> 1) & 1)) << 1) | ((1 - ((a >> 2) & 1)) << 2);
                    for (int d = 0; d < 3; ++d) acc += B[24 * cc + 3 * p + d] * ueb[d * nc + cc];
                }
                rhs[id] += acc;
            }
    free(mask);
}

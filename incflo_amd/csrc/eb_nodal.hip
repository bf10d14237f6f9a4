// eb_nodal.hip -- host side of the embedded-boundary nodal multigrid: cylinder
// geometry from boundary integrals, per-level cut-cell tables, Galerkin
// coarsening of cut elements, V-cycle and the FP64 MLMG outer loop (same
// stopping rule as the regular solver).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "eb_nodal.h"
#include "mgx_ctx.h"

namespace mgx {

// ---------------------------------------------------------------------------
// geometry

namespace {

const double kGL16X[16] = {-0.9894009349916499, -0.9445750230732326, -0.8656312023878318,
                           -0.7554044083550030, -0.6178762444026438, -0.4580167776572274,
                           -0.2816035507792589, -0.0950125098376374, 0.0950125098376374,
                           0.2816035507792589,  0.4580167776572274,  0.6178762444026438,
                           0.7554044083550030,  0.8656312023878318,  0.9445750230732326,
                           0.9894009349916499};
const double kGL16W[16] = {0.0271524594117541, 0.0622535239386479, 0.0951585116824928,
                           0.1246289712555339, 0.1495959888165767, 0.1691565193950025,
                           0.1826034150449236, 0.1894506104550685, 0.1894506104550685,
                           0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                           0.1246289712555339, 0.0951585116824928, 0.0622535239386479,
                           0.0271524594117541};

// Calls f(t, w) for the Gauss-Legendre points (angle t, weight w in angle) of
// the arcs of the circle that lie inside the cell [x0,x0+hx]x[y0,y0+hy]
// (counter-clockwise, pieces of at most pi/8, 16 points each).
template <class F>
void for_arc_points(double x0, double y0, double hx, double hy, double cx, double cy, double r,
                    F&& f) {
    const double x1 = x0 + hx, y1 = y0 + hy;
    const double twopi = 2.0 * M_PI;
    auto norm = [&](double t) {
        while (t < 0.0) t += twopi;
        while (t >= twopi) t -= twopi;
        return t;
    };
    std::vector<double> th;
    auto at_x = [&](double xe) {
        const double d = (xe - cx) / r;
        if (std::fabs(d) >= 1.0) return;
        const double a = std::acos(d);
        for (double t : {a, -a}) {
            const double y = cy + r * std::sin(t);
            if (y >= y0 && y <= y1) th.push_back(norm(t));
        }
    };
    auto at_y = [&](double ye) {
        const double d = (ye - cy) / r;
        if (std::fabs(d) >= 1.0) return;
        const double a = std::asin(d);
        for (double t : {a, M_PI - a}) {
            const double x = cx + r * std::cos(t);
            if (x >= x0 && x <= x1) th.push_back(norm(t));
        }
    };
    at_x(x0);
    at_x(x1);
    at_y(y0);
    at_y(y1);
    std::sort(th.begin(), th.end());
    th.erase(std::unique(th.begin(), th.end(), [](double a, double b) { return b - a < 1e-14; }),
             th.end());
    auto inside = [&](double t) {
        const double x = cx + r * std::cos(t), y = cy + r * std::sin(t);
        return x >= x0 && x <= x1 && y >= y0 && y <= y1;
    };
    auto arc = [&](double ta, double tb) {
        // split into pieces of at most pi/8 for the quadrature
        const int np = std::max(1, (int)std::ceil((tb - ta) / (M_PI / 8)));
        for (int s = 0; s < np; ++s) {
            const double a = ta + (tb - ta) * s / np, b = ta + (tb - ta) * (s + 1) / np;
            for (int g = 0; g < 16; ++g)
                f(0.5 * (a + b) + 0.5 * (b - a) * kGL16X[g], 0.5 * (b - a) * kGL16W[g]);
        }
    };
    if (th.empty()) {
        if (inside(0.0)) arc(0.0, twopi);  // the whole circle lies in the cell
    } else {
        for (size_t s = 0; s < th.size(); ++s) {
            const double ta = th[s], tb = (s + 1 < th.size()) ? th[s + 1] : th[0] + twopi;
            if (tb - ta < 1e-15) continue;
            if (inside(0.5 * (ta + tb))) arc(ta, tb);
        }
    }
}

// moments (unit cell) of the part of the cell [x0,x0+hx]x[y0,y0+hy] inside the
// disk: Green's theorem with Q = hx X^{p+1} Y^q / (p+1), counter-clockwise
// boundary = arcs of the circle inside the cell + the cell's right edge inside
// the disk (the left edge has X = 0, horizontal edges dy = 0).
void covered_moments(double x0, double y0, double hx, double hy, double cx, double cy, double r,
                     double Mc[9]) {
    for (int q = 0; q < 9; ++q) Mc[q] = 0.0;
    const double x1 = x0 + hx, y1 = y0 + hy;
    for_arc_points(x0, y0, hx, hy, cx, cy, r, [&](double t, double w) {
        const double X = (cx + r * std::cos(t) - x0) / hx, Y = (cy + r * std::sin(t) - y0) / hy;
        const double dy = r * std::cos(t);
        double Xp = X;  // X^{p+1}
        for (int p = 0; p < 3; ++p) {
            double Yq = 1.0;
            for (int q = 0; q < 3; ++q) {
                Mc[p * 3 + q] += w * hx * Xp * Yq / (p + 1) * dy;
                Yq *= Y;
            }
            Xp *= X;
        }
    });
    // right edge x = x1, upwards, the part inside the disk
    const double dxe = x1 - cx;
    if (std::fabs(dxe) < r) {
        const double s = std::sqrt(r * r - dxe * dxe);
        const double ya = std::max(cy - s, y0), yb = std::min(cy + s, y1);
        if (yb > ya) {
            const double Ya = (ya - y0) / hy, Yb = (yb - y0) / hy;
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q)
                    Mc[p * 3 + q] += hx / (p + 1) * hy * (std::pow(Yb, q + 1) - std::pow(Ya, q + 1)) / (q + 1);
        }
    }
    for (int q = 0; q < 9; ++q) Mc[q] /= hx * hy;
}

// (1/V) int over the body surface in the cell of N_p n dA, n out of the fluid
// (toward the axis), corners p (x bit 0, y bit 1, z bit 2), [p][d]: the
// z-factor of N_p integrates to 1/2 over the cell height and n_z = 0
void surface_weights(double x0, double y0, double hx, double hy, double cx, double cy, double r,
                     double B[24]) {
    double B2[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for_arc_points(x0, y0, hx, hy, cx, cy, r, [&](double t, double w) {
        const double X = (cx + r * std::cos(t) - x0) / hx, Y = (cy + r * std::sin(t) - y0) / hy;
        const double nx = -std::cos(t), ny = -std::sin(t), ds = r * w;
        for (int c = 0; c < 4; ++c) {
            const double N = ((c & 1) ? X : 1.0 - X) * ((c & 2) ? Y : 1.0 - Y);
            B2[c][0] += N * nx * ds;
            B2[c][1] += N * ny * ds;
        }
    });
    for (int p = 0; p < 8; ++p) {
        B[3 * p + 0] = 0.5 * B2[p & 3][0] / (hx * hy);
        B[3 * p + 1] = 0.5 * B2[p & 3][1] / (hx * hy);
        B[3 * p + 2] = 0.0;
    }
}

// integral over the fluid of prod_e f_e(x_e), each f_e a polynomial with
// coefficients c[e][0..2]
double poly_int(const double* mom, const double c[3][3]) {
    double s = 0.0;
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < 3; ++r) {
                const double w = c[0][p] * c[1][q] * c[2][r];
                if (w != 0.0) s += w * mom[p + 3 * q + 9 * r];
            }
    return s;
}

// coefficients of l_a * l_b (or l_a alone, or 1) along one direction
void shape_poly(int a, int b, double c[3]) {
    c[0] = c[1] = c[2] = 0.0;
    if (a < 0) { c[0] = 1.0; return; }
    if (b < 0) {
        if (a == 0) { c[0] = 1.0; c[1] = -1.0; }
        else c[1] = 1.0;
        return;
    }
    if (a == 0 && b == 0) { c[0] = 1.0; c[1] = -2.0; c[2] = 1.0; }
    else if (a == 1 && b == 1) c[2] = 1.0;
    else { c[1] = 1.0; c[2] = -1.0; }
}

void div_weights(const double* mom, const double dx[3], double W[24]) {
    for (int p = 0; p < 8; ++p)
        for (int d = 0; d < 3; ++d) {
            double c[3][3];
            for (int e = 0; e < 3; ++e) shape_poly(e == d ? -1 : ((p >> e) & 1), -1, c[e]);
            W[3 * p + d] = (((p >> d) & 1) ? 1.0 : -1.0) / dx[d] * poly_int(mom, c);
        }
}

void full_moments(double mom[27]) {
    for (int q = 0; q < 27; ++q) mom[q] = 1.0 / ((q % 3 + 1) * ((q / 3) % 3 + 1) * (q / 9 + 1));
}

}  // namespace

void eb_cylinder_moments(const NodalGeom& g, const double prob_lo[3], double cx, double cy,
                         double r, std::vector<int>& type, std::vector<double>& mom) {
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    type.assign(nc, 1);
    mom.assign(27 * nc, 0.0);
    const double hx = g.dx[0], hy = g.dx[1];
    for (int j = 0; j < g.n[1]; ++j)
        for (int i = 0; i < g.n[0]; ++i) {
            const double x0 = prob_lo[0] + i * hx, y0 = prob_lo[1] + j * hy;
            // nearest and farthest points of the cell from the axis
            const double px = std::min(std::max(cx, x0), x0 + hx), py = std::min(std::max(cy, y0), y0 + hy);
            const double dmin = (px - cx) * (px - cx) + (py - cy) * (py - cy);
            double dmax = 0.0;
            for (int q = 0; q < 4; ++q) {
                const double x = x0 + hx * (q & 1), y = y0 + hy * (q >> 1);
                dmax = std::max(dmax, (x - cx) * (x - cx) + (y - cy) * (y - cy));
            }
            int t;
            double m2[9];
            if (dmin >= r * r) {
                t = 1;
            } else if (dmax <= r * r) {
                t = 0;
            } else {
                double Mc[9];
                covered_moments(x0, y0, hx, hy, cx, cy, r, Mc);
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) m2[p * 3 + q] = 1.0 / ((p + 1) * (q + 1)) - Mc[p * 3 + q];
                t = (m2[0] <= 1e-14) ? 0 : (m2[0] >= 1.0 - 1e-14 ? 1 : 2);
            }
            for (int k = 0; k < g.n[2]; ++k) {
                const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
                type[c] = t;
                double* m = &mom[27 * c];
                if (t == 1) full_moments(m);
                else if (t == 2)
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q)
                            for (int s = 0; s < 3; ++s) m[p + 3 * q + 9 * s] = m2[p * 3 + q] / (s + 1);
            }
        }
}

void eb_cylinder_surface_weights(const NodalGeom& g, const double prob_lo[3], double cx,
                                 double cy, double r, const std::vector<int>& type,
                                 std::vector<double>& B) {
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    B.assign(24 * nc, 0.0);
    for (int j = 0; j < g.n[1]; ++j)
        for (int i = 0; i < g.n[0]; ++i) {
            if (type[i + (long)g.n[0] * j] != 2) continue;  // the type is z-independent
            double b[24];
            surface_weights(prob_lo[0] + i * g.dx[0], prob_lo[1] + j * g.dx[1], g.dx[0], g.dx[1], cx,
                            cy, r, b);
            for (int k = 0; k < g.n[2]; ++k) {
                const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
                std::copy(b, b + 24, &B[24 * c]);
            }
        }
}

void eb_cylinder_apertures(const int n[3], const double prob_lo[3], const double dx[3], double cx,
                           double cy, double r, std::vector<double> ap[3],
                           std::vector<unsigned char>& covered) {
    // fraction of the segment [y0, y0 + h] (at abscissa x) outside the disk
    auto seg = [&](double x, double c_par, double c_perp, double y0, double h) {
        const double d = x - c_perp;
        if (std::fabs(d) >= r) return 1.0;
        const double s = std::sqrt(r * r - d * d);
        const double lo = std::max(c_par - s, y0), hi = std::min(c_par + s, y0 + h);
        return hi > lo ? 1.0 - (hi - lo) / h : 1.0;
    };
    NodalGeom g{};
    for (int d = 0; d < 3; ++d) {
        g.n[d] = n[d];
        g.dx[d] = dx[d];
    }
    std::vector<int> type;
    std::vector<double> mom;
    eb_cylinder_moments(g, prob_lo, cx, cy, r, type, mom);
    const long nc = (long)n[0] * n[1] * n[2];
    covered.assign(nc, 0);
    for (long c = 0; c < nc; ++c) covered[c] = type[c] == 0;
    for (int d = 0; d < 3; ++d) {
        const int f0 = n[0] + (d == 0), f1 = n[1] + (d == 1), f2 = n[2] + (d == 2);
        ap[d].assign((size_t)f0 * f1 * f2, 1.0);
        for (int k = 0; k < f2; ++k)
            for (int j = 0; j < f1; ++j)
                for (int i = 0; i < f0; ++i) {
                    double a;
                    if (d == 0) a = seg(prob_lo[0] + i * dx[0], cy, cx, prob_lo[1] + j * dx[1], dx[1]);
                    else if (d == 1) a = seg(prob_lo[1] + j * dx[1], cx, cy, prob_lo[0] + i * dx[0], dx[0]);
                    else a = mom[27 * (i + (long)n[0] * (j + (long)n[1] * std::min(k, n[2] - 1)))];
                    ap[d][i + (long)f0 * (j + (long)f1 * k)] = a;
                }
    }
}

void eb_element_from_moments(const double* mom, const double dx[3], double E[64]) {
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
                double c[3][3];
                for (int e = 0; e < 3; ++e)
                    shape_poly(e == d ? -1 : ((a >> e) & 1), e == d ? -1 : ((b >> e) & 1), c[e]);
                const double sa = ((a >> d) & 1) ? 1.0 : -1.0, sb = ((b >> d) & 1) ? 1.0 : -1.0;
                s += sa * sb * poly_int(mom, c) / (dx[d] * dx[d]);
            }
            E[a * 8 + b] = s;
        }
}

void eb_regular_element(const double dx[3], double E[64]) {
    double m[27];
    full_moments(m);
    eb_element_from_moments(m, dx, E);
}

// ---------------------------------------------------------------------------
// device helpers local to the EB solver

namespace {

__global__ void eb_subtract_masked_kernel(const unsigned char* __restrict__ mask, double* __restrict__ a,
                                          long n, const double* __restrict__ s) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n && mask[id]) a[id] -= s[0];
}

// one-workgroup BiCGStab (van der Vorst 1992) on the coarsest EB level, FP64,
// vectors in global memory; inactive nodes hold 0; mean removed when singular
// The EB nodal operator in FP64 for the bottom solve: inactive nodes are held
// at zero and the singular mean is over the active ones; its smoother is the
// V-cycle's weighted Jacobi sweep.
struct EBBottomOp {
    EBLev L;
    double nact;
    double omega;
    __device__ long size() const { return (long)L.g.nn[0] * L.g.nn[1] * L.g.nn[2]; }
    __device__ bool held(long id) const { return !L.mask[id]; }
    __device__ double nfree() const { return nact; }
    __device__ void node(long id, int& i, int& j, int& k) const {
        i = (int)(id % L.g.nn[0]);
        j = (int)((id / L.g.nn[0]) % L.g.nn[1]);
        k = (int)(id / ((long)L.g.nn[0] * L.g.nn[1]));
    }
    __device__ void apply(const double* in, double* out, long n) const {
        __syncthreads();
        for (long id = threadIdx.x; id < n; id += blockDim.x) {
            int i, j, k;
            node(id, i, j, k);
            double dg;
            out[id] = L.mask[id] ? eb_apply_node(L, in, nullptr, i, j, k, dg) : 0.0;
        }
        __syncthreads();
    }
    __device__ void smooth(const BottomVecs& V, long n) const {
        __syncthreads();
        for (long id = threadIdx.x; id < n; id += blockDim.x) {
            int i, j, k;
            node(id, i, j, k);
            double u = 0.0;
            if (L.mask[id]) {
                double dg;
                const double lx = eb_apply_node(L, V.x, nullptr, i, j, k, dg);
                if (dg != 0.0) u = V.x[id] + omega * (V.b[id] - lx) / dg;
            }
            V.t[id] = u;
        }
        __syncthreads();
        for (long id = threadIdx.x; id < n; id += blockDim.x) V.x[id] = V.t[id];
        __syncthreads();
    }
};

__global__ void __launch_bounds__(1024) eb_bottom_kernel(EBLev L, const double* __restrict__ b_in,
                                                         double* __restrict__ x_out, double* work,
                                                         int kind, int maxiter, double rtol,
                                                         double omega, int singular, double nact) {
    __shared__ double red[16];
    const EBBottomOp op{L, nact, omega};
    bottom_solve_body(op, b_in, x_out, work, kind, maxiter, rtol, singular, nullptr, red);
}

void node_mask(const NodalGeom& g, const std::vector<int>& cidx, std::vector<unsigned char>& m) {
    m.assign(g.nnode(), 0);
    for (int k = 0; k < g.nn[2]; ++k)
        for (int j = 0; j < g.nn[1]; ++j)
            for (int i = 0; i < g.nn[0]; ++i) {
                const int n3[3] = {i, j, k};
                bool dir = false, act = false;
                for (int d = 0; d < 3; ++d) {
                    if (!g.per[d] && n3[d] == 0 && g.dlo[d]) dir = true;
                    if (!g.per[d] && n3[d] == g.n[d] && g.dhi[d]) dir = true;
                }
                for (int a = 0; a < 8 && !act; ++a) {
                    int c[3];
                    bool ok = true;
                    for (int d = 0; d < 3; ++d) {
                        int cc = n3[d] - 1 + ((a >> d) & 1);
                        if (g.per[d]) cc = (cc + g.n[d]) % g.n[d];
                        else if (cc < 0 || cc >= g.n[d]) ok = false;
                        c[d] = cc;
                    }
                    if (ok && cidx[c[0] + (long)g.n[0] * (c[1] + (long)g.n[1] * c[2])] != -2) act = true;
                }
                m[i + (long)g.nn[0] * (j + (long)g.nn[1] * k)] = (act && !dir) ? 1 : 0;
            }
}

template <class T>
void upload(DevBuf& buf, const std::vector<T>& h) {
    buf.alloc(std::max<size_t>(1, h.size()) * sizeof(T));
    if (!h.empty()) MGX_HIP(hipMemcpy(buf.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace

void EBNodal::setup(const NodalGeom& g0, int nlev, const std::vector<int>& type,
                    const std::vector<double>& mom, hipStream_t s) {
    (void)s;
    L.clear();
    L.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        EBLevelData& lv = L[l];
        lv.g = (l == 0) ? g0 : coarsen(L[l - 1].g);
        eb_regular_element(lv.g.dx, lv.E);
        const long nc = (long)lv.g.n[0] * lv.g.n[1] * lv.g.n[2];
        lv.cidx_h.assign(nc, -1);
        lv.cut_cells.clear();
        lv.children.clear();
        if (l == 0) {
            std::vector<double> Eg, W, V;
            for (long c = 0; c < nc; ++c) {
                if (type[c] == 0) { lv.cidx_h[c] = -2; continue; }
                if (type[c] == 1) continue;
                lv.cidx_h[c] = (int)lv.cut_cells.size();
                lv.cut_cells.push_back((int)c);
                double E[64], w[24];
                eb_element_from_moments(&mom[27 * c], lv.g.dx, E);
                div_weights(&mom[27 * c], lv.g.dx, w);
                Eg.insert(Eg.end(), E, E + 64);
                W.insert(W.end(), w, w + 24);
                V.push_back(mom[27 * c]);
            }
            upload(lv.Egeom, Eg);
            upload(lv.W, W);
            upload(lv.V, V);
        } else {
            const EBLevelData& f = L[l - 1];
            for (int k = 0; k < lv.g.n[2]; ++k)
                for (int j = 0; j < lv.g.n[1]; ++j)
                    for (int i = 0; i < lv.g.n[0]; ++i) {
                        int ch[8], ncov = 0, nreg = 0;
                        for (int q = 0; q < 8; ++q) {
                            const int fi = 2 * i + (q & 1), fj = 2 * j + ((q >> 1) & 1), fk = 2 * k + ((q >> 2) & 1);
                            ch[q] = fi + f.g.n[0] * (fj + f.g.n[1] * fk);
                            const int t = f.cidx_h[ch[q]];
                            ncov += (t == -2);
                            nreg += (t == -1);
                        }
                        const long c = i + (long)lv.g.n[0] * (j + (long)lv.g.n[1] * k);
                        if (ncov == 8) lv.cidx_h[c] = -2;
                        else if (nreg == 8) lv.cidx_h[c] = -1;
                        else {
                            lv.cidx_h[c] = (int)lv.cut_cells.size();
                            lv.cut_cells.push_back((int)c);
                            lv.children.insert(lv.children.end(), ch, ch + 8);
                        }
                    }
            upload(lv.child, lv.children);
        }
        lv.ncut = (int)lv.cut_cells.size();
        upload(lv.cidx, lv.cidx_h);
        upload(lv.cutc, lv.cut_cells);
        node_mask(lv.g, lv.cidx_h, lv.mask_h);
        upload(lv.mask, lv.mask_h);
        lv.K.alloc(std::max(1, lv.ncut) * 64 * sizeof(double));
        const long nn = lv.g.nnode();
        if (l > 0) {
            lv.sig_buf.alloc(lv.g.nsig() * sizeof(double));
            MGX_HIP(hipMemset(lv.sig_buf.p, 0, lv.g.nsig() * sizeof(double)));
        }
        lv.b_buf.alloc(nn * sizeof(double));
        lv.x_buf.alloc(nn * sizeof(double));
        lv.x2_buf.alloc(nn * sizeof(double));
        lv.r_buf.alloc(nn * sizeof(double));
        MGX_HIP(hipMemset(lv.x_buf.p, 0, nn * sizeof(double)));
        MGX_HIP(hipMemset(lv.x2_buf.p, 0, nn * sizeof(double)));
    }
    bottom_work.alloc(7 * L.back().g.nnode() * sizeof(double));
    ready = true;
}

EBLev EBNodal::lev(int l) const {
    const EBLevelData& lv = L[l];
    EBLev e{};
    e.g = kgeom(lv.g);
    std::memcpy(e.E, lv.E, sizeof(e.E));
    e.sig = (l == 0) ? sig0 : lv.sig_buf.as<double>();
    e.cidx = lv.cidx.as<int>();
    e.K = lv.K.as<double>();
    e.mask = lv.mask.as<unsigned char>();
    return e;
}

void EBNodal::set_sigma(const double* s0, hipStream_t s) {
    sig0 = s0;
    for (size_t l = 1; l < L.size(); ++l)
        sigma_coarsen_kernel<double><<<grid3(L[l].g.n[0], L[l].g.n[1], L[l].g.n[2]), block3(), 0, s>>>(
            kgeom(L[l - 1].g), kgeom(L[l].g), (l == 1) ? sig0 : L[l - 1].sig_buf.as<double>(),
            L[l].sig_buf.as<double>());
    if (L[0].ncut > 0)
        eb_scale_kernel<<<(64 * L[0].ncut + 255) / 256, 256, 0, s>>>(
            kgeom(L[0].g), L[0].cutc.as<int>(), L[0].ncut, L[0].Egeom.as<double>(), sig0,
            L[0].K.as<double>());
    for (size_t l = 1; l < L.size(); ++l)
        if (L[l].ncut > 0)
            eb_galerkin_kernel<<<L[l].ncut, 64, 0, s>>>(lev((int)l - 1), L[l].child.as<int>(),
                                                        L[l].ncut, L[l].K.as<double>());
}

void EBNodal::smooth(int l, int sweeps, bool from_zero, const MGOpts& o, hipStream_t s) {
    EBLevelData& lv = L[l];
    const long nn = lv.g.nnode();
    double* x = lv.x_buf.as<double>();
    double* x2 = lv.x2_buf.as<double>();
    if (from_zero) zero_kernel<double><<<grid1d(nn), 256, 0, s>>>(x, nn);
    for (int q = 0; q < sweeps; ++q) {
        EBArgs a{};
        a.L = lev(l);
        a.x = x;
        a.b = lv.b_buf.as<double>();
        a.xout = x2;
        a.omega = nodal_omega(lv.g, o.omega);
        eb_stencil_kernel<<<grid3(lv.g.nn[0], lv.g.nn[1], lv.g.nn[2]), block3(), 0, s>>>(a, EB_JACOBI);
        std::swap(x, x2);
    }
    // keep the result in x_buf
    if (x != lv.x_buf.as<double>())
        MGX_HIP(hipMemcpyAsync(lv.x_buf.p, x, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void EBNodal::vcycle(const MGOpts& o, hipStream_t s) {
    const int nl = (int)L.size();
    for (int l = 0; l < nl - 1; ++l) {
        EBLevelData& lv = L[l];
        smooth(l, o.nu1, true, o, s);
        EBArgs a{};
        a.L = lev(l);
        a.x = lv.x_buf.as<double>();
        a.b = lv.b_buf.as<double>();
        a.rout = lv.r_buf.as<double>();
        eb_stencil_kernel<<<grid3(lv.g.nn[0], lv.g.nn[1], lv.g.nn[2]), block3(), 0, s>>>(a, EB_RESID);
        EBLevelData& lc = L[l + 1];
        nodal_restrict_kernel<double><<<grid3(lc.g.nn[0], lc.g.nn[1], lc.g.nn[2]), block3(), 0, s>>>(
            kgeom(lv.g), kgeom(lc.g), lv.r_buf.as<double>(), lc.b_buf.as<double>());
        eb_mask_kernel<<<grid1d(lc.g.nnode()), 256, 0, s>>>(lc.mask.as<unsigned char>(),
                                                           lc.b_buf.as<double>(), lc.g.nnode());
    }
    {
        EBLevelData& lb = L[nl - 1];
        long nact = 0;
        for (unsigned char m : lb.mask_h) nact += m;
        const long nb = lb.g.nnode();
        const int threads = nb >= 1024 ? 1024 : (nb >= 256 ? 256 : 64);
        eb_bottom_kernel<<<1, threads, 0, s>>>(lev(nl - 1), lb.b_buf.as<double>(), lb.x_buf.as<double>(),
                                               bottom_work.as<double>(), o.bottom_solver,
                                               o.bottom_maxiter, o.bottom_rtol, nodal_omega(lb.g, o.omega),
                                               lb.g.singular() ? 1 : 0, (double)nact);
    }
    for (int l = nl - 2; l >= 0; --l) {
        EBLevelData& lv = L[l];
        nodal_prolong_add_kernel<double><<<grid3(lv.g.nn[0], lv.g.nn[1], lv.g.nn[2]), block3(), 0, s>>>(
            kgeom(lv.g), kgeom(L[l + 1].g), L[l + 1].x_buf.as<double>(), lv.x_buf.as<double>());
        eb_mask_kernel<<<grid1d(lv.g.nnode()), 256, 0, s>>>(lv.mask.as<unsigned char>(),
                                                           lv.x_buf.as<double>(), lv.g.nnode());
        smooth(l, o.nu2, false, o, s);
    }
}

static double read_u64_norm(unsigned long long* dev, unsigned long long* h, hipStream_t s) {
    MGX_HIP(hipMemcpyAsync(h, dev, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    double d;
    std::memcpy(&d, h, sizeof(d));
    return d;
}

void EBNodal::solve(double* b, double* phi_a, double* phi_b, double*& phi, double rtol,
                    double atol, const MGOpts& o, DevBuf& scal, DevBuf& part,
                    unsigned long long* h_pinned, mgx_solve_info* info, hipStream_t s) {
    EBLevelData& l0 = L[0];
    const NodalGeom& g = l0.g;
    const long nn = g.nnode();
    const unsigned char* mask = l0.mask.as<unsigned char>();
    long nact = 0;
    for (unsigned char m : l0.mask_h) nact += m;
    unsigned long long* nrm = scal.as<unsigned long long>();
    // b: inactive rows 0, mean over active nodes removed when singular
    eb_mask_kernel<<<grid1d(nn), 256, 0, s>>>(mask, b, nn);
    eb_mask_kernel<<<grid1d(nn), 256, 0, s>>>(mask, phi, nn);
    if (g.singular() && nact > 0) {
        const int np = (int)std::min<long>(1024, std::max<long>(1, (nn + 255) / 256));
        partial_sum_kernel<double><<<np, 256, 0, s>>>(b, nn, part.as<double>());
        final_sum_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, scal.as<double>() + 33,
                                           1.0 / (double)nact);
        eb_subtract_masked_kernel<<<grid1d(nn), 256, 0, s>>>(mask, b, nn, scal.as<double>() + 33);
    }
    zero_u64_kernel<<<1, 64, 0, s>>>(nrm, 1);
    max_abs_kernel<double><<<(int)std::min<long>(1024, (nn + 255) / 256), 256, 0, s>>>(b, nn, nrm);
    const double bnorm = read_u64_norm(nrm, h_pinned, s);
    auto residual = [&]() {
        zero_u64_kernel<<<1, 64, 0, s>>>(nrm + 1, 1);
        EBArgs a{};
        a.L = lev(0);
        a.x = phi;
        a.b = b;
        a.rout = l0.b_buf.as<double>();
        a.norm = nrm + 1;
        eb_stencil_kernel<<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, s>>>(a, EB_RESID);
        return read_u64_norm(nrm + 1, h_pinned + 1, s);
    };
    const double r0 = residual();
    const double tol = std::max(atol, rtol * bnorm);
    double rn = r0;
    int it = 0;
    bool conv = std::isfinite(bnorm) && rn <= tol;
    while (!conv && it < o.maxiter && std::isfinite(rn)) {
        ++it;
        vcycle(o, s);
        double* pout = (phi == phi_a) ? phi_b : phi_a;
        zero_u64_kernel<<<1, 64, 0, s>>>(nrm + 1, 1);
        EBArgs a{};
        a.L = lev(0);
        a.x = phi;
        a.cor = l0.x_buf.as<double>();
        a.b = b;
        a.xout = pout;
        a.rout = l0.b_buf.as<double>();
        a.norm = nrm + 1;
        eb_stencil_kernel<<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, s>>>(a, EB_UPDRES);
        phi = pout;
        rn = read_u64_norm(nrm + 1, h_pinned + 1, s);
        if (o.verbose)
            fprintf(stderr, "mgx EB nodal MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                    bnorm > 0 ? rn / bnorm : rn);
        conv = rn <= tol;
    }
    if (g.singular() && nact > 0) {
        const int np = (int)std::min<long>(1024, std::max<long>(1, (nn + 255) / 256));
        partial_sum_kernel<double><<<np, 256, 0, s>>>(phi, nn, part.as<double>());
        final_sum_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, scal.as<double>() + 34,
                                           1.0 / (double)nact);
        eb_subtract_masked_kernel<<<grid1d(nn), 256, 0, s>>>(mask, phi, nn, scal.as<double>() + 34);
    }
    MGX_HIP(hipStreamSynchronize(s));
    if (info) {
        info->iters = it;
        info->vcycles = it;
        info->converged = conv ? 1 : 0;
        info->nlevels = (int)L.size();
        info->fallback_used = 0;
        info->bnorm = bnorm;
        info->resnorm0 = r0;
        info->resnorm = rn;
        info->rel_resid = bnorm > 0 ? rn / bnorm : rn;
    }
}

}  // namespace mgx

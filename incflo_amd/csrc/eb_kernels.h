// eb_kernels.h -- embedded-boundary (cut-cell) nodal operator on the GPU.
//
// Every cell of a level is regular (sigma * E_ref, E_ref the FD-normalised Q1
// element of the level's spacing), covered (no contribution) or cut (its own
// 8x8 element matrix K, sigma included, in a table).  On the finest level a
// cut cell's K = sigma * E_geom, E_geom assembled from the fluid moments of
// the cell; on coarser levels K is the Galerkin product of the 8 children,
//   K_C = 1/8 sum_f P_f^T K_f P_f  (P_f: trilinear weights of the coarse
//   corners at the child's corners),
// which is the exact coarse element when sigma is constant.  Nodes whose
// cells are all covered, and Dirichlet nodes, are inactive and hold 0.
//   (L x)_n = - sum_{cells c of n} sum_m K_c[n_local][m] x_m
#pragma once

#include <hip/hip_runtime.h>

#include "box_kernels.h"
#include "nodal_kernels.h"

namespace mgx {

struct EBLev {
    KGeom g;
    double E[64];               // regular element (sigma excluded)
    const double* sig;          // ghosted cell sigma (regular cells)
    const int* cidx;            // per cell: -1 regular, -2 covered, >= 0 cut index
    const double* K;            // cut elements, 64 per cut cell (sigma included)
    const unsigned char* mask;  // per node: 1 active
};

enum { EB_APPLY = 0, EB_RESID = 1, EB_JACOBI = 2, EB_UPDRES = 3 };

struct EBArgs {
    EBLev L;
    const double* x;
    const double* cor;   // EB_UPDRES: x + cor is the new iterate
    const double* b;
    double* xout;
    double* rout;
    unsigned long long* norm;
    double omega;
};

// wrapped cell index along d of node i's side a; -1 outside a non-periodic domain
__device__ __forceinline__ int eb_cell(const KGeom& g, int d, int i, int a) {
    int c = i - 1 + a;
    if (g.per[d]) return c < 0 ? c + g.n[d] : (c >= g.n[d] ? c - g.n[d] : c);
    return (c < 0 || c >= g.n[d]) ? -1 : c;
}

__device__ __forceinline__ int eb_node(const KGeom& g, int d, int i) {
    if (g.per[d]) return i < 0 ? i + g.nn[d] : (i >= g.nn[d] ? i - g.nn[d] : i);
    return i;
}

// (L x)_n and the diagonal of L at node (i, j, k); x + cor when cor != nullptr
__device__ inline double eb_apply_node(const EBLev& L, const double* x, const double* cor, int i,
                                       int j, int k, double& diag) {
    const KGeom& g = L.g;
    double acc = 0.0, dg = 0.0;
    for (int a = 0; a < 8; ++a) {
        const int ax = a & 1, ay = (a >> 1) & 1, az = (a >> 2) & 1;
        const int c0 = eb_cell(g, 0, i, ax), c1 = eb_cell(g, 1, j, ay), c2 = eb_cell(g, 2, k, az);
        if (c0 < 0 || c1 < 0 || c2 < 0) continue;
        const long cl = c0 + (long)g.n[0] * (c1 + (long)g.n[1] * c2);
        const int ci = L.cidx[cl];
        if (ci == -2) continue;
        const double* Kr;
        double s = 1.0;
        if (ci >= 0) {
            Kr = L.K + 64L * ci;
        } else {
            Kr = L.E;
            s = L.sig[(c0 + 1) + (long)g.sg[0] * ((c1 + 1) + (long)g.sg[1] * (c2 + 1))];
            if (s == 0.0) continue;
        }
        const int p = (1 - ax) | ((1 - ay) << 1) | ((1 - az) << 2);
        double e = 0.0;
        for (int m = 0; m < 8; ++m) {
            const int mi = eb_node(g, 0, i - 1 + ax + (m & 1));
            const int mj = eb_node(g, 1, j - 1 + ay + ((m >> 1) & 1));
            const int mk = eb_node(g, 2, k - 1 + az + ((m >> 2) & 1));
            const long mid = mi + (long)g.nn[0] * (mj + (long)g.nn[1] * mk);
            double v = x[mid];
            if (cor) v += cor[mid];
            e += Kr[p * 8 + m] * v;
        }
        acc -= s * e;
        dg -= s * Kr[p * 9];
    }
    diag = dg;
    return acc;
}

static __global__ void __launch_bounds__(256) eb_stencil_kernel(EBArgs a, int mode) {
    int i, j, k;
    const bool valid = grid_point(a.L.g.nn[0], a.L.g.nn[1], i, j, k);
    double r = 0.0;
    if (valid) {
        const long id = i + (long)a.L.g.nn[0] * (j + (long)a.L.g.nn[1] * k);
        if (a.L.mask[id]) {
            double dg;
            const double lx =
                eb_apply_node(a.L, a.x, mode == EB_UPDRES ? a.cor : nullptr, i, j, k, dg);
            if (mode == EB_APPLY) {
                a.xout[id] = lx;
            } else if (mode == EB_JACOBI) {
                a.xout[id] = (dg != 0.0) ? a.x[id] + a.omega * (a.b[id] - lx) / dg : 0.0;
            } else {
                r = a.b[id] - lx;
                a.rout[id] = r;
                if (mode == EB_UPDRES) a.xout[id] = a.x[id] + a.cor[id];
            }
        } else if (mode == EB_APPLY || mode == EB_JACOBI) {
            a.xout[id] = 0.0;
        } else {
            a.rout[id] = 0.0;
            if (mode == EB_UPDRES) a.xout[id] = 0.0;
        }
    }
    if ((mode == EB_RESID || mode == EB_UPDRES) && a.norm) {
        // one atomic per wave; every lane stays active through the shuffles
        double v = nabs(r);
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) atomic_max_nonneg(a.norm, v);
    }
}

// x = 0 on inactive nodes (after prolongation)
static __global__ void eb_mask_kernel(const unsigned char* __restrict__ mask, double* __restrict__ x,
                                      long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n && !mask[id]) x[id] = 0.0;
}

// finest-level cut elements: K = sigma_c * E_geom
static __global__ void eb_scale_kernel(KGeom g, const int* __restrict__ cells, int ncut,
                                       const double* __restrict__ Egeom, const double* __restrict__ sig,
                                       double* __restrict__ K) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 64L * ncut) return;
    const int c = (int)(t / 64);
    const int cl = cells[c];
    const int c0 = cl % g.n[0], c1 = (cl / g.n[0]) % g.n[1], c2 = cl / (g.n[0] * g.n[1]);
    const double s = sig[(c0 + 1) + (long)g.sg[0] * ((c1 + 1) + (long)g.sg[1] * (c2 + 1))];
    K[t] = s * Egeom[t];
}

// Galerkin coarse element of each coarse cut cell from its 8 children
// (children: fine cell linear indices, child order x fastest)
static __global__ void eb_galerkin_kernel(EBLev F, const int* __restrict__ children, int ncut,
                                          double* __restrict__ Kc) {
    const int c = blockIdx.x;
    if (c >= ncut) return;
    const int t = threadIdx.x;  // 64 threads: entry (a, b) of the coarse element
    const int a = t >> 3, b = t & 7;
    double acc = 0.0;
    for (int f = 0; f < 8; ++f) {
        const int cl = children[8 * c + f];
        const int ci = F.cidx[cl];
        if (ci == -2) continue;
        const double* Kf;
        double s = 1.0;
        if (ci >= 0) {
            Kf = F.K + 64L * ci;
        } else {
            const int c0 = cl % F.g.n[0], c1 = (cl / F.g.n[0]) % F.g.n[1], c2 = cl / (F.g.n[0] * F.g.n[1]);
            s = F.sig[(c0 + 1) + (long)F.g.sg[0] * ((c1 + 1) + (long)F.g.sg[1] * (c2 + 1))];
            Kf = F.E;
        }
        const int fx = f & 1, fy = (f >> 1) & 1, fz = (f >> 2) & 1;
        // P[p][corner] = prod_d w(corner_d, (child_d + p_d) / 2)
        double Pa[8], Pb[8];
        for (int p = 0; p < 8; ++p) {
            const double tx = 0.5 * (fx + (p & 1)), ty = 0.5 * (fy + ((p >> 1) & 1)),
                         tz = 0.5 * (fz + ((p >> 2) & 1));
            Pa[p] = ((a & 1) ? tx : 1.0 - tx) * (((a >> 1) & 1) ? ty : 1.0 - ty) *
                    (((a >> 2) & 1) ? tz : 1.0 - tz);
            Pb[p] = ((b & 1) ? tx : 1.0 - tx) * (((b >> 1) & 1) ? ty : 1.0 - ty) *
                    (((b >> 2) & 1) ? tz : 1.0 - tz);
        }
        double e = 0.0;
        for (int p = 0; p < 8; ++p) {
            if (Pa[p] == 0.0) continue;
            double row = 0.0;
            for (int q = 0; q < 8; ++q) row += Kf[p * 8 + q] * Pb[q];
            e += Pa[p] * row;
        }
        acc += s * e;
    }
    Kc[64L * c + t] = 0.125 * acc;
}

// EB nodal divergence rhs (ghosted canonical cell velocity ug, 3 comps):
// regular cells use the standard weights (+-1/(4h) each), cut cells their
// table W (24 per cut cell: [corner][d]), covered cells nothing.
static __global__ void __launch_bounds__(256) eb_rhs_kernel(KGeom g, const int* __restrict__ cidx,
                                                            const double* __restrict__ W,
                                                            const unsigned char* __restrict__ mask,
                                                            const double* __restrict__ ug,
                                                            double* __restrict__ rhs, double fx,
                                                            double fy, double fz) {
    const double fr[3] = {fx, fy, fz};  // 1 / (4 h_d): regular-cell weights
    int i, j, k;
    if (!grid_point(g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    if (!mask[id]) { rhs[id] = 0.0; return; }
    const long gs = (long)g.sg[0] * g.sg[1] * g.sg[2];
    double acc = 0.0;
    for (int a = 0; a < 8; ++a) {
        const int cd[3] = {a & 1, (a >> 1) & 1, (a >> 2) & 1};
        const int n3[3] = {i, j, k};
        int c[3], cg[3];
        bool inside = true;
        for (int d = 0; d < 3; ++d) {
            int cc = n3[d] - 1 + cd[d];
            if (g.per[d]) cc = cc < 0 ? cc + g.n[d] : (cc >= g.n[d] ? cc - g.n[d] : cc);
            else if (cc < 0 || cc >= g.n[d]) inside = false;
            c[d] = cc;
            cg[d] = cc + 1;
        }
        const int p = (1 - cd[0]) | ((1 - cd[1]) << 1) | ((1 - cd[2]) << 2);
        const long gc = cg[0] + (long)g.sg[0] * (cg[1] + (long)g.sg[1] * cg[2]);
        int ci = -1;
        if (inside) ci = cidx[c[0] + (long)g.n[0] * (c[1] + (long)g.n[1] * c[2])];
        if (ci == -2) continue;
        for (int d = 0; d < 3; ++d) {
            double w;
            if (ci >= 0) w = W[24L * ci + 3 * p + d];
            else w = ((p >> d) & 1) ? fr[d] : -fr[d];
            acc -= ug[d * gs + gc] * w;
        }
    }
    rhs[id] = acc;
}

// EB inflow velocity: rhs += sum over the cut cells around each active node of
// B[corner] . u_eb (B: 24 per cut cell, (1/V) int over the body surface of
// N_p n_d dA with n out of the fluid; ueb: 3 x cells, x fastest)
static __global__ void __launch_bounds__(256) eb_inflow_rhs_kernel(KGeom g, const int* __restrict__ cidx,
                                                                   const double* __restrict__ B,
                                                                   const unsigned char* __restrict__ mask,
                                                                   const double* __restrict__ ueb,
                                                                   double* __restrict__ rhs) {
    int i, j, k;
    if (!grid_point(g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    if (!mask[id]) return;
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    double acc = 0.0;
    for (int a = 0; a < 8; ++a) {
        const int cd[3] = {a & 1, (a >> 1) & 1, (a >> 2) & 1};
        const int n3[3] = {i, j, k};
        int c[3];
        bool inside = true;
        for (int d = 0; d < 3; ++d) {
            int cc = n3[d] - 1 + cd[d];
            if (g.per[d]) cc = cc < 0 ? cc + g.n[d] : (cc >= g.n[d] ? cc - g.n[d] : cc);
            else if (cc < 0 || cc >= g.n[d]) inside = false;
            c[d] = cc;
        }
        if (!inside) continue;
        const long cl = c[0] + (long)g.n[0] * (c[1] + (long)g.n[1] * c[2]);
        const int ci = cidx[cl];
        if (ci < 0) continue;
        const int p = (1 - cd[0]) | ((1 - cd[1]) << 1) | ((1 - cd[2]) << 2);
        for (int d = 0; d < 3; ++d) acc += B[24L * ci + 3 * p + d] * ueb[d * nc + cl];
    }
    rhs[id] += acc;
}

// canonical cells (ncomp x cells, no ghosts) <- boxed cells (valid region)
static __global__ void __launch_bounds__(256) gather_cells_kernel(KGeom g, const BoxInfo* boxes,
                                                                  const Tile* tiles,
                                                                  const double* __restrict__ in,
                                                                  int ng, int ncomp,
                                                                  double* __restrict__ out) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    const long o = field_offset(b, ng, ncomp, false);
    const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
    for (int q = 0; q < ncomp; ++q) out[q * nc + c] = in[o + box_index(b, ng, false, i, j, k, q)];
}

}  // namespace mgx

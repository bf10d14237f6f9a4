// box_kernels.h -- kernels that move data between the caller's boxed fields
// (AMReX Array4 layout, boxes back to back in one allocation) and the
// solver's canonical level arrays, driven by a device box-offset table and a
// tile table (one 256-cell chunk of one box per workgroup), so that all boxes
// of uneven sizes are handled by a single launch.
#pragma once

#include "mgx.h"
#include <hip/hip_runtime.h>

#include "nodal_kernels.h"

namespace mgx {

struct BoxInfo {
    int lo[3];
    int len[3];      // valid cells per direction
    int rlo[3];      // iteration region lo / extent for "grown" cell loops
    int rlen[3];
    long pc[4];      // prefix sums over earlier boxes for cell-field offsets
    long pn[4];      // ... for nodal-field offsets
    long pf[3][4];   // ... for face-field offsets (faces normal to x, y, z)
};

struct Tile {
    int box;
    int start;
};

enum { REG_CELL = 0, REG_CELL_GROWN = 1, REG_NODE = 2, REG_FACE0 = 3, REG_NREG = 6 };

__device__ __forceinline__ long field_offset(const BoxInfo& b, int ng, int ncomp, bool nodal) {
    const long* P = nodal ? b.pn : b.pc;
    const long g = ng;
    return (long)ncomp * (P[0] + 2 * g * P[1] + 4 * g * g * P[2] + 8 * g * g * g * P[3]);
}

__device__ __forceinline__ long face_offset(const BoxInfo& b, int ng, int d) {
    const long* P = b.pf[d];
    const long g = ng;
    return P[0] + 2 * g * P[1] + 4 * g * g * P[2] + 8 * g * g * g * P[3];
}

// index of face (i,j,k) (normal d) inside box b of a face field with ghost width ng
__device__ __forceinline__ long box_index_face(const BoxInfo& b, int ng, int d, int i, int j,
                                               int k) {
    const long D0 = b.len[0] + (d == 0) + 2 * ng, D1 = b.len[1] + (d == 1) + 2 * ng;
    return (i - b.lo[0] + ng) + D0 * ((j - b.lo[1] + ng) + D1 * (long)(k - b.lo[2] + ng));
}

// index of (i,j,k,comp) inside box b of a field with ghost width ng
__device__ __forceinline__ long box_index(const BoxInfo& b, int ng, bool nodal, int i, int j, int k,
                                          int comp) {
    const int e = nodal ? 1 : 0;
    const long D0 = b.len[0] + e + 2 * ng, D1 = b.len[1] + e + 2 * ng, D2 = b.len[2] + e + 2 * ng;
    return (i - b.lo[0] + ng) + D0 * ((j - b.lo[1] + ng) + D1 * ((k - b.lo[2] + ng) + D2 * comp));
}

// Decode the (box, i, j, k) handled by this thread; false when past the region.
__device__ __forceinline__ bool tile_point(const BoxInfo* boxes, const Tile* tiles, int region,
                                           int& bidx, int& i, int& j, int& k) {
    const Tile t = tiles[blockIdx.x];
    bidx = t.box;
    const BoxInfo& b = boxes[bidx];
    int lo[3], len[3];
    for (int d = 0; d < 3; ++d) {
        if (region == REG_CELL_GROWN) { lo[d] = b.rlo[d]; len[d] = b.rlen[d]; }
        else if (region >= REG_FACE0) { lo[d] = b.lo[d]; len[d] = b.len[d] + (region - REG_FACE0 == d); }
        else { lo[d] = b.lo[d]; len[d] = b.len[d] + (region == REG_NODE ? 1 : 0); }
    }
    const int lin = t.start + (int)threadIdx.x;
    const int vol = len[0] * len[1] * len[2];
    if (lin >= vol) return false;
    const int q = lin / len[0];
    i = lo[0] + (lin - q * len[0]);
    j = lo[1] + q % len[1];
    k = lo[2] + q / len[1];
    return true;
}

struct GatherVelArgs {
    KGeom g;
    const BoxInfo* boxes;
    const Tile* tiles;
    const double* vel; int vel_ng;
    const double* vel_old; int vold_ng;
    const double* gp; int gp_ng;
    const double* rho; int rho_ng;
    double const_rho;  // > 0: sigma = scale / const_rho
    double scale;      // scaling_factor s
    int sigma_from_rho;
    double* ug;        // canonical ghosted cell velocity (3 comps)
    double* sig;       // canonical ghosted sigma (level 0)
    int ghost_rule;    // 0: take domain-face ghosts from vel; 1: 0 / inflow table
    double inflow[18];
    int bclo[3], bchi[3];
};

// u* = u (+ (s/rho) gp) (- u_old) on valid cells; sigma = s/rho; ghost cells
// outside non-periodic faces from the caller or the inflow rule.
static __global__ void __launch_bounds__(256) gather_vel_kernel(GatherVelArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL_GROWN, bi, i, j, k)) return;
    const BoxInfo& b = a.boxes[bi];
    const KGeom& g = a.g;
    const long ug3 = (long)g.sg[0] * g.sg[1] * g.sg[2];
    const bool inside = i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2];
    const long boff_v = field_offset(b, a.vel_ng, 3, false);
    if (inside) {
        double s_rho = 0.0, sigma = 0.0;
        if (a.rho) {
            const double r = a.rho[field_offset(b, a.rho_ng, 1, false) +
                                   box_index(b, a.rho_ng, false, i, j, k, 0)];
            s_rho = a.scale / r;
        } else {
            s_rho = a.scale / a.const_rho;
        }
        sigma = (a.const_rho > 0.0) ? a.scale / a.const_rho : s_rho;
        for (int c = 0; c < 3; ++c) {
            double u = a.vel[boff_v + box_index(b, a.vel_ng, false, i, j, k, c)];
            if (a.gp) u += s_rho * a.gp[field_offset(b, a.gp_ng, 3, false) +
                                        box_index(b, a.gp_ng, false, i, j, k, c)];
            if (a.vel_old) u -= a.vel_old[field_offset(b, a.vold_ng, 3, false) +
                                          box_index(b, a.vold_ng, false, i, j, k, c)];
            write_cell_images(a.ug + c * ug3, g, i, j, k, u);
        }
        if (a.sigma_from_rho) write_cell_images(a.sig, g, i, j, k, sigma);
    } else {
        // ghost cell outside at least one non-periodic face
        int nonper = 0, face = -1;
        const int c3[3] = {i, j, k};
        for (int d = 0; d < 3; ++d)
            if (c3[d] < 0 || c3[d] >= g.n[d]) { ++nonper; face = 2 * d + (c3[d] < 0 ? 0 : 1); }
        for (int c = 0; c < 3; ++c) {
            double u;
            if (!a.ghost_rule) {
                u = a.vel[boff_v + box_index(b, a.vel_ng, false, i, j, k, c)];
            } else if (nonper == 1) {
                const int d = face >> 1;
                const int bc = (face & 1) ? a.bchi[d] : a.bclo[d];
                u = (bc == MGX_BC_INFLOW) ? a.inflow[face * 3 + c] : 0.0;
            } else {
                u = 0.0;
            }
            // periodic images in the other (periodic) directions
            int xs[2] = {i + 1, -1}, ys[2] = {j + 1, -1}, zs[2] = {k + 1, -1};
            if (g.per[0]) xs[1] = (i == 0) ? g.n[0] + 1 : (i == g.n[0] - 1 ? 0 : -1);
            if (g.per[1]) ys[1] = (j == 0) ? g.n[1] + 1 : (j == g.n[1] - 1 ? 0 : -1);
            if (g.per[2]) zs[1] = (k == 0) ? g.n[2] + 1 : (k == g.n[2] - 1 ? 0 : -1);
            for (int q = 0; q < 8; ++q) {
                const int X = xs[q & 1], Y = ys[(q >> 1) & 1], Z = zs[(q >> 2) & 1];
                if (X < 0 || Y < 0 || Z < 0) continue;
                a.ug[c * ug3 + X + (long)g.sg[0] * (Y + (long)g.sg[1] * Z)] = u;
            }
        }
    }
}

// (D u)_n = sum_d 1/(4 h_d) [ sum_{+d cells} u_d - sum_{-d cells} u_d ]
static __global__ void __launch_bounds__(256) nodal_rhs_kernel(KGeom g, const double* __restrict__ ug,
                                                        double* __restrict__ rhs, double fx,
                                                        double fy, double fz,
                                                        const int4* tl = nullptr) {
    int i, j, k;
    if (!grid_point_tl(tl, g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    const long ug3 = (long)g.sg[0] * g.sg[1] * g.sg[2];
    double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int aa = q & 1, bb = (q >> 1) & 1, cc = (q >> 2) & 1;
        const long gi = (i + aa) + (long)g.sg[0] * ((j + bb) + (long)g.sg[1] * (k + cc));
        const double ux = ug[gi], uy = ug[ug3 + gi], uz = ug[2 * ug3 + gi];
        sx += aa ? ux : -ux;
        sy += bb ? uy : -uy;
        sz += cc ? uz : -uz;
    }
    rhs[id] = sx * fx + sy * fy + sz * fz;
}

// sigma from a boxed cell field
static __global__ void __launch_bounds__(256) gather_sigma_kernel(KGeom g, const BoxInfo* boxes,
                                                           const Tile* tiles, const double* sigma,
                                                           int ng, double* sig) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    write_cell_images(sig, g, i, j, k,
                      sigma[field_offset(b, ng, 1, false) + box_index(b, ng, false, i, j, k, 0)]);
}

__device__ __forceinline__ void cell_gradient(const KGeom& g, const double* phi, int i, int j,
                                              int k, const double f[3], double gr[3]) {
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int aa = q & 1, bb = (q >> 1) & 1, cc = (q >> 2) & 1;
        int ii = i + aa, jj = j + bb, kk = k + cc;
        if (ii >= g.nn[0]) ii -= g.nn[0];
        if (jj >= g.nn[1]) jj -= g.nn[1];
        if (kk >= g.nn[2]) kk -= g.nn[2];
        const double v = phi[ii + (long)g.nn[0] * (jj + (long)g.nn[1] * kk)];
        gx += aa ? v : -v;
        gy += bb ? v : -v;
        gz += cc ? v : -v;
    }
    gr[0] = gx * f[0];
    gr[1] = gy * f[1];
    gr[2] = gz * f[2];
}

struct UpdateArgs {
    KGeom g;
    const BoxInfo* boxes;
    const Tile* tiles;
    const double* phi;     // canonical nodes
    const double* ug;      // canonical u* (NULL: read vel instead)
    const double* sig;     // canonical ghosted sigma
    double f[3];           // 1/(4 h_d)
    double* vel; int vel_ng;          // u_new written here (NULL: skip)
    const double* vel_old; int vold_ng;
    double* gp; int gp_ng; int gp_add;   // G(phi) written / added (NULL: skip)
    // embedded boundary (NULL: none): per cell -1 regular / -2 covered / cut
    // index; cut cells' gradient weights W[24] ([corner][d]) and volume fraction
    const int* eb_cidx;
    const double* eb_W;
    const double* eb_V;
    // canonical ghosted copies of u_new / G phi at the valid cells (NULL:
    // none; ucan may be ug itself: each cell reads u* before writing)
    double* ucan;
    double* gcan;
};

// u_new = u* - sigma G(phi) (+ u_old); gp = G(phi) or gp += G(phi)
static __global__ void __launch_bounds__(256) update_cells_kernel(UpdateArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = a.boxes[bi];
    const KGeom& g = a.g;
    double gr[3];
    int ci = -1;
    if (a.eb_cidx) ci = a.eb_cidx[i + (long)g.n[0] * (j + (long)g.n[1] * k)];
    if (ci >= 0) {
        // cut cell: average gradient of the Q1 phi over the fluid part
        gr[0] = gr[1] = gr[2] = 0.0;
        for (int q = 0; q < 8; ++q) {
            int ii = i + (q & 1), jj = j + ((q >> 1) & 1), kk = k + ((q >> 2) & 1);
            if (ii >= g.nn[0]) ii -= g.nn[0];
            if (jj >= g.nn[1]) jj -= g.nn[1];
            if (kk >= g.nn[2]) kk -= g.nn[2];
            const double v = a.phi[ii + (long)g.nn[0] * (jj + (long)g.nn[1] * kk)];
            for (int d = 0; d < 3; ++d) gr[d] += a.eb_W[24L * ci + 3 * q + d] * v;
        }
        const double vf = a.eb_V[ci];
        for (int d = 0; d < 3; ++d) gr[d] /= vf;
    } else if (ci == -2) {
        gr[0] = gr[1] = gr[2] = 0.0;
    } else {
        cell_gradient(g, a.phi, i, j, k, a.f, gr);
    }
    if (a.vel) {
        const long ug3 = (long)g.sg[0] * g.sg[1] * g.sg[2];
        const long gi = (i + 1) + (long)g.sg[0] * ((j + 1) + (long)g.sg[1] * (k + 1));
        const double sigma = a.sig[gi];
        const long off = field_offset(b, a.vel_ng, 3, false);
        for (int c = 0; c < 3; ++c) {
            const long vi = off + box_index(b, a.vel_ng, false, i, j, k, c);
            double u = (a.ug ? a.ug[c * ug3 + gi] : a.vel[vi]) - sigma * gr[c];
            if (a.vel_old) u += a.vel_old[field_offset(b, a.vold_ng, 3, false) +
                                          box_index(b, a.vold_ng, false, i, j, k, c)];
            if (ci == -2) u = 0.0;  // covered cells carry no velocity
            a.vel[vi] = u;
            if (a.ucan) a.ucan[c * ug3 + gi] = u;
        }
    }
    if (a.gcan) {
        const long ug3 = (long)g.sg[0] * g.sg[1] * g.sg[2];
        const long gi = (i + 1) + (long)g.sg[0] * ((j + 1) + (long)g.sg[1] * (k + 1));
        for (int c = 0; c < 3; ++c) a.gcan[c * ug3 + gi] = gr[c];
    }
    if (a.gp) {
        const long off = field_offset(b, a.gp_ng, 3, false);
        for (int c = 0; c < 3; ++c) {
            const long gi = off + box_index(b, a.gp_ng, false, i, j, k, c);
            a.gp[gi] = a.gp_add ? a.gp[gi] + gr[c] : gr[c];
        }
    }
}

// boxed nodal field <- canonical nodes (periodic node n == node 0)
static __global__ void __launch_bounds__(256) scatter_nodes_kernel(KGeom g, const BoxInfo* boxes,
                                                            const Tile* tiles, const double* phi,
                                                            double* out, int ng, int add) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_NODE, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const int ii = (i >= g.nn[0]) ? i - g.nn[0] : i;
    const int jj = (j >= g.nn[1]) ? j - g.nn[1] : j;
    const int kk = (k >= g.nn[2]) ? k - g.nn[2] : k;
    const double v = phi[ii + (long)g.nn[0] * (jj + (long)g.nn[1] * kk)];
    const long o = field_offset(b, ng, 1, true) + box_index(b, ng, true, i, j, k, 0);
    out[o] = add ? out[o] + v : v;
}

// canonical nodes <- boxed nodal field
static __global__ void __launch_bounds__(256) gather_nodes_kernel(KGeom g, const BoxInfo* boxes,
                                                           const Tile* tiles, const double* in,
                                                           int ng, double* phi) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_NODE, bi, i, j, k)) return;
    if (i >= g.nn[0] || j >= g.nn[1] || k >= g.nn[2]) return;  // periodic duplicate
    const BoxInfo& b = boxes[bi];
    phi[i + (long)g.nn[0] * (j + (long)g.nn[1] * k)] =
        in[field_offset(b, ng, 1, true) + box_index(b, ng, true, i, j, k, 0)];
}

// canonical node flags (1: value != 0) <- boxed nodal int field
static __global__ void __launch_bounds__(256) gather_node_flags_kernel(KGeom g, const BoxInfo* boxes,
                                                                     const Tile* tiles, const int* in,
                                                                     int ng, unsigned char* out) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_NODE, bi, i, j, k)) return;
    if (i >= g.nn[0] || j >= g.nn[1] || k >= g.nn[2]) return;  // periodic duplicate
    const BoxInfo& b = boxes[bi];
    out[i + (long)g.nn[0] * (j + (long)g.nn[1] * k)] =
        in[field_offset(b, ng, 1, true) + box_index(b, ng, true, i, j, k, 0)] != 0;
}

// ---- face and cell fields of the MAC projection -----------------------------
// canonical face index (faces along d: n_d + 1 of them)
__device__ __forceinline__ long canon_face(const KGeom& g, int d, int i, int j, int k) {
    const long f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1);
    return i + f0 * (j + f1 * (long)k);
}

// canonical faces <- boxed face field (dir = d), or the reverse when to_box
// The ghost cell just outside the domain face (d, side) of a boxed cell field
// (ng >= 1), one value per boundary face (plane element t: x fastest over the
// two remaining directions), read from the box holding the adjacent valid
// cell; NaN where no box holds it.
static __global__ void __launch_bounds__(256) gather_face_ghost_kernel(KGeom g, const BoxInfo* boxes,
                                                                       int nboxes,
                                                                       const double* __restrict__ in,
                                                                       int ng, int d, int side,
                                                                       double* __restrict__ out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
    const long m0 = g.n[a0];
    if (t >= m0 * g.n[a1]) return;
    int p[3];
    p[d] = side ? g.n[d] - 1 : 0;
    p[a0] = (int)(t % m0);
    p[a1] = (int)(t / m0);
    double v = __builtin_nan("");
    for (int bi = 0; bi < nboxes; ++bi) {
        const BoxInfo& b = boxes[bi];
        if (p[0] < b.lo[0] || p[0] >= b.lo[0] + b.len[0] || p[1] < b.lo[1] ||
            p[1] >= b.lo[1] + b.len[1] || p[2] < b.lo[2] || p[2] >= b.lo[2] + b.len[2])
            continue;
        int q[3] = {p[0], p[1], p[2]};
        q[d] += side ? 1 : -1;
        v = in[field_offset(b, ng, 1, false) + box_index(b, ng, false, q[0], q[1], q[2], 0)];
        break;
    }
    out[t] = v;
}

static __global__ void __launch_bounds__(256) face_copy_kernel(KGeom g, const BoxInfo* boxes,
                                                              const Tile* tiles, int d, double* boxed,
                                                              int ng, double* canon, int to_box) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_FACE0 + d, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long o = face_offset(b, ng, d) + box_index_face(b, ng, d, i, j, k);
    const long c = canon_face(g, d, i, j, k);
    if (to_box) boxed[o] = canon[c];
    else canon[c] = boxed[o];
}

// canonical cells (no ghosts) <- boxed cell field (valid cells), or the reverse
static __global__ void __launch_bounds__(256) cell_copy_kernel(KGeom g, const BoxInfo* boxes,
                                                              const Tile* tiles, double* boxed,
                                                              int ng, double* canon, int to_box) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long o = field_offset(b, ng, 1, false) + box_index(b, ng, false, i, j, k, 0);
    const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
    if (to_box) boxed[o] = canon[c];
    else canon[c] = boxed[o];
}

}  // namespace mgx

namespace mgx {

// ---- AMR synchronisation (refinement ratio 2): average_down_faces and
// average_down of cell fields.  One thread per fine face (cell) whose indices
// are all even: it averages the 4 fine faces (8 fine cells) over one coarse
// face (cell) and writes every coarse box that holds that face (cell).

__device__ __forceinline__ bool box_holds(const BoxInfo& b, int i, int j, int k, int d) {
    // d < 0: cell (i,j,k); else face (i,j,k) normal to d (lo..hi+1 along d)
    const int p[3] = {i, j, k};
    for (int e = 0; e < 3; ++e) {
        const int hi = b.lo[e] + b.len[e] - 1 + (e == d ? 1 : 0);
        if (p[e] < b.lo[e] || p[e] > hi) return false;
    }
    return true;
}

static __global__ void __launch_bounds__(256) average_down_faces_kernel(
    KGeom gf, const BoxInfo* fboxes, const Tile* ftiles, int d, const double* __restrict__ fine,
    int fng, const BoxInfo* cboxes, int ncboxes, double* __restrict__ crse, int cng, int ncomp) {
    int bi, i, j, k;
    if (!tile_point(fboxes, ftiles, REG_FACE0 + d, bi, i, j, k)) return;
    if ((i | j | k) & 1) return;
    const BoxInfo& b = fboxes[bi];
    const long fvol = (long)(b.len[0] + (d == 0) + 2 * fng) * (b.len[1] + (d == 1) + 2 * fng) *
                      (b.len[2] + (d == 2) + 2 * fng);
    const long fo = (long)ncomp * face_offset(b, fng, d);
    const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    for (int cb = 0; cb < ncboxes; ++cb) {
        const BoxInfo& c = cboxes[cb];
        if (!box_holds(c, I, J, K, d)) continue;
        const long cvol = (long)(c.len[0] + (d == 0) + 2 * cng) * (c.len[1] + (d == 1) + 2 * cng) *
                          (c.len[2] + (d == 2) + 2 * cng);
        const long co = (long)ncomp * face_offset(c, cng, d) + box_index_face(c, cng, d, I, J, K);
        for (int q = 0; q < ncomp; ++q) {
            double s = 0.0;
            for (int t = 0; t < 4; ++t) {
                int p[3] = {i, j, k};
                p[e1] += t & 1;
                p[e2] += t >> 1;
                s += fine[fo + q * fvol + box_index_face(b, fng, d, p[0], p[1], p[2])];
            }
            crse[co + q * cvol] = 0.25 * s;
        }
    }
}

static __global__ void __launch_bounds__(256) average_down_cells_kernel(
    KGeom gf, const BoxInfo* fboxes, const Tile* ftiles, const double* __restrict__ fine, int fng,
    const BoxInfo* cboxes, int ncboxes, double* __restrict__ crse, int cng, int ncomp) {
    int bi, i, j, k;
    if (!tile_point(fboxes, ftiles, REG_CELL, bi, i, j, k)) return;
    if ((i | j | k) & 1) return;
    const BoxInfo& b = fboxes[bi];
    const long fo = field_offset(b, fng, ncomp, false);
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    for (int cb = 0; cb < ncboxes; ++cb) {
        const BoxInfo& c = cboxes[cb];
        if (!box_holds(c, I, J, K, -1)) continue;
        const long co = field_offset(c, cng, ncomp, false);
        for (int q = 0; q < ncomp; ++q) {
            double s = 0.0;
            for (int t = 0; t < 8; ++t)
                s += fine[fo + box_index(b, fng, false, i + (t & 1), j + ((t >> 1) & 1), k + (t >> 2), q)];
            crse[co + box_index(c, cng, false, I, J, K, q)] = 0.125 * s;
        }
        break;  // a cell belongs to one box
    }
}

}  // namespace mgx

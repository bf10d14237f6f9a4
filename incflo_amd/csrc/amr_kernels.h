// amr_kernels.h -- kernels of the composite nodal solve over AMR levels
// (refinement ratio 2 between neighbours; the reference projects all AMR
// levels in one NodalProjector,
// src/projection/incflo_apply_nodal_projection.cpp:124-215).  Each kernel
// works on one pair of neighbouring levels ("coarse" and "fine" below); a
// level between two others is the fine level of one pair and the coarse
// level of the next, where its unknowns are its AMR_DOF nodes not surrounded
// by cells under the finer level.
//
// Composite problem: conforming Q1 elements on the composite mesh.  F is the
// union of the fine boxes; its fine nodes are unknowns (AMR_DOF), fine nodes
// on the coarse-fine boundary are fixed by the coarse nodes (AMR_BDRY:
// coincident ones take the coarse value, the others the trilinear
// interpolation of the coarse face / edge), fine nodes away from F are unused
// (AMR_OUT).  Coarse nodes are unknowns unless every cell around them is
// covered by F.  Both levels live on dense canonical arrays (the fine one at
// twice the coarse resolution over the whole domain); the fine cell
// coefficient is 0 outside F, so the fine stencil sees only the elements of F.
//   fine unknown m:   (A x)_m = (L_f x~)_m
//   coarse unknown n: (A x)_n = (L_c^out x_c)_n + R(L_f x~ on dF)_n
// with L_c^out the coarse operator with covered cells' coefficient 0 and R the
// full-weighting restriction (P^T / 8).
#pragma once

#include <hip/hip_runtime.h>

#include "box_kernels.h"
#include "nodal_kernels.h"

namespace mgx {

enum { AMR_OUT = 0, AMR_DOF = 1, AMR_BDRY = 2 };

// trilinear interpolation of the coarse nodes at fine node (i, j, k) (the
// weights of nodal_prolong_add_kernel)
template <class T>
__device__ __forceinline__ T amr_interp(const KGeom& gc, const T* __restrict__ xc, int i, int j,
                                        int k) {
    int lo[3], hi[3];
    const int f[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = f[d] >> 1;
        hi[d] = (f[d] & 1) ? lo[d] + 1 : lo[d];
        if (gc.per[d] && hi[d] >= gc.nn[d]) hi[d] -= gc.nn[d];
    }
    T acc = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int I = (q & 1) ? hi[0] : lo[0];
        const int J = (q & 2) ? hi[1] : lo[1];
        const int K = (q & 4) ? hi[2] : lo[2];
        acc += xc[I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K)];
    }
    return (T)0.125 * acc;
}

// fine field on the closure of F: dF nodes from the coarse field, 0 outside
template <class T>
static __global__ void __launch_bounds__(256) amr_fill_kernel(KGeom gf, KGeom gc,
                                                              const unsigned char* __restrict__ ft,
                                                              const T* __restrict__ xc,
                                                              T* __restrict__ xf,
                                                              const int4* tl = nullptr) {
    int i, j, k;
    if (!grid_point_tl(tl, gf.nn[0], gf.nn[1], i, j, k)) return;
    const long id = i + (long)gf.nn[0] * (j + (long)gf.nn[1] * k);
    const int t = ft[id];
    if (t == AMR_BDRY) xf[id] = amr_interp(gc, xc, i, j, k);
    else if (t == AMR_OUT) xf[id] = 0;
}

// coarse correction onto the fine level: unknowns += P e, dF = P e, outside 0
template <class T>
static __global__ void __launch_bounds__(256) amr_prolong_kernel(KGeom gf, KGeom gc,
                                                                 const unsigned char* __restrict__ ft,
                                                                 const T* __restrict__ ec,
                                                                 T* __restrict__ ef,
                                                                 const int4* tl = nullptr) {
    int i, j, k;
    if (!grid_point_tl(tl, gf.nn[0], gf.nn[1], i, j, k)) return;
    const long id = i + (long)gf.nn[0] * (j + (long)gf.nn[1] * k);
    const int t = ft[id];
    if (t == AMR_OUT) { ef[id] = 0; return; }
    const T p = amr_interp(gc, ec, i, j, k);
    ef[id] = (t == AMR_DOF) ? ef[id] + p : p;
}

// max over the block (256 threads, every one calls this), one atomic per block
__device__ __forceinline__ void wave_max_atomic(unsigned long long* norm, double v) {
    __shared__ double red[4];
    const double m = block_max<256>(v, red);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_nonneg(norm, m);
}

// coarse unknowns: out = base + sgn w, w the restriction of the finer
// level's dF values (NULL: none); every other coarse node 0; optional max
// |out| (every lane reaches the shuffles); tl: the coarse level's tile list
// (a level above level 0), else the whole coarse grid.  A coarse level that
// is itself a fine level (ftc) also writes its own dF values of base to wb
// (0 elsewhere), for the restriction onto the level below it.
template <class TO>
static __global__ void __launch_bounds__(256) amr_coarse_combine_kernel(
    KGeom gc, const unsigned char* __restrict__ cdof, const double* __restrict__ base,
    const double* __restrict__ w, double sgn, TO* __restrict__ out, unsigned long long* norm,
    const int4* tl = nullptr, const unsigned char* __restrict__ ftc = nullptr, double* __restrict__ wb = nullptr) {
    int I, J, K;
    const bool valid = grid_point_tl(tl, gc.nn[0], gc.nn[1], I, J, K);
    double r = 0.0;
    if (valid) {
        const long id = I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K);
        if (cdof[id] && !is_dir(gc, I, J, K)) r = base[id] + (w ? sgn * w[id] : 0.0);
        out[id] = (TO)r;
        if (wb) wb[id] = ftc[id] == AMR_BDRY ? base[id] : 0.0;
    }
    if (norm) wave_max_atomic(norm, nabs(r));
}

// w = y on dF (the level's AMR_BDRY nodes), 0 elsewhere, over the level's tiles
template <class T>
static __global__ void __launch_bounds__(256) amr_mask_bdry_kernel(KGeom g, const unsigned char* __restrict__ ft,
                                                                   const T* __restrict__ y, T* __restrict__ w,
                                                                   const int4* tl) {
    int i, j, k;
    if (!grid_point_tl(tl, g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    w[id] = ft[id] == AMR_BDRY ? y[id] : (T)0;
}

// fine unknowns: r = b - y (cast to TO), 0 elsewhere; optional max |r|;
// wb (optional): y on dF, 0 elsewhere (for the restriction onto the level
// below)
template <class TO>
static __global__ void __launch_bounds__(256) amr_fine_resid_kernel(
    KGeom gf, const unsigned char* __restrict__ ft, const double* __restrict__ b,
    const double* __restrict__ y, TO* __restrict__ out, unsigned long long* norm,
    const int4* tl = nullptr, double* __restrict__ wb = nullptr) {
    int i, j, k;
    const bool valid = grid_point_tl(tl, gf.nn[0], gf.nn[1], i, j, k);
    double r = 0.0;
    if (valid) {
        const long id = i + (long)gf.nn[0] * (j + (long)gf.nn[1] * k);
        const int t = ft[id];
        if (t == AMR_DOF && !is_dir(gf, i, j, k)) r = b[id] - (y ? y[id] : 0.0);
        out[id] = (TO)r;
        if (wb) wb[id] = t == AMR_BDRY ? y[id] : 0.0;
    }
    if (norm) wave_max_atomic(norm, nabs(r));
}

// right-hand side of the coarse-grid correction (all coarse nodes):
//   b_C = [n unknown] r_c + w,  w = R(r_f - L_f e)
// with r the composite residual and e the fine correction after the
// pre-smoothing (r_f is 0 on dF, and L_f e is 0 away from F's closure);
// nodes under F have no dF node around them
template <class T>
static __global__ void __launch_bounds__(256) amr_coarse_rhs_kernel(
    KGeom gc, const unsigned char* __restrict__ cdof, const T* __restrict__ rc, const T* __restrict__ w,
    T* __restrict__ bC, const int4* tl = nullptr) {
    int I, J, K;
    if (!grid_point_tl(tl, gc.nn[0], gc.nn[1], I, J, K)) return;
    const long id = I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K);
    if (is_dir(gc, I, J, K)) { bC[id] = 0; return; }
    bC[id] = (T)((cdof[id] ? (double)rc[id] : 0.0) + (double)w[id]);
}

// One Jacobi sweep on the coarse unknowns of the composite correction
// equation, after the fine post-smoothing: with yc = L_c,out ec (coarse
// elements outside F) and w = R(yf on dF) (yf = L_f e, the fine correction,
// dF following the coarse one), t = r_c - (yc + w) and
// ec += omega t / diag, diag of the coarse operator with the full coarse
// sigma (near dF the composite diagonal differs by up to ~20%, which a damped
// Jacobi sweep tolerates).  Smoothing the coarse unknowns next to F with the
// composite operator is what the coarse correction alone leaves undone.
template <class T>
static __global__ void __launch_bounds__(256) amr_coarse_jacobi_kernel(
    KGeom gc, const unsigned char* __restrict__ cdof, const T* __restrict__ rc, const T* __restrict__ yc,
    const T* __restrict__ w, const T* __restrict__ sig, double diag4, double omega, T* __restrict__ ec,
    const int4* tl = nullptr) {
    int I, J, K;
    if (!grid_point_tl(tl, gc.nn[0], gc.nn[1], I, J, K)) return;
    const long id = I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K);
    if (!cdof[id] || is_dir(gc, I, J, K)) return;
    const double acc = (double)yc[id] + (double)w[id];
    double ss = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        ss += (double)sig[(I + (q & 1)) + (long)gc.sg[0] * ((J + ((q >> 1) & 1)) + (long)gc.sg[1] * (K + ((q >> 2) & 1)))];
    const double diag = -diag4 * ss;
    if (diag != 0.0) ec[id] = (T)((double)ec[id] + omega * ((double)rc[id] - acc) / diag);
}

// dst += src where mask (and not a Dirichlet node of g)
template <class T>
static __global__ void __launch_bounds__(256) amr_add_kernel(KGeom g, const unsigned char* __restrict__ mask,
                                                             unsigned char want, const T* __restrict__ src,
                                                             double* __restrict__ dst,
                                                             const int4* tl = nullptr) {
    int i, j, k;
    if (!grid_point_tl(tl, g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    if (mask[id] == want && !is_dir(g, i, j, k)) dst[id] += (double)src[id];
}

// v = v where mask == want, else 0
static __global__ void __launch_bounds__(256) amr_select_kernel(KGeom g, const unsigned char* __restrict__ mask,
                                                                unsigned char want, double* __restrict__ v) {
    int i, j, k;
    if (!grid_point(g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    if (mask[id] != want) v[id] = 0.0;
}

// per-block sums of v over the nodes whose mask equals want (mask NULL: all),
// in a fixed order (deterministic)
static __global__ void __launch_bounds__(256) amr_sum_kernel(const double* __restrict__ v,
                                                             const unsigned char* __restrict__ mask,
                                                             unsigned char want, long n,
                                                             double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256)
        if (!mask || mask[q] == want) s += v[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[0] = (wa * sum(pa) + wb * sum(pb)) / count  (one thread, fixed order)
static __global__ void amr_mean_kernel(const double* __restrict__ pa, int na, double wa,
                                       const double* __restrict__ pb, int nb, double wb,
                                       double count, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int q = 0; q < na; ++q) a += pa[q];
    for (int q = 0; q < nb; ++q) b += pb[q];
    out[0] = (wa * a + wb * b) / count;
}

// out[0] = sum_l 8^(nparts-1-l) sum(part[l np .. (l+1) np)) / count: the
// weighted mean of a right-hand side over the unknowns of nparts levels in the
// symmetric scaling (a level's rows times 8 per level below the finest), one
// thread, fixed order
static __global__ void amr_mean_levels_kernel(const double* __restrict__ part, int np, int nparts,
                                              double count, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = 0.0, w = 1.0;
    for (int l = nparts - 1; l >= 0; --l) {
        double a = 0.0;
        for (int q = 0; q < np; ++q) a += part[(long)l * np + q];
        acc += w * a;
        w *= 8.0;
    }
    out[0] = acc / count;
}

// v -= f * mu[0] where mask == want (mask NULL: everywhere), or where mask != want when invert
static __global__ void __launch_bounds__(256) amr_shift_kernel(double* __restrict__ v,
                                                               const unsigned char* __restrict__ mask,
                                                               unsigned char want, int invert, long n,
                                                               const double* __restrict__ mu, double f) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    if (mask && ((mask[q] == want) == (invert != 0))) return;
    v[q] -= f * mu[0];
}

// coarse nodes under F (not unknowns) take the coincident fine value; ftc:
// the coarse level's own node types when it is itself a fine level (only its
// AMR_DOF nodes can lie under F)
static __global__ void __launch_bounds__(256) amr_inject_kernel(KGeom gf, KGeom gc,
                                                                const unsigned char* __restrict__ cdof,
                                                                const double* __restrict__ xf,
                                                                double* __restrict__ xc,
                                                                const unsigned char* __restrict__ ftc = nullptr,
                                                                const int4* tl = nullptr) {
    int I, J, K;
    if (!grid_point_tl(tl, gc.nn[0], gc.nn[1], I, J, K)) return;
    const long id = I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K);
    if (cdof[id] || (ftc && ftc[id] != AMR_DOF)) return;
    xc[id] = xf[2 * I + (long)gf.nn[0] * (2 * J + (long)gf.nn[1] * 2 * K)];
}

// coarse cell coefficient with covered cells set to 0 (ghost images included)
static __global__ void __launch_bounds__(256) amr_sigma_out_kernel(KGeom gc, const unsigned char* __restrict__ cov,
                                                                   const double* __restrict__ sig,
                                                                   double* __restrict__ out) {
    int i, j, k;
    if (!grid_point(gc.n[0], gc.n[1], i, j, k)) return;
    const long c = i + (long)gc.n[0] * (j + (long)gc.n[1] * k);
    const double v = cov[c] ? 0.0 : sig[(i + 1) + (long)gc.sg[0] * ((j + 1) + (long)gc.sg[1] * (k + 1))];
    write_cell_images(out, gc, i, j, k, v);
}

// covered coarse cells (ghost images included) of a ghosted 3-comp velocity -> 0
static __global__ void __launch_bounds__(256) amr_zero_covered_kernel(KGeom gc, const unsigned char* __restrict__ cov,
                                                                      double* __restrict__ ug) {
    int i, j, k;
    if (!grid_point(gc.n[0], gc.n[1], i, j, k)) return;
    if (!cov[i + (long)gc.n[0] * (j + (long)gc.n[1] * k)]) return;
    const long ug3 = (long)gc.sg[0] * gc.sg[1] * gc.sg[2];
    for (int c = 0; c < 3; ++c) write_cell_images(ug + c * ug3, gc, i, j, k, 0.0);
}

// average_down onto covered coarse cells of the caller's coarse boxes:
// u_c = mean of the 8 projected fine children, gp_c = mean of their G phi
// (src/projection/incflo_apply_nodal_projection.cpp:258-266), read from the
// finer level's canonical copies (written by its update, and by its own
// average-down where a still finer level lies under it); the coarse level's
// canonical copies are written too when a still coarser level follows.
struct AvgDownArgs {
    KGeom gc, gf;
    const BoxInfo* boxes;   // coarse layout
    const Tile* tiles;
    const unsigned char* cov;
    const double* uf;       // canonical ghosted fine u_new (3 comps)
    const double* gf_;      // canonical ghosted fine G phi (3 comps; NULL: no gp)
    double* vel; int vel_ng;
    double* gp; int gp_ng;
    double* uc;             // canonical ghosted coarse u_new / G phi (NULL: not kept)
    double* gc_;
};

static __global__ void __launch_bounds__(256) amr_avgdown_kernel(AvgDownArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    if (!a.cov[i + (long)a.gc.n[0] * (j + (long)a.gc.n[1] * k)]) return;
    const BoxInfo& b = a.boxes[bi];
    const KGeom& gf = a.gf;
    const long f3 = (long)gf.sg[0] * gf.sg[1] * gf.sg[2];
    const long c3 = (long)a.gc.sg[0] * a.gc.sg[1] * a.gc.sg[2];
    const long ci = (i + 1) + (long)a.gc.sg[0] * ((j + 1) + (long)a.gc.sg[1] * (k + 1));
    double us[3] = {0.0, 0.0, 0.0}, gs[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 8; ++q) {
        const int fi = 2 * i + (q & 1), fj = 2 * j + ((q >> 1) & 1), fk = 2 * k + ((q >> 2) & 1);
        const long gi = (fi + 1) + (long)gf.sg[0] * ((fj + 1) + (long)gf.sg[1] * (fk + 1));
        for (int c = 0; c < 3; ++c) {
            us[c] += a.uf[c * f3 + gi];
            if (a.gf_) gs[c] += a.gf_[c * f3 + gi];
        }
    }
    for (int c = 0; c < 3; ++c) {
        us[c] *= 0.125;
        gs[c] *= 0.125;
    }
    if (a.vel) {
        const long off = field_offset(b, a.vel_ng, 3, false);
        for (int c = 0; c < 3; ++c) a.vel[off + box_index(b, a.vel_ng, false, i, j, k, c)] = us[c];
    }
    if (a.gp && a.gf_) {
        const long off = field_offset(b, a.gp_ng, 3, false);
        for (int c = 0; c < 3; ++c) a.gp[off + box_index(b, a.gp_ng, false, i, j, k, c)] = gs[c];
    }
    if (a.uc)
        for (int c = 0; c < 3; ++c) a.uc[c * c3 + ci] = us[c];
    if (a.gc_ && a.gf_)
        for (int c = 0; c < 3; ++c) a.gc_[c * c3 + ci] = gs[c];
}

}  // namespace mgx

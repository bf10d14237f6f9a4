// amr_mac.hip -- composite MAC projection over AMR levels (refinement ratio 2
// between neighbours) and its C ABI.
//
// incflo projects the MAC velocities of all AMR levels with one MacProjector
// (src/convection/incflo_compute_MAC_projected_velocities.cpp:96-133, 272-300)
// and the projector leaves the fine faces averaged down onto the coarse level
// (:301-302).  The composite problem here is a conservative finite-volume
// discretisation on the composite mesh:
//   unknowns: each level's cells not covered by the next finer level (level 0
//   covers the domain);
//   every face between two unknowns of one level carries the two-point flux
//     beta * (phi_j - phi_i) / distance between the cell centres,
//   a face on a level's boundary links its cell to the cell of the level
//   below behind it at distance 1.5 h, so a cell of the level below next to
//   the finer level sees the 4 finer fluxes on each such face (reflux);
//   right-hand side: div u per unknown, faces under a finer level taking the
//   mean of its faces (from the finest level down).
// The volume-weighted operator is symmetric.  Solver: composite V-cycles
// inside an FP64 outer iteration on the composite residual (MLMG's stopping
// rule ||r||_inf <= max(atol, rtol ||b||_inf)), recursing from the finest
// level: red-black Gauss-Seidel on a level's correction with the level
// below's at 0, the residual restricted to the level below (cells under the
// level: the mean of their 8 residuals), the level below's correction (level
// 0: one V-cycle of its MAC hierarchy), piecewise-constant prolongation,
// red-black Gauss-Seidel against it, then one red-black Gauss-Seidel sweep of
// the level below's unknowns with the composite rows.  MGX_MODE_MIXED runs
// the V-cycle in FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "amr_kernels.h"
#include "mac_mg.h"
#include "mgx_ctx.h"

namespace mgx {

template <class T>
struct AmacCoef {
    T ih2[3];   // 1 / h^2 of the level
    T ihf2[3];  // 1 / h^2 of the next finer level
};

__device__ __forceinline__ int wrap_idx(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

enum { AM_RESID = 0, AM_GSRB = 1 };

// One level's rows of the composite operator.
//   fm:  the level's cells (NULL: every cell, level 0)
//   cov: its cells under the next finer level (NULL: none)
//   comp = 1, composite rows: covered cells are no unknowns, and a face to a
//        covered neighbour links to the 4 finer cells next to it (reflux,
//        values xf, finer coefficients bfx..bfz);
//   comp = 0, the level's correction equation: every cell of fm is an
//        unknown and the finer level is ignored.
//   A face to a cell outside fm links to the level below's cell behind it at
//   distance 1.5 h (values xc; NULL: 0).  Domain faces: Dirichlet (value 0 at
//   h/2) or Neumann.  x NULL reads 0.
// RESID: r = b - L x on the unknowns (0 elsewhere), ||r||_inf into norm.
// GSRB:  one colour of red-black Gauss-Seidel on x, in place.
// tl: the 16 x 16-cell blocks of one plane that hold cells of fm (levels
// above 0; no other cell is visited, and none of them is read by the passes
// of that level), NULL: every cell on the dense 64 x 4 grid (level 0)
// FM / CV: the level has a region mask / a finer level (compile-time, so that
// level 0 and the finest level carry no mask tests)
template <class T, class TX, class TO, int MODE, bool FM, bool CV>
static __global__ void amac_level_kernel(MGeomK g, MGeomK gc, MGeomK gf, AmacCoef<T> c,
                                         const unsigned char* __restrict__ fm,
                                         const unsigned char* __restrict__ cov, int comp,
                                         const T* __restrict__ bx, const T* __restrict__ by,
                                         const T* __restrict__ bz, const T* __restrict__ bfx,
                                         const T* __restrict__ bfy, const T* __restrict__ bfz, TX* x,
                                         const TX* __restrict__ xc, const TX* __restrict__ xf,
                                         const T* __restrict__ b, TO* __restrict__ r,
                                         unsigned long long* norm, int color,
                                         const int4* __restrict__ tl) {
    __shared__ double red[4];
    int i, j, k;
    double m = 0.0;
    bool in;
    if (MODE == AM_GSRB) {
        // the threads cover the cells of one colour: a block of blockDim.x
        // threads spans 2 blockDim.x cells of a row (launched with half the x
        // threads of the other passes)
        int i0;
        if (tl) {
            const int4 t = tl[blockIdx.x];
            i0 = t.x * 2 * (int)blockDim.x;
            j = t.y * (int)blockDim.y + threadIdx.y;
            k = t.z;
        } else {
            i0 = blockIdx.x * 2 * (int)blockDim.x;
            j = blockIdx.y * (int)blockDim.y + threadIdx.y;
            k = blockIdx.z;
        }
        i = i0 + 2 * (int)threadIdx.x + ((j + k + color) & 1);  // i0 is even
        in = i < g.n[0] && j < g.n[1];
    } else {
        in = grid_point_tl(tl, g.n[0], g.n[1], i, j, k);
    }
    if (in) {
        const long id = cidx(g, i, j, k);
        const bool unk = (!FM || fm[id]) && !(CV && comp && cov[id]);
        const bool mine = MODE != AM_GSRB || (((i + j + k + color) & 1) == 0);
        if (!unk) {
            if (MODE == AM_RESID) r[id] = (TO)0;
        } else if (mine) {
            const int p[3] = {i, j, k};
            const T* beta[3] = {bx, by, bz};
            const T* betaf[3] = {bfx, bfy, bfz};
            const T xp = x ? (T)x[id] : (T)0;
            T acc = 0, dg = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    int q[3] = {p[0], p[1], p[2]};
                    q[d] += s ? 1 : -1;
                    int f[3] = {p[0], p[1], p[2]};
                    if (s) f[d] += 1;
                    const T bet = beta[d][fidx(g, d, f[0], f[1], f[2])];
                    if (q[d] < 0 || q[d] >= g.n[d]) {
                        if (!g.per[d]) {
                            if (s ? g.dhi[d] : g.dlo[d]) {
                                acc -= 2 * c.ih2[d] * bet * xp;
                                dg -= 2 * c.ih2[d] * bet;
                            }
                            continue;
                        }
                        q[d] = wrap_idx(q[d], g.n[d]);
                    }
                    const long qid = cidx(g, q[0], q[1], q[2]);
                    if (FM && !fm[qid]) {
                        // coarse-fine face: the cell of the level below behind it
                        const T t = bet * c.ih2[d] * (T)(2.0 / 3.0);
                        const T xq = xc ? (T)xc[cidx(gc, q[0] >> 1, q[1] >> 1, q[2] >> 1)] : (T)0;
                        acc += t * (xq - xp);
                        dg -= t;
                    } else if (CV && comp && cov[qid]) {
                        // the 4 finer cells of q next to the face, each
                        // through its own face (the one the finer rows use)
                        const int fd = s ? 2 * q[d] : 2 * q[d] + 1;
                        const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
#pragma unroll
                        for (int a = 0; a < 2; ++a)
#pragma unroll
                            for (int e = 0; e < 2; ++e) {
                                int ff[3];
                                ff[d] = fd;
                                ff[t0] = 2 * p[t0] + a;
                                ff[t1] = 2 * p[t1] + e;
                                int fa[3] = {ff[0], ff[1], ff[2]};
                                if (!s) fa[d] += 1;  // hi face of the finer cell (a periodic
                                                     // seam's face n_d holds face 0's value)
                                const T bf = betaf[d][fidx(gf, d, fa[0], fa[1], fa[2])];
                                const T t = bf * c.ihf2[d] * (T)(1.0 / 12.0);
                                const T xq = xf ? (T)xf[cidx(gf, ff[0], ff[1], ff[2])] : (T)0;
                                acc += t * (xq - xp);
                                dg -= t;
                            }
                    } else {
                        const T t = c.ih2[d] * bet;
                        acc += t * ((x ? (T)x[qid] : (T)0) - xp);
                        dg -= t;
                    }
                }
            if (MODE == AM_GSRB) {
                if (dg != (T)0) x[id] = (TX)(xp + (b[id] - acc) / dg);
            } else {
                const T rv = b[id] - acc;
                r[id] = (TO)rv;
                m = nabs((double)rv);
            }
        }
    }
    if (MODE == AM_RESID && norm) {
        const double bm = block_max<256>(m, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_nonneg(norm, bm);
    }
}

// coarse right-hand side of the correction: covered cells the mean of their
// 8 fine residuals, the others their own residual
template <class T>
static __global__ void amac_restrict_kernel(MGeomK gc, MGeomK gf, const unsigned char* __restrict__ cov,
                                            const T* __restrict__ rc, const T* __restrict__ rf,
                                            T* __restrict__ out) {
    int I, J, K;
    if (!grid_point(gc.n[0], gc.n[1], I, J, K)) return;
    const long id = cidx(gc, I, J, K);
    if (!cov[id]) {
        out[id] = rc[id];
        return;
    }
    T s = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        s += rf[cidx(gf, 2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + ((q >> 2) & 1))];
    out[id] = (T)0.125 * s;
}

// e_f += e_c (parent) on F (tl: F's blocks, as for amac_level_kernel)
template <class T>
static __global__ void amac_prolong_kernel(MGeomK gf, MGeomK gc, const unsigned char* __restrict__ fm,
                                           const T* __restrict__ ec, T* __restrict__ ef,
                                           const int4* __restrict__ tl) {
    int i, j, k;
    if (!grid_point_tl(tl, gf.n[0], gf.n[1], i, j, k)) return;
    const long id = cidx(gf, i, j, k);
    if (fm[id]) ef[id] += ec[cidx(gc, i >> 1, j >> 1, k >> 1)];
}

// x += e on the unknowns of a level above 0, over its blocks (tl)
template <class T>
static __global__ void amac_add_tl_kernel(MGeomK g, const unsigned char* __restrict__ unk,
                                          const T* __restrict__ e, double* __restrict__ x,
                                          const int4* __restrict__ tl) {
    int i, j, k;
    if (!grid_point_tl(tl, g.n[0], g.n[1], i, j, k)) return;
    const long id = cidx(g, i, j, k);
    if (unk[id]) x[id] += (double)e[id];
}

// part[blockIdx.x] = sum of v over the unknowns (unk) of a level above 0,
// its blocks (tl, nb of them) taken in a grid-stride loop; the partials are
// those of amr_sum_kernel (one per workgroup, fixed order)
static __global__ void __launch_bounds__(256) amac_sum_tl_kernel(MGeomK g, const double* __restrict__ v,
                                                                 const unsigned char* __restrict__ unk,
                                                                 const int4* __restrict__ tl, int nb,
                                                                 double* __restrict__ part) {
    __shared__ double red[256];
    const int tid = threadIdx.x + 16 * threadIdx.y;
    double s = 0.0;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int4 t = tl[b];
        const int i = t.x * 16 + threadIdx.x, j = t.y * 16 + threadIdx.y;
        if (i < g.n[0] && j < g.n[1]) {
            const long id = cidx(g, i, j, t.z);
            if (unk[id]) s += v[id];
        }
    }
    red[tid] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = red[0];
}

// v -= f mu[0] on the unknowns of a level above 0, over its blocks (tl)
static __global__ void amac_shift_tl_kernel(MGeomK g, double* __restrict__ v, const unsigned char* __restrict__ unk,
                                            const double* __restrict__ mu, double f, const int4* __restrict__ tl) {
    int i, j, k;
    if (!grid_point_tl(tl, g.n[0], g.n[1], i, j, k)) return;
    const long id = cidx(g, i, j, k);
    if (unk[id]) v[id] -= f * mu[0];
}

// e = 0 on a level's blocks (tl); cells outside them are never written
template <class T>
static __global__ void amac_zero_tl_kernel(MGeomK g, T* __restrict__ e, const int4* __restrict__ tl) {
    int i, j, k;
    if (!grid_point_tl(tl, g.n[0], g.n[1], i, j, k)) return;
    e[cidx(g, i, j, k)] = (T)0;
}

// x += e where mask == want
template <class T>
static __global__ void amac_add_kernel(const unsigned char* __restrict__ mask, unsigned char want,
                                       const T* __restrict__ e, double* __restrict__ x, long n) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < n && mask[q] == want) x[q] += (double)e[q];
}

// covered coarse cells <- mean of their 8 fine cells
static __global__ void amac_cover_mean_kernel(MGeomK gc, MGeomK gf, const unsigned char* __restrict__ cov,
                                              const double* __restrict__ xf, double* __restrict__ xc) {
    int I, J, K;
    if (!grid_point(gc.n[0], gc.n[1], I, J, K)) return;
    const long id = cidx(gc, I, J, K);
    if (!cov[id]) return;
    double s = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        s += xf[cidx(gf, 2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + ((q >> 2) & 1))];
    xc[id] = 0.125 * s;
}

// faces of direction d of a level above 0 with a cell of the level (fm) on
// either side: u -= beta grad phi between two of its cells, the coarse-fine
// flux to the cell of the level below behind the face (distance 1.5 h), or
// at a Dirichlet domain face 2 beta phi / h; other faces keep u
static __global__ void amac_fine_faces_kernel(MGeomK gf, MGeomK gc, int d,
                                              const unsigned char* __restrict__ fm,
                                              const double* __restrict__ beta,
                                              const double* __restrict__ xf,
                                              const double* __restrict__ xc, double ih,
                                              double* __restrict__ u) {
    const int f0 = gf.n[0] + (d == 0), f1 = gf.n[1] + (d == 1), f2 = gf.n[2] + (d == 2);
    int i, j, k;
    if (!grid_point(f0, f1, i, j, k)) return;
    if (k >= f2) return;
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    lo[d] -= 1;
    const long fid = fidx(gf, d, i, j, k);
    if (lo[d] < 0 || hi[d] >= gf.n[d]) {
        if (!gf.per[d]) {
            // a domain face: only a Dirichlet face of a cell of the level changes
            const bool low = lo[d] < 0;
            const int* in = low ? hi : lo;
            const long iid = cidx(gf, in[0], in[1], in[2]);
            if (!fm[iid] || !(low ? gf.dlo[d] : gf.dhi[d])) return;
            const double g = (low ? xf[iid] : -xf[iid]) * 2.0 * ih;
            u[fid] -= beta[fid] * g;
            return;
        }
        lo[d] = wrap_idx(lo[d], gf.n[d]);
        hi[d] = wrap_idx(hi[d], gf.n[d]);
    }
    const long il = cidx(gf, lo[0], lo[1], lo[2]), ihh = cidx(gf, hi[0], hi[1], hi[2]);
    const bool fl = fm[il] != 0, fh = fm[ihh] != 0;
    if (!fl && !fh) return;
    const double bet = beta[fid];
    double g;
    if (fl && fh) g = (xf[ihh] - xf[il]) * ih;
    else if (fl) g = (xc[cidx(gc, hi[0] >> 1, hi[1] >> 1, hi[2] >> 1)] - xf[il]) * ih * (2.0 / 3.0);
    else g = (xf[ihh] - xc[cidx(gc, lo[0] >> 1, lo[1] >> 1, lo[2] >> 1)]) * ih * (2.0 / 3.0);
    u[fid] -= bet * g;
}

// coarse faces of direction d under the fine level (a covered cell on either
// side) <- mean of the 4 fine faces on them
static __global__ void amac_avgdown_faces_kernel(MGeomK gc, MGeomK gf, int d,
                                                 const unsigned char* __restrict__ cov,
                                                 const double* __restrict__ uf,
                                                 double* __restrict__ uc) {
    const int f0 = gc.n[0] + (d == 0), f1 = gc.n[1] + (d == 1);
    int I, J, K;
    if (!grid_point(f0, f1, I, J, K)) return;
    int lo[3] = {I, J, K}, hi[3] = {I, J, K};
    lo[d] -= 1;
    bool under = false;
    if (lo[d] >= 0 || gc.per[d]) under = under || cov[cidx(gc, wrap_idx(lo[0], gc.n[0]), wrap_idx(lo[1], gc.n[1]),
                                                         wrap_idx(lo[2], gc.n[2]))];
    if (hi[d] < gc.n[d] || gc.per[d]) under = under || cov[cidx(gc, wrap_idx(hi[0], gc.n[0]), wrap_idx(hi[1], gc.n[1]),
                                                              wrap_idx(hi[2], gc.n[2]))];
    if (!under) return;
    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
    const int C[3] = {I, J, K};
    double s = 0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            int f[3];
            f[d] = 2 * C[d];
            f[t0] = 2 * C[t0] + a;
            f[t1] = 2 * C[t1] + e;
            s += uf[fidx(gf, d, f[0], f[1], f[2])];
        }
    uc[fidx(gc, d, I, J, K)] = 0.25 * s;
}

// periodic direction d: face n_d takes face 0's value
static __global__ void amac_periodic_face_kernel(MGeomK g, int d, double* __restrict__ u) {
    const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    const long np = (long)g.n[a0] * g.n[a1];
    if (q >= np) return;
    int f[3];
    f[a0] = (int)(q % g.n[a0]);
    f[a1] = (int)(q / g.n[a0]);
    f[d] = 0;
    const long i0 = fidx(g, d, f[0], f[1], f[2]);
    f[d] = g.n[d];
    u[fidx(g, d, f[0], f[1], f[2])] = u[i0];
}

// periodic direction d of the fine level: a face on the seam is held by the
// box of the fine cell next to it, index 0 when cell 0 is in F, else index n_d;
// both entries take the held value
static __global__ void amac_periodic_merge_kernel(MGeomK g, int d, const unsigned char* __restrict__ fm,
                                                  double* __restrict__ u) {
    const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    const long np = (long)g.n[a0] * g.n[a1];
    if (q >= np) return;
    int f[3];
    f[a0] = (int)(q % g.n[a0]);
    f[a1] = (int)(q / g.n[a0]);
    f[d] = 0;
    const long i0 = fidx(g, d, f[0], f[1], f[2]);
    const bool held0 = fm[cidx(g, f[0], f[1], f[2])] != 0;
    f[d] = g.n[d];
    const long in = fidx(g, d, f[0], f[1], f[2]);
    const double v = held0 ? u[i0] : u[in];
    u[i0] = v;
    u[in] = v;
}

static int amac_nlevels(const MacGeom& g, int max_coarsening_level) {
    NodalGeom ng{};
    for (int d = 0; d < 3; ++d) ng.n[d] = g.n[d];
    return nodal_nlevels(ng, max_coarsening_level);
}

// one level of the composite MAC problem (dense canonical arrays at its
// resolution; level 0 covers the domain)
struct AmacLev {
    MacGeom g{};
    bool finer = false;              // a finer level lies above
    DevBuf fm, cov, unk;             // cells of the level (l >= 1), under the finer level, unknowns
    DevBuf beta[3], beta32[3];
    DevBuf x, b, r, r2, e;           // FP64 iterate, right-hand side, V-cycle work (FP64-sized)
    DevBuf R;                        // rhs of the level's correction (l >= 1; FP64-sized)
    TileSet tiles;                   // blocks holding cells of the level (l >= 1)
    long nunk = 0;
    template <class T> const T* betaT(int d) const {
        return std::is_same<T, float>::value ? (const T*)beta32[d].p : (const T*)beta[d].p;
    }
};

class AmrMac {
   public:
    MGOpts opts;
    int nlev = 1;
    MacHierarchy<double> H64;
    MacHierarchy<float> H32;
    std::vector<std::unique_ptr<AmacLev>> lev;  // level 0 first
    DevBuf hb;                       // level-0 right-hand side of the hierarchy
    DevBuf part, scal, ibuf;
    unsigned long long* h_pinned = nullptr;
    double ndof = 0.0;               // unknowns in units of finest cells
    bool beta_set = false, h64_beta = false, h32_beta = false;

    ~AmrMac() {
        if (h_pinned) (void)hipHostFree(h_pinned);
    }
    int L() const { return (int)lev.size() - 1; }
    AmacLev& at(int l) { return *lev[l]; }

    // lboxes[l - 1]: level l's boxes (6 ints each, level-l cell indices)
    void setup(const MacGeom& g0, const std::vector<std::vector<int>>& lboxes, const MGOpts& o) {
        // the boxes and the periodicity are fixed per handle, so a set-up for
        // new boundary types or options keeps the coefficient arrays
        const int L = (int)lboxes.size();
        const bool keep = beta_set && (int)lev.size() == L + 1;
        if ((int)lev.size() != L + 1) {
            lev.clear();
            for (int l = 0; l <= L; ++l) lev.emplace_back(new AmacLev());
        }
        opts = o;
        nlev = amac_nlevels(g0, o.max_coarsening_level);
        std::vector<std::vector<unsigned char>> fm(L + 1), cov(L + 1);
        for (int l = 0; l <= L; ++l) {
            MacGeom& g = at(l).g;
            g = g0;
            for (int d = 0; d < 3; ++d) {
                g.n[d] = g0.n[d] << l;
                g.dx[d] = g0.dx[d] / (double)(1 << l);
            }
            at(l).finer = l < L;
            fm[l].assign(g.ncell(), l == 0 ? 1 : 0);
            if (l > 0) {
                const std::vector<int>& bx = lboxes[l - 1];
                for (size_t b = 0; b + 5 < bx.size(); b += 6)
                    for (int k = bx[b + 2]; k <= bx[b + 5]; ++k)
                        for (int j = bx[b + 1]; j <= bx[b + 4]; ++j)
                            for (int i = bx[b + 0]; i <= bx[b + 3]; ++i)
                                fm[l][i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k)] = 1;
            }
        }
        ndof = 0.0;
        for (int l = 0; l <= L; ++l) {
            const MacGeom& g = at(l).g;
            cov[l].assign(g.ncell(), 0);
            if (l < L) {
                const MacGeom& gf = at(l + 1).g;
                for (int k = 0; k < g.n[2]; ++k)
                    for (int j = 0; j < g.n[1]; ++j)
                        for (int i = 0; i < g.n[0]; ++i) {
                            const size_t id = i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
                            cov[l][id] = fm[l + 1][2 * i + (size_t)gf.n[0] * (2 * j + (size_t)gf.n[1] * 2 * k)];
                            MGX_REQUIRE(!cov[l][id] || fm[l][id], MGX_ERR_ARG,
                                        "AMR MAC projector: a level reaches outside the level below it");
                        }
                // proper nesting: every cell around a covered one lies in the level
                if (l > 0)
                    for (int k = 0; k < g.n[2]; ++k)
                        for (int j = 0; j < g.n[1]; ++j)
                            for (int i = 0; i < g.n[0]; ++i) {
                                if (!cov[l][i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k)]) continue;
                                for (int dk = -1; dk <= 1; ++dk)
                                    for (int dj = -1; dj <= 1; ++dj)
                                        for (int di = -1; di <= 1; ++di) {
                                            int q[3] = {i + di, j + dj, k + dk};
                                            bool in = true;
                                            for (int d = 0; d < 3; ++d) {
                                                if (g.per[d]) q[d] = (q[d] + g.n[d]) % g.n[d];
                                                else if (q[d] < 0 || q[d] >= g.n[d]) in = false;
                                            }
                                            MGX_REQUIRE(!in || fm[l][q[0] + (size_t)g.n[0] * (q[1] + (size_t)g.n[1] * q[2])],
                                                        MGX_ERR_ARG,
                                                        "AMR MAC projector: a level is not properly nested in "
                                                        "the level below it");
                                        }
                            }
            }
            std::vector<unsigned char> unk(g.ncell());
            long cnt = 0;
            for (size_t q = 0; q < unk.size(); ++q) {
                unk[q] = fm[l][q] && !cov[l][q];
                cnt += unk[q];
            }
            AmacLev& v = at(l);
            v.nunk = cnt;
            ndof += (double)cnt * std::pow(8.0, L - l);
            if (l > 0) {
                v.fm.alloc(fm[l].size());
                MGX_HIP(hipMemcpy(v.fm.p, fm[l].data(), fm[l].size(), hipMemcpyHostToDevice));
                v.tiles.build(g.n, fm[l]);
            }
            if (l < L) {
                v.cov.alloc(cov[l].size());
                MGX_HIP(hipMemcpy(v.cov.p, cov[l].data(), cov[l].size(), hipMemcpyHostToDevice));
            }
            v.unk.alloc(unk.size());
            MGX_HIP(hipMemcpy(v.unk.p, unk.data(), unk.size(), hipMemcpyHostToDevice));
            for (int d = 0; d < 3; ++d) {
                v.beta[d].alloc(g.nface(d) * sizeof(double));
                v.beta32[d].alloc(g.nface(d) * sizeof(float));
                if (!keep) MGX_HIP(hipMemset(v.beta[d].p, 0, g.nface(d) * sizeof(double)));
            }
            for (DevBuf* bb : {&v.x, &v.b, &v.r, &v.r2, &v.e, &v.R}) {
                bb->alloc(g.ncell() * sizeof(double));
                MGX_HIP(hipMemset(bb->p, 0, g.ncell() * sizeof(double)));
            }
        }
        scal.alloc(64 * sizeof(double));
        part.alloc(4096 * sizeof(double));
        ibuf.alloc(16 * sizeof(int));
        if (!h_pinned) MGX_HIP(hipHostMalloc((void**)&h_pinned, 64 * sizeof(unsigned long long), 0));
        H64.built = H32.built = false;
        beta_set = keep;
        h64_beta = h32_beta = false;
    }

    // the levels' beta (FP64 canonical faces) were filled by the caller
    void beta_changed(hipStream_t s) {
        for (auto& v : lev)
            for (int d = 0; d < 3; ++d)
                cast_kernel<float, double><<<grid1d(v->g.nface(d)), 256, 0, s>>>(
                    v->beta[d].as<double>(), v->beta32[d].as<float>(), v->g.nface(d));
        beta_set = true;
        h64_beta = h32_beta = false;
    }

    void zero_norm(int slot, hipStream_t s) {
        zero_u64_kernel<<<1, 64, 0, s>>>(scal.as<unsigned long long>() + slot, 1);
    }
    double read_norm(int slot, hipStream_t s) {
        MGX_HIP(hipMemcpyAsync(h_pinned + slot, scal.as<unsigned long long>() + slot,
                               sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        MGX_HIP(hipStreamSynchronize(s));
        double v;
        std::memcpy(&v, &h_pinned[slot], sizeof(v));
        return v;
    }

    template <class T>
    AmacCoef<T> coef(int l) const {
        AmacCoef<T> c;
        const MacGeom& g = lev[l]->g;
        for (int d = 0; d < 3; ++d) {
            c.ih2[d] = (T)(1.0 / (g.dx[d] * g.dx[d]));
            c.ihf2[d] = (T)(4.0 / (g.dx[d] * g.dx[d]));
        }
        return c;
    }

    // the neighbours' geometry (itself where there is none)
    MGeomK kcs(int l) const { return (l > 0 ? lev[l - 1] : lev[l])->g.k(); }
    MGeomK kfn(int l) const { return (l < L() ? lev[l + 1] : lev[l])->g.k(); }

    // one launch of the level kernel on level l
    template <class T, class TX, class TO, int MODE>
    void level_pass(int l, int comp, TX* x, const TX* xc, const TX* xf, const T* b, TO* r,
                    unsigned long long* norm, int color, hipStream_t s) {
        AmacLev& v = at(l);
        const T* bf[3] = {nullptr, nullptr, nullptr};
        if (l < L())
            for (int d = 0; d < 3; ++d) bf[d] = at(l + 1).betaT<T>(d);
        dim3 gr = grid3(v.g.n[0], v.g.n[1], v.g.n[2]), blk = block3();
        const int4* tl = nullptr;
        if (l > 0) {
            // only the blocks that hold cells of the level
            int nb = 0;
            tl = v.tiles.blocks(nb);
            if (nb == 0) return;
            gr = dim3(nb);
            blk = dim3(TileSet::kTX, TileSet::kTY);
        }
        if (MODE == AM_GSRB) blk.x /= 2;  // one thread per cell of the colour
        const unsigned char* fm = l > 0 ? v.fm.as<unsigned char>() : nullptr;
        const unsigned char* cv = l < L() ? v.cov.as<unsigned char>() : nullptr;
#define MGX_AMAC_LAUNCH(FMB, CVB)                                                                 \
    amac_level_kernel<T, TX, TO, MODE, FMB, CVB><<<gr, blk, 0, s>>>(                              \
        v.g.k(), kcs(l), kfn(l), coef<T>(l), fm, cv, comp, v.betaT<T>(0), v.betaT<T>(1),         \
        v.betaT<T>(2), bf[0], bf[1], bf[2], x, xc, xf, b, r, norm, color, tl)
        if (fm && cv) MGX_AMAC_LAUNCH(true, true);
        else if (fm) MGX_AMAC_LAUNCH(true, false);
        else if (cv) MGX_AMAC_LAUNCH(false, true);
        else MGX_AMAC_LAUNCH(false, false);
#undef MGX_AMAC_LAUNCH
    }

    // v_l -= the weighted mean over the unknowns of every level (slot: where
    // the mean lands in scal); vals[l] a per-level canonical cell array
    void remove_mean(const std::vector<double*>& vals, int slot, hipStream_t s) {
        const int np = 512;
        // levels above 0: over their blocks only (their unknowns lie in them)
        for (int l = 0; l <= L(); ++l) {
            int nb = 0;
            const int4* tl = l > 0 ? at(l).tiles.blocks(nb) : nullptr;
            if (l > 0)
                amac_sum_tl_kernel<<<np, dim3(TileSet::kTX, TileSet::kTY), 0, s>>>(
                    at(l).g.k(), vals[l], at(l).unk.as<unsigned char>(), tl, nb, part.as<double>() + (long)l * np);
            else
                amr_sum_kernel<<<np, 256, 0, s>>>(vals[l], at(l).unk.as<unsigned char>(), 1, at(l).g.ncell(),
                                                  part.as<double>() + (long)l * np);
        }
        amr_mean_levels_kernel<<<1, 64, 0, s>>>(part.as<double>(), np, L() + 1, ndof, scal.as<double>() + slot);
        for (int l = 0; l <= L(); ++l) {
            if (l > 0) {
                int nb = 0;
                const int4* tl = at(l).tiles.blocks(nb);
                if (nb > 0)
                    amac_shift_tl_kernel<<<nb, dim3(TileSet::kTX, TileSet::kTY), 0, s>>>(
                        at(l).g.k(), vals[l], at(l).unk.as<unsigned char>(), scal.as<double>() + slot, 1.0, tl);
            } else {
                amr_shift_kernel<<<grid1d(at(l).g.ncell()), 256, 0, s>>>(vals[l], at(l).unk.as<unsigned char>(), 1,
                                                                         0, at(l).g.ncell(),
                                                                         scal.as<double>() + slot, 1.0);
            }
        }
    }

    // FP64 composite residual of the levels' x into their r (stored as TO); ||r||_inf
    template <class TO>
    double residual(hipStream_t s) {
        zero_norm(1, s);
        for (int l = 0; l <= L(); ++l)
            level_pass<double, double, TO, AM_RESID>(
                l, 1, at(l).x.as<double>(), l > 0 ? at(l - 1).x.as<double>() : nullptr,
                l < L() ? at(l + 1).x.as<double>() : nullptr, at(l).b.as<double>(), at(l).r.as<TO>(),
                scal.as<unsigned long long>() + 1, 0, s);
        return read_norm(1, s);
    }

    // the correction of level l >= 1 from its right-hand side R_l: red-black
    // Gauss-Seidel with the level below's correction at 0, the residual
    // restricted (cells under level l: the mean of their 8 residuals; the
    // level below's unknowns: its composite rows with this correction), the
    // level below's correction (level 0: one V-cycle of its hierarchy),
    // piecewise-constant prolongation, red-black Gauss-Seidel against it, and
    // one red-black Gauss-Seidel sweep of the level below's unknowns with the
    // composite rows (this level's values held)
    template <class T>
    void cycle_level(MacHierarchy<T>& H, int l, hipStream_t s) {
        AmacLev& v = at(l);
        AmacLev& c = at(l - 1);
        T* e = v.e.as<T>();
        const T* R = l == L() ? v.r.as<T>() : v.R.as<T>();
        {
            // e is read on the level's cells only, and nothing writes it elsewhere
            int nb = 0;
            const int4* tl = v.tiles.blocks(nb);
            if (nb > 0)
                amac_zero_tl_kernel<T><<<nb, dim3(TileSet::kTX, TileSet::kTY), 0, s>>>(v.g.k(), e, tl);
        }
        for (int q = 0; q < opts.nu1; ++q)
            for (int color = 0; color < 2; ++color)
                level_pass<T, T, T, AM_GSRB>(l, 0, e, nullptr, nullptr, R, nullptr, nullptr, color, s);
        level_pass<T, T, T, AM_RESID>(l, 0, e, nullptr, nullptr, R, v.r2.as<T>(), nullptr, 0, s);
        level_pass<T, T, T, AM_RESID>(l - 1, 1, nullptr, nullptr, e, c.r.as<T>(), c.r2.as<T>(), nullptr, 0, s);
        T* Rc = l == 1 ? H.L[0].b : c.R.as<T>();
        amac_restrict_kernel<T><<<grid3(c.g.n[0], c.g.n[1], c.g.n[2]), block3(), 0, s>>>(
            c.g.k(), v.g.k(), c.cov.as<unsigned char>(), c.r2.as<T>(), v.r2.as<T>(), Rc);
        T* ec;
        if (l == 1) {
            H.vcycle(opts, s, ibuf.as<int>());
            ec = H.L[0].x;
        } else {
            cycle_level<T>(H, l - 1, s);
            ec = c.e.as<T>();
        }
        {
            int nb = 0;
            const int4* tl = v.tiles.blocks(nb);
            if (nb > 0)
                amac_prolong_kernel<T><<<nb, dim3(TileSet::kTX, TileSet::kTY), 0, s>>>(
                    v.g.k(), c.g.k(), v.fm.as<unsigned char>(), ec, e, tl);
        }
        for (int q = 0; q < opts.nu2; ++q)
            for (int color = 0; color < 2; ++color)
                level_pass<T, T, T, AM_GSRB>(l, 0, e, ec, nullptr, R, nullptr, nullptr, color, s);
        const T* ecc = l >= 3 ? at(l - 2).e.as<T>() : (l == 2 ? H.L[0].x : nullptr);
        for (int color = 0; color < 2; ++color)
            level_pass<T, T, T, AM_GSRB>(l - 1, 1, ec, ecc, e, c.r.as<T>(), nullptr, nullptr, color, s);
    }

    // one composite V-cycle on the residuals in the levels' r; the
    // corrections are added to the levels' x
    template <class T>
    void vcycle(MacHierarchy<T>& H, hipStream_t s) {
        if (L() == 0) {
            MGX_HIP(hipMemcpyAsync(H.L[0].b, at(0).r.p, at(0).g.ncell() * sizeof(T), hipMemcpyDeviceToDevice, s));
            H.vcycle(opts, s, ibuf.as<int>());
        } else {
            cycle_level<T>(H, L(), s);
        }
        for (int l = 0; l <= L(); ++l) {
            const T* e = l == 0 ? H.L[0].x : at(l).e.as<T>();
            if (l == 0) {
                amac_add_kernel<T><<<grid1d(at(l).g.ncell()), 256, 0, s>>>(at(l).unk.as<unsigned char>(), 1, e,
                                                                            at(l).x.as<double>(), at(l).g.ncell());
            } else {
                int nb = 0;
                const int4* tl = at(l).tiles.blocks(nb);
                if (nb > 0)
                    amac_add_tl_kernel<T><<<nb, dim3(TileSet::kTX, TileSet::kTY), 0, s>>>(
                        at(l).g.k(), at(l).unk.as<unsigned char>(), e, at(l).x.as<double>(), tl);
            }
        }
    }

    template <class T>
    void prepare(MacHierarchy<T>& H, bool& have_beta, hipStream_t s) {
        if (!H.built || (int)H.L.size() != nlev) {
            H.build(at(0).g, nlev);
            have_beta = false;
        }
        if (!have_beta) {
            const double* b0[3] = {at(0).beta[0].as<double>(), at(0).beta[1].as<double>(),
                                   at(0).beta[2].as<double>()};
            H.set_beta(b0, s);
            have_beta = true;
        }
        hb.alloc(at(0).g.ncell() * sizeof(double));
        H.L[0].b = hb.as<T>();
    }

    // composite solve of the right-hand side in the levels' b; result in
    // their x with the cells under a finer level set to the mean of its 8
    // cells (from the finest level down) and, when singular, the weighted
    // mean removed
    void solve(double rtol, double atol, int mode, mgx_solve_info* info, hipStream_t s) {
        MGX_REQUIRE(beta_set, MGX_ERR_STATE, "AMR MAC projector: beta not set");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_UNSUPPORTED,
                    "AMR MAC projector: FP64 and mixed modes only");
        const bool mixed = mode == MGX_MODE_MIXED;
        if (mixed) prepare(H32, h32_beta, s);
        else prepare(H64, h64_beta, s);
        MGX_HIP(hipMemsetAsync(ibuf.p, 0, sizeof(int), s));
        const bool sing = at(0).g.singular();
        std::vector<double*> xs, bs;
        for (auto& v : lev) {
            xs.push_back(v->x.as<double>());
            bs.push_back(v->b.as<double>());
        }
        if (sing) remove_mean(bs, 40, s);
        for (auto& v : lev) MGX_HIP(hipMemsetAsync(v->x.p, 0, v->g.ncell() * sizeof(double), s));
        const double bnorm = residual<double>(s);  // x = 0: ||b||_inf over the unknowns
        const double tol = std::max(atol, rtol * bnorm);
        const double r0 = mixed ? residual<float>(s) : bnorm;
        double rn = r0;
        int it = 0;
        bool conv = std::isfinite(bnorm) && rn <= tol;
        while (!conv && it < opts.maxiter && std::isfinite(rn)) {
            ++it;
            if (mixed) {
                vcycle<float>(H32, s);
                rn = residual<float>(s);
            } else {
                vcycle<double>(H64, s);
                rn = residual<double>(s);
            }
            if (opts.verbose)
                fprintf(stderr, "mgx AMR MAC MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                        bnorm > 0 ? rn / bnorm : rn);
            conv = rn <= tol;
        }
        if (sing) remove_mean(xs, 41, s);
        for (int l = L() - 1; l >= 0; --l)
            amac_cover_mean_kernel<<<grid3(at(l).g.n[0], at(l).g.n[1], at(l).g.n[2]), block3(), 0, s>>>(
                at(l).g.k(), at(l + 1).g.k(), at(l).cov.as<unsigned char>(), at(l + 1).x.as<double>(),
                at(l).x.as<double>());
        MGX_HIP(hipStreamSynchronize(s));
        if (info) {
            info->iters = it;
            info->vcycles = it;
            info->converged = conv ? 1 : 0;
            info->nlevels = nlev + L();
            info->fallback_used = 0;
            info->bnorm = bnorm;
            info->resnorm0 = r0;
            info->resnorm = rn;
            info->rel_resid = bnorm > 0 ? rn / bnorm : rn;
        }
    }
};

}  // namespace mgx

using namespace mgx;

struct mgx_amr_mac_s {
    mgx_ctx_s* ctx = nullptr;
    int n_cell[3] = {0, 0, 0};
    double prob_lo[3] = {0, 0, 0}, prob_hi[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    std::vector<int> cboxes;
    std::vector<std::vector<int>> lboxes;   // levels 1..
    int bclo[3] = {0, 0, 0}, bchi[3] = {0, 0, 0};
    bool bc_set = false;
    MGOpts opts;
    bool dirty = true;
    AmrMac solver;
    std::vector<std::unique_ptr<BoxLayout>> lay;  // per level
    std::vector<std::unique_ptr<DevBuf>> u;       // per level and direction: 3 l + d
    hipStream_t stream() const { return ctx->stream; }
    int nfine() const { return (int)lboxes.size(); }

    mgx_amr_mac_s() {
        opts.maxiter = 200;
        opts.bottom_maxiter = 200;
    }

    static KGeom kgeom_of(const MacGeom& g) {
        KGeom k{};
        for (int d = 0; d < 3; ++d) {
            k.n[d] = g.n[d];
            k.per[d] = g.per[d];
            k.nn[d] = g.n[d];
            k.sg[d] = g.n[d] + 2;
        }
        return k;
    }

    void ensure_setup() {
        if (!dirty) return;
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(periodic[d] || bc_set, MGX_ERR_STATE,
                        "AMR MAC projector: setDomainBC is required for non-periodic domains");
        MacGeom g{};
        for (int d = 0; d < 3; ++d) {
            g.n[d] = n_cell[d];
            g.per[d] = periodic[d];
            g.dlo[d] = (!periodic[d] && bclo[d] == MGX_BC_DIRICHLET) ? 1 : 0;
            g.dhi[d] = (!periodic[d] && bchi[d] == MGX_BC_DIRICHLET) ? 1 : 0;
            g.rlo[d] = g.rhi[d] = 0;
            g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
        }
        g.robin_fixes = false;
        solver.setup(g, lboxes, opts);
        const int L = nfine();
        lay.clear();
        u.clear();
        int zero[3] = {0, 0, 0};
        for (int l = 0; l <= L; ++l) {
            const MacGeom& gl = solver.at(l).g;
            int nl[3] = {gl.n[0], gl.n[1], gl.n[2]};
            double dxl[3] = {gl.dx[0], gl.dx[1], gl.dx[2]};
            lay.emplace_back(new BoxLayout());
            if (l == 0) lay[0]->build(cboxes.data(), (int)cboxes.size() / 6, make_nodal_geom(nl, periodic, zero, zero, dxl));
            else lay[l]->build(lboxes[l - 1].data(), (int)lboxes[l - 1].size() / 6,
                               make_nodal_geom(nl, periodic, zero, zero, dxl), true);
            for (int d = 0; d < 3; ++d) {
                u.emplace_back(new DevBuf());
                u.back()->alloc(gl.nface(d) * sizeof(double));
            }
        }
        dirty = false;
    }

    // boxed face field of level l <-> canonical faces
    void faces(int l, int d, const double* boxed, int ng, double* canon, int to_box) {
        BoxLayout& Lb = *lay[l];
        const KGeom k = kgeom_of(solver.at(l).g);
        face_copy_kernel<<<Lb.ntiles[REG_FACE0 + d], 256, 0, stream()>>>(
            k, Lb.dev_boxes(), Lb.tiles(REG_FACE0 + d), d, const_cast<double*>(boxed), ng, canon, to_box);
    }
    void cells_out(int l, double* boxed, int ng, double* canon) {
        BoxLayout& Lb = *lay[l];
        const KGeom k = kgeom_of(solver.at(l).g);
        cell_copy_kernel<<<Lb.ntiles[REG_CELL], 256, 0, stream()>>>(k, Lb.dev_boxes(), Lb.tiles(REG_CELL),
                                                                    boxed, ng, canon, 1);
    }
    // level l >= 1, periodic direction d: the seam faces take the held value
    void merge_periodic(int l, double* uu, int d) {
        const MacGeom& g = solver.at(l).g;
        if (!g.per[d] || l == 0) return;
        const long np = (long)g.n[d == 0 ? 1 : 0] * g.n[d == 2 ? 1 : 2];
        amac_periodic_merge_kernel<<<grid1d(np), 256, 0, stream()>>>(g.k(), d,
                                                                    solver.at(l).fm.as<unsigned char>(), uu);
    }
    // a periodic direction's face n_d equals face 0 on the canonical array
    void sync_periodic(const MacGeom& g, double* uu, int d) {
        if (!g.per[d]) return;
        const long np = (long)g.n[d == 0 ? 1 : 0] * g.n[d == 2 ? 1 : 2];
        amac_periodic_face_kernel<<<grid1d(np), 256, 0, stream()>>>(g.k(), d, uu);
    }
    // faces of every level under the next finer one <- the mean of its faces, finest first
    void average_down_all(hipStream_t s) {
        for (int l = nfine() - 1; l >= 0; --l) {
            const MacGeom& gc = solver.at(l).g;
            for (int d = 0; d < 3; ++d)
                amac_avgdown_faces_kernel<<<grid3(gc.n[0] + (d == 0), gc.n[1] + (d == 1), gc.n[2] + (d == 2)),
                                            block3(), 0, s>>>(gc.k(), solver.at(l + 1).g.k(), d,
                                                              solver.at(l).cov.as<unsigned char>(),
                                                              u[3 * (l + 1) + d]->as<double>(),
                                                              u[3 * l + d]->as<double>());
        }
    }
};

extern "C" {

MGX_API int mgx_amr_mac_create(mgx_ctx ctx, const mgx_level_desc* levels, int nlevels,
                               mgx_amr_mac* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    auto h = std::make_unique<mgx_amr_mac_s>();
    int status = guarded(ctx, [&] {
        MGX_REQUIRE(levels != nullptr && nlevels >= 1 && nlevels <= MGX_AMR_MAX_LEVELS, MGX_ERR_UNSUPPORTED,
                    "AMR MAC projector: 1 to MGX_AMR_MAX_LEVELS levels (ratio 2)");
        const mgx_level_desc& c = levels[0];
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(c.n_cell[d] >= 2 && c.n_cell[d] % 2 == 0, MGX_ERR_ARG,
                        "AMR MAC projector: coarse extents must be even and >= 2");
            h->n_cell[d] = c.n_cell[d];
            h->prob_lo[d] = c.prob_lo[d];
            h->prob_hi[d] = c.prob_hi[d];
            h->periodic[d] = c.is_periodic[d] ? 1 : 0;
        }
        MGX_REQUIRE(c.nboxes > 0 && c.boxes, MGX_ERR_ARG, "AMR MAC projector: empty coarse box list");
        h->cboxes.assign(c.boxes, c.boxes + 6 * c.nboxes);
        for (int l = 1; l < nlevels; ++l) {
            const mgx_level_desc& p = levels[l - 1];
            const mgx_level_desc& f = levels[l];
            for (int d = 0; d < 3; ++d) {
                MGX_REQUIRE(f.n_cell[d] == 2 * p.n_cell[d], MGX_ERR_UNSUPPORTED,
                            "AMR MAC projector: refinement ratio 2 only");
                MGX_REQUIRE(f.prob_lo[d] == c.prob_lo[d] && f.prob_hi[d] == c.prob_hi[d] &&
                                (f.is_periodic[d] != 0) == (c.is_periodic[d] != 0),
                            MGX_ERR_ARG, "AMR MAC projector: levels must share the domain");
            }
            MGX_REQUIRE(f.nboxes >= 0 && (f.nboxes == 0 || f.boxes), MGX_ERR_ARG,
                        "AMR MAC projector: bad fine box list");
            MGX_REQUIRE(f.nboxes > 0 || l == nlevels - 1, MGX_ERR_ARG,
                        "AMR MAC projector: only the finest level may be empty");
            for (int b = 0; b < f.nboxes; ++b)
                for (int d = 0; d < 3; ++d) {
                    const int lo = f.boxes[6 * b + d], hi = f.boxes[6 * b + 3 + d];
                    MGX_REQUIRE(lo >= 0 && hi < f.n_cell[d] && lo <= hi, MGX_ERR_ARG,
                                "AMR MAC projector: fine box outside the domain");
                    MGX_REQUIRE(lo % 2 == 0 && hi % 2 == 1, MGX_ERR_ARG,
                                "AMR MAC projector: fine boxes must align with coarse cells");
                }
            if (f.nboxes > 0) h->lboxes.emplace_back(f.boxes, f.boxes + 6 * f.nboxes);
        }
        h->ctx = ctx;
    });
    if (status == MGX_OK) *out = h.release();
    return status;
}

MGX_API int mgx_amr_mac_destroy(mgx_amr_mac h) {
    if (!h) return MGX_ERR_ARG;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_amr_mac_set_domain_bc(mgx_amr_mac h, const int lo[3], const int hi[3]) {
    if (!h || !lo || !hi) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        for (int d = 0; d < 3; ++d) {
            if (h->periodic[d]) continue;
            for (int v : {lo[d], hi[d]})
                MGX_REQUIRE(v == MGX_BC_DIRICHLET || v == MGX_BC_NEUMANN || v == MGX_BC_INFLOW,
                            MGX_ERR_UNSUPPORTED,
                            "AMR MAC projector: Dirichlet, Neumann or inflow faces only");
            h->bclo[d] = lo[d];
            h->bchi[d] = hi[d];
        }
        h->bc_set = true;
        h->dirty = true;
    });
}

MGX_API int mgx_amr_mac_set_options(mgx_amr_mac h, const mgx_mlmg_opts* o) {
    if (!h || !o) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGOpts m;
        to_opts(o, m);
        h->opts = m;
        h->dirty = true;
    });
}

MGX_API int mgx_amr_mac_set_beta(mgx_amr_mac h, const double* const* beta, int ng) {
    if (!h || !beta) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        const int L = h->nfine();
        for (int q = 0; q < 3 * (L + 1); ++q)
            MGX_REQUIRE(beta[q] != nullptr, MGX_ERR_ARG,
                        "AMR MAC projector: beta needed on every level and direction");
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "beta ghost width out of range");
        h->ensure_setup();
        AmrMac& S = h->solver;
        for (int l = 0; l <= L; ++l)
            for (int d = 0; d < 3; ++d) {
                AmacLev& v = S.at(l);
                if (l > 0) MGX_HIP(hipMemsetAsync(v.beta[d].p, 0, v.g.nface(d) * sizeof(double), h->stream()));
                h->faces(l, d, beta[3 * l + d], ng, v.beta[d].as<double>(), 0);
                h->merge_periodic(l, v.beta[d].as<double>(), d);
            }
        S.beta_changed(h->stream());
    });
}

MGX_API int mgx_amr_mac_project(mgx_amr_mac h, double* const* umac, int umac_ng, double* const* phi,
                                int phi_ng, double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h || !umac) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        const int L = h->nfine();
        for (int q = 0; q < 3 * (L + 1); ++q)
            MGX_REQUIRE(umac[q] != nullptr, MGX_ERR_ARG,
                        "AMR MAC projector: face velocity needed on every level and direction");
        MGX_REQUIRE(umac_ng >= 0 && umac_ng <= 8 && phi_ng >= 0, MGX_ERR_ARG, "bad ghost widths");
        h->ensure_setup();
        AmrMac& S = h->solver;
        MGX_REQUIRE(S.beta_set, MGX_ERR_STATE, "AMR MAC projector: beta not set");
        hipStream_t s = h->stream();
        for (int l = 0; l <= L; ++l)
            for (int d = 0; d < 3; ++d) {
                double* ud = h->u[3 * l + d]->as<double>();
                if (l > 0) MGX_HIP(hipMemsetAsync(ud, 0, S.at(l).g.nface(d) * sizeof(double), s));
                h->faces(l, d, umac[3 * l + d], umac_ng, ud, 0);
                h->merge_periodic(l, ud, d);
            }
        // right-hand side: each level's cells from its faces, with the faces
        // under a finer level replaced by the mean of its faces (finest first)
        h->average_down_all(s);
        for (int l = 0; l <= L; ++l) {
            const MacGeom& g = S.at(l).g;
            mac_div_kernel<<<grid3(g.n[0], g.n[1], g.n[2]), block3(), 0, s>>>(
                g.k(), h->u[3 * l]->as<double>(), h->u[3 * l + 1]->as<double>(), h->u[3 * l + 2]->as<double>(),
                1.0 / g.dx[0], 1.0 / g.dx[1], 1.0 / g.dx[2], S.at(l).b.as<double>());
        }
        S.solve(rtol, atol, mode, info, s);
        // u -= beta grad phi on every level, then the faces averaged down
        for (int l = 0; l <= L; ++l) {
            const MacGeom& g = S.at(l).g;
            for (int d = 0; d < 3; ++d) {
                double* ud = h->u[3 * l + d]->as<double>();
                const dim3 gr = grid3(g.n[0] + (d == 0), g.n[1] + (d == 1), g.n[2] + (d == 2));
                if (l == 0)
                    mac_face_update_kernel<<<gr, block3(), 0, s>>>(g.k(), d, S.at(0).beta[d].as<double>(),
                                                                   S.at(0).x.as<double>(), 1.0 / g.dx[d], ud, 0);
                else
                    amac_fine_faces_kernel<<<gr, block3(), 0, s>>>(
                        g.k(), S.at(l - 1).g.k(), d, S.at(l).fm.as<unsigned char>(), S.at(l).beta[d].as<double>(),
                        S.at(l).x.as<double>(), S.at(l - 1).x.as<double>(), 1.0 / g.dx[d], ud);
                h->sync_periodic(g, ud, d);
            }
        }
        h->average_down_all(s);
        for (int l = 0; l <= L; ++l)
            for (int d = 0; d < 3; ++d) h->faces(l, d, umac[3 * l + d], umac_ng, h->u[3 * l + d]->as<double>(), 1);
        if (phi)
            for (int l = 0; l <= L; ++l)
                if (phi[l]) h->cells_out(l, phi[l], phi_ng, S.at(l).x.as<double>());
        MGX_HIP(hipStreamSynchronize(s));
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "AMR MAC projector: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

}  // extern "C"

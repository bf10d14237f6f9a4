// amr_mac.hip -- two-level composite MAC projection (AMR, refinement ratio 2)
// and its C ABI.
//
// incflo projects the MAC velocities of all AMR levels with one MacProjector
// (src/convection/incflo_compute_MAC_projected_velocities.cpp:96-133, 272-300)
// and the projector leaves the fine faces averaged down onto the coarse level
// (:301-302).  The composite problem here is a conservative finite-volume
// discretisation on the composite mesh:
//   unknowns: fine cells inside the fine region F, coarse cells outside it;
//   every face between two unknowns carries the two-point flux
//     beta * (phi_j - phi_i) / distance between the cell centres,
//   so a fine face on the coarse-fine boundary links its fine cell to the
//   coarse cell behind it at distance 1.5 h_f, and a coarse cell next to F
//   sees the 4 fine fluxes on each coarse-fine face (reflux);
//   right-hand side: div u per unknown, the coarse-fine faces of a coarse
//   cell taking the mean of the fine faces on them.
// The volume-weighted operator is symmetric.  Solver: composite V-cycles
// inside an FP64 outer iteration on the composite residual (MLMG's stopping
// rule ||r||_inf <= max(atol, rtol ||b||_inf)):
//   red-black Gauss-Seidel on the fine unknowns with the coarse correction
//   held at 0, the composite residual restricted to the whole coarse grid
//   (covered cells: the mean of their 8 fine residuals), one V-cycle of the
//   coarse MAC hierarchy, piecewise-constant prolongation onto F, red-black
//   Gauss-Seidel on the fine unknowns against the new coarse correction.
// MGX_MODE_MIXED runs the V-cycle in FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "amr_kernels.h"
#include "mac_mg.h"
#include "mgx_ctx.h"

namespace mgx {

template <class T>
struct AmacCoef {
    T ihc2[3];  // 1 / h_c^2
    T ihf2[3];  // 1 / h_f^2
};

__device__ __forceinline__ int wrap_idx(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

enum { AM_RESID = 0, AM_GSRB = 1 };

// Fine rows of the composite operator on the cells of F (fine mask fm).
// RESID: r = b - L x (x on F, xc behind the coarse-fine faces; xc NULL = 0),
//        ||r||_inf into norm; 0 outside F.
// GSRB:  one colour of red-black Gauss-Seidel on x, in place.
template <class T, class TX, class TO, int MODE>
static __global__ void amac_fine_kernel(MGeomK gf, MGeomK gc, AmacCoef<T> c,
                                        const unsigned char* __restrict__ fm, const T* __restrict__ bx,
                                        const T* __restrict__ by, const T* __restrict__ bz, TX* x,
                                        const TX* __restrict__ xc, const T* __restrict__ b,
                                        TO* __restrict__ r, unsigned long long* norm, int color) {
    __shared__ double red[4];
    int i, j, k;
    double m = 0.0;
    if (grid_point(gf.n[0], gf.n[1], i, j, k)) {
        const long id = cidx(gf, i, j, k);
        const bool in = fm[id] != 0;
        const bool mine = MODE != AM_GSRB || (((i + j + k + color) & 1) == 0);
        if (in && mine) {
            const int p[3] = {i, j, k};
            const T* beta[3] = {bx, by, bz};
            const T xp = (T)x[id];
            T acc = 0, dg = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    int q[3] = {p[0], p[1], p[2]};
                    q[d] += s ? 1 : -1;
                    if (q[d] < 0 || q[d] >= gf.n[d]) {
                        if (!gf.per[d]) continue;  // F stays off non-periodic faces
                        q[d] = wrap_idx(q[d], gf.n[d]);
                    }
                    int f[3] = {p[0], p[1], p[2]};
                    if (s) f[d] += 1;
                    const T bet = beta[d][fidx(gf, d, f[0], f[1], f[2])];
                    const long qid = cidx(gf, q[0], q[1], q[2]);
                    if (fm[qid]) {
                        const T t = bet * c.ihf2[d];
                        acc += t * ((T)x[qid] - xp);
                        dg -= t;
                    } else {
                        const T t = bet * c.ihf2[d] * (T)(2.0 / 3.0);
                        const T xq = xc ? (T)xc[cidx(gc, q[0] >> 1, q[1] >> 1, q[2] >> 1)] : (T)0;
                        acc += t * (xq - xp);
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
        } else if (MODE == AM_RESID) {
            r[id] = (TO)0;
        }
    }
    if (MODE == AM_RESID && norm) {
        const double bm = block_max<256>(m, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_nonneg(norm, bm);
    }
}

// Coarse rows on the uncovered cells.
// RESID: r = b - L (xc, xf); xc NULL = 0; covered cells get r = 0.
// GSRB:  one colour of red-black Gauss-Seidel on xc (uncovered cells, the
//        fine values xf held), in place: smooths the coarse cells next to the
//        fine region with the composite rows.
// Domain faces: Dirichlet (value 0 at h_c/2) or Neumann as on the single level.
template <class T, class TX, class TO, int MODE>
static __global__ void amac_coarse_kernel(MGeomK gc, MGeomK gf, AmacCoef<T> c,
                                          const unsigned char* __restrict__ cov,
                                          const T* __restrict__ bcx, const T* __restrict__ bcy,
                                          const T* __restrict__ bcz, const T* __restrict__ bfx,
                                          const T* __restrict__ bfy, const T* __restrict__ bfz,
                                          TX* xc, const TX* __restrict__ xf,
                                          const T* __restrict__ b, TO* __restrict__ r,
                                          unsigned long long* norm, int color) {
    __shared__ double red[4];
    int I, J, K;
    double m = 0.0;
    if (grid_point(gc.n[0], gc.n[1], I, J, K)) {
        const long id = cidx(gc, I, J, K);
        const bool mine = MODE != AM_GSRB || (((I + J + K + color) & 1) == 0);
        if (cov[id]) {
            if (MODE == AM_RESID) r[id] = (TO)0;
        } else if (mine) {
            const int P[3] = {I, J, K};
            const T* betac[3] = {bcx, bcy, bcz};
            const T* betaf[3] = {bfx, bfy, bfz};
            const T xp = xc ? (T)xc[id] : (T)0;
            T acc = 0, dg = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    int Q[3] = {P[0], P[1], P[2]};
                    Q[d] += s ? 1 : -1;
                    int F[3] = {P[0], P[1], P[2]};
                    if (s) F[d] += 1;
                    const T bet = betac[d][fidx(gc, d, F[0], F[1], F[2])];
                    if (Q[d] < 0 || Q[d] >= gc.n[d]) {
                        if (!gc.per[d]) {
                            if (s ? gc.dhi[d] : gc.dlo[d]) {
                                acc -= 2 * c.ihc2[d] * bet * xp;
                                dg -= 2 * c.ihc2[d] * bet;
                            }
                            continue;
                        }
                        Q[d] = wrap_idx(Q[d], gc.n[d]);
                    }
                    const long qid = cidx(gc, Q[0], Q[1], Q[2]);
                    if (!cov[qid]) {
                        const T xq = xc ? (T)xc[qid] : (T)0;
                        acc += c.ihc2[d] * bet * (xq - xp);
                        dg -= c.ihc2[d] * bet;
                        continue;
                    }
                    // coarse-fine face: the 4 fine cells of Q next to it, each
                    // through its own face (the one the fine rows use)
                    const int fd = s ? 2 * Q[d] : 2 * Q[d] + 1;
                    const int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
#pragma unroll
                    for (int a = 0; a < 2; ++a)
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            int f[3];
                            f[d] = fd;
                            f[t0] = 2 * P[t0] + a;
                            f[t1] = 2 * P[t1] + e;
                            int fa[3] = {f[0], f[1], f[2]};
                            if (!s) fa[d] += 1;  // hi face of the fine cell
                            const T bf = betaf[d][fidx(gf, d, fa[0], fa[1], fa[2])];
                            const T t = bf * c.ihf2[d] * (T)(1.0 / 12.0);
                            acc += t * ((T)xf[cidx(gf, f[0], f[1], f[2])] - xp);
                            dg -= t;
                        }
                }
            if (MODE == AM_GSRB) {
                if (dg != (T)0) xc[id] = (TX)(xp + (b[id] - acc) / dg);
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

// e_f += e_c (parent) on F
template <class T>
static __global__ void amac_prolong_kernel(MGeomK gf, MGeomK gc, const unsigned char* __restrict__ fm,
                                           const T* __restrict__ ec, T* __restrict__ ef) {
    int i, j, k;
    if (!grid_point(gf.n[0], gf.n[1], i, j, k)) return;
    const long id = cidx(gf, i, j, k);
    if (fm[id]) ef[id] += ec[cidx(gc, i >> 1, j >> 1, k >> 1)];
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

// fine faces of direction d with a cell of F on either side: u -= beta grad phi
// (between two F cells) or the coarse-fine flux (F cell to the coarse cell
// behind the face, distance 1.5 h_f); other fine faces keep u
static __global__ void amac_fine_faces_kernel(MGeomK gf, MGeomK gc, int d,
                                              const unsigned char* __restrict__ fm,
                                              const double* __restrict__ beta,
                                              const double* __restrict__ xf,
                                              const double* __restrict__ xc, double ih,
                                              double* __restrict__ u) {
    const int f0 = gf.n[0] + (d == 0), f1 = gf.n[1] + (d == 1);
    int i, j, k;
    if (!grid_point(f0, f1, i, j, k)) return;
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    lo[d] -= 1;
    if (lo[d] < 0 || hi[d] >= gf.n[d]) {
        if (!gf.per[d]) return;
        lo[d] = wrap_idx(lo[d], gf.n[d]);
        hi[d] = wrap_idx(hi[d], gf.n[d]);
    }
    const long il = cidx(gf, lo[0], lo[1], lo[2]), ihh = cidx(gf, hi[0], hi[1], hi[2]);
    const bool fl = fm[il] != 0, fh = fm[ihh] != 0;
    if (!fl && !fh) return;
    const long fid = fidx(gf, d, i, j, k);
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

class AmrMac {
   public:
    MacGeom gc{}, gf{};
    MGOpts opts;
    int nlev = 1;
    MacHierarchy<double> H64;
    MacHierarchy<float> H32;
    DevBuf fmask, cmask;             // fine cells of F, covered coarse cells
    DevBuf betac[3], betaf[3], betaf32[3];
    DevBuf xc, xf, bc, bf;           // FP64 iterate and right-hand side
    DevBuf rc, rf, rc2, rf2, ef;     // V-cycle work (FP64-sized, reused for FP32)
    DevBuf hb;                       // level-0 right-hand side of the coarse hierarchy
    DevBuf part, scal, ibuf;
    unsigned long long* h_pinned = nullptr;
    double ndof = 0.0;               // unknowns in units of fine cells (coarse: 8)
    bool beta_set = false, h64_beta = false, h32_beta = false;

    ~AmrMac() {
        if (h_pinned) (void)hipHostFree(h_pinned);
    }

    void setup(const MacGeom& g0, const std::vector<int>& fboxes, const MGOpts& o) {
        // the boxes and the periodicity are fixed per handle, so a set-up for
        // new boundary types or options keeps the coefficient arrays
        const bool keep = beta_set;
        gc = g0;
        gf = g0;
        for (int d = 0; d < 3; ++d) {
            gf.n[d] = 2 * g0.n[d];
            gf.dx[d] = 0.5 * g0.dx[d];
        }
        opts = o;
        nlev = amac_nlevels(gc, o.max_coarsening_level);
        std::vector<unsigned char> fm(gf.ncell(), 0), cm(gc.ncell(), 0);
        for (size_t b = 0; b + 5 < fboxes.size(); b += 6)
            for (int k = fboxes[b + 2]; k <= fboxes[b + 5]; ++k)
                for (int j = fboxes[b + 1]; j <= fboxes[b + 4]; ++j)
                    for (int i = fboxes[b + 0]; i <= fboxes[b + 3]; ++i)
                        fm[i + (size_t)gf.n[0] * (j + (size_t)gf.n[1] * k)] = 1;
        long nfu = 0, ncu = 0;
        for (int k = 0; k < gc.n[2]; ++k)
            for (int j = 0; j < gc.n[1]; ++j)
                for (int i = 0; i < gc.n[0]; ++i) {
                    const size_t id = i + (size_t)gc.n[0] * (j + (size_t)gc.n[1] * k);
                    cm[id] = fm[2 * i + (size_t)gf.n[0] * (2 * j + (size_t)gf.n[1] * 2 * k)];
                    ncu += cm[id] ? 0 : 1;
                }
        for (unsigned char v : fm) nfu += v;
        ndof = (double)nfu + 8.0 * (double)ncu;
        fmask.alloc(fm.size());
        cmask.alloc(cm.size());
        MGX_HIP(hipMemcpy(fmask.p, fm.data(), fm.size(), hipMemcpyHostToDevice));
        MGX_HIP(hipMemcpy(cmask.p, cm.data(), cm.size(), hipMemcpyHostToDevice));
        for (int d = 0; d < 3; ++d) {
            betac[d].alloc(gc.nface(d) * sizeof(double));
            betaf[d].alloc(gf.nface(d) * sizeof(double));
            betaf32[d].alloc(gf.nface(d) * sizeof(float));
            if (!keep) MGX_HIP(hipMemset(betaf[d].p, 0, gf.nface(d) * sizeof(double)));
        }
        for (DevBuf* b : {&xc, &bc, &rc, &rc2}) b->alloc(gc.ncell() * sizeof(double));
        for (DevBuf* b : {&xf, &bf, &rf, &rf2, &ef}) b->alloc(gf.ncell() * sizeof(double));
        scal.alloc(64 * sizeof(double));
        part.alloc(4096 * sizeof(double));
        ibuf.alloc(16 * sizeof(int));
        if (!h_pinned) MGX_HIP(hipHostMalloc((void**)&h_pinned, 64 * sizeof(unsigned long long), 0));
        H64.built = H32.built = false;
        beta_set = keep;
        h64_beta = h32_beta = false;
    }

    // betac / betaf (FP64 canonical faces) were filled by the caller
    void beta_changed(hipStream_t s) {
        for (int d = 0; d < 3; ++d)
            cast_kernel<float, double><<<grid1d(gf.nface(d)), 256, 0, s>>>(
                betaf[d].as<double>(), betaf32[d].as<float>(), gf.nface(d));
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
    AmacCoef<T> coef() const {
        AmacCoef<T> c;
        for (int d = 0; d < 3; ++d) {
            c.ihc2[d] = (T)(1.0 / (gc.dx[d] * gc.dx[d]));
            c.ihf2[d] = (T)(1.0 / (gf.dx[d] * gf.dx[d]));
        }
        return c;
    }

    // v -= volume-weighted mean over the unknowns of both levels (coarse cell
    // = 8 fine cells); slot: where the mean lands in scal
    void remove_mean(double* vc, double* vf, int slot, hipStream_t s) {
        const int np = 512;
        amr_sum_kernel<<<np, 256, 0, s>>>(vc, cmask.as<unsigned char>(), 0, gc.ncell(), part.as<double>());
        amr_sum_kernel<<<np, 256, 0, s>>>(vf, fmask.as<unsigned char>(), 1, gf.ncell(),
                                          part.as<double>() + np);
        amr_mean_kernel<<<1, 64, 0, s>>>(part.as<double>(), np, 8.0, part.as<double>() + np, np, 1.0,
                                         ndof, scal.as<double>() + slot);
        amr_shift_kernel<<<grid1d(gc.ncell()), 256, 0, s>>>(vc, cmask.as<unsigned char>(), 0, 0,
                                                            gc.ncell(), scal.as<double>() + slot, 1.0);
        amr_shift_kernel<<<grid1d(gf.ncell()), 256, 0, s>>>(vf, fmask.as<unsigned char>(), 1, 0,
                                                            gf.ncell(), scal.as<double>() + slot, 1.0);
    }

    // FP64 composite residual of (xc, xf) into (rc, rf) stored as TO; ||r||_inf
    template <class TO>
    double residual(hipStream_t s) {
        const MGeomK kc = gc.k(), kf = gf.k();
        const AmacCoef<double> c = coef<double>();
        zero_norm(1, s);
        amac_fine_kernel<double, double, TO, AM_RESID><<<grid3(gf.n[0], gf.n[1], gf.n[2]), block3(), 0, s>>>(
            kf, kc, c, fmask.as<unsigned char>(), betaf[0].as<double>(), betaf[1].as<double>(),
            betaf[2].as<double>(), xf.as<double>(), xc.as<double>(), bf.as<double>(), rf.as<TO>(),
            scal.as<unsigned long long>() + 1, 0);
        amac_coarse_kernel<double, double, TO, AM_RESID><<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            kc, kf, c, cmask.as<unsigned char>(), betac[0].as<double>(), betac[1].as<double>(),
            betac[2].as<double>(), betaf[0].as<double>(), betaf[1].as<double>(), betaf[2].as<double>(),
            xc.as<double>(), xf.as<double>(), bc.as<double>(), rc.as<TO>(),
            scal.as<unsigned long long>() + 1, 0);
        return read_norm(1, s);
    }

    template <class T>
    void gsrb_fine(const T* const bfT[3], T* e, const T* ec, const T* b, hipStream_t s) {
        const MGeomK kc = gc.k(), kf = gf.k();
        const AmacCoef<T> c = coef<T>();
        for (int color = 0; color < 2; ++color)
            amac_fine_kernel<T, T, T, AM_GSRB><<<grid3(gf.n[0], gf.n[1], gf.n[2]), block3(), 0, s>>>(
                kf, kc, c, fmask.as<unsigned char>(), bfT[0], bfT[1], bfT[2], e, ec, b, nullptr,
                nullptr, color);
    }

    // one composite V-cycle on the residual in (rc, rf); the correction is
    // added to (xc, xf)
    template <class T>
    void vcycle(MacHierarchy<T>& H, const T* const bfT[3], hipStream_t s) {
        const MGeomK kc = gc.k(), kf = gf.k();
        const AmacCoef<T> c = coef<T>();
        const long nf = gf.ncell(), nc = gc.ncell();
        T* e = ef.as<T>();
        zero_kernel<T><<<grid1d(nf), 256, 0, s>>>(e, nf);
        for (int q = 0; q < opts.nu1; ++q) gsrb_fine<T>(bfT, e, nullptr, rf.as<T>(), s);
        // composite residual of (0, e): fine rows and the coarse rows next to F
        amac_fine_kernel<T, T, T, AM_RESID><<<grid3(gf.n[0], gf.n[1], gf.n[2]), block3(), 0, s>>>(
            kf, kc, c, fmask.as<unsigned char>(), bfT[0], bfT[1], bfT[2], e, nullptr, rf.as<T>(),
            rf2.as<T>(), nullptr, 0);
        const T* const bcT[3] = {H.L[0].beta[0], H.L[0].beta[1], H.L[0].beta[2]};
        amac_coarse_kernel<T, T, T, AM_RESID><<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            kc, kf, c, cmask.as<unsigned char>(), bcT[0], bcT[1], bcT[2], bfT[0], bfT[1], bfT[2],
            nullptr, e, rc.as<T>(), rc2.as<T>(), nullptr, 0);
        amac_restrict_kernel<T><<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            kc, kf, cmask.as<unsigned char>(), rc2.as<T>(), rf2.as<T>(), H.L[0].b);
        H.vcycle(opts, s, ibuf.as<int>());
        amac_prolong_kernel<T><<<grid3(gf.n[0], gf.n[1], gf.n[2]), block3(), 0, s>>>(
            kf, kc, fmask.as<unsigned char>(), H.L[0].x, e);
        for (int q = 0; q < opts.nu2; ++q) gsrb_fine<T>(bfT, e, H.L[0].x, rf.as<T>(), s);
        // the coarse cells next to F with the composite rows (fine values held)
        for (int color = 0; color < 2; ++color)
            amac_coarse_kernel<T, T, T, AM_GSRB><<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
                kc, kf, c, cmask.as<unsigned char>(), bcT[0], bcT[1], bcT[2], bfT[0], bfT[1], bfT[2],
                H.L[0].x, e, rc.as<T>(), nullptr, nullptr, color);
        amac_add_kernel<T><<<grid1d(nf), 256, 0, s>>>(fmask.as<unsigned char>(), 1, e, xf.as<double>(), nf);
        amac_add_kernel<T><<<grid1d(nc), 256, 0, s>>>(cmask.as<unsigned char>(), 0, H.L[0].x,
                                                      xc.as<double>(), nc);
    }

    template <class T>
    void prepare(MacHierarchy<T>& H, bool& have_beta, hipStream_t s) {
        if (!H.built || (int)H.L.size() != nlev) {
            H.build(gc, nlev);
            have_beta = false;
        }
        if (!have_beta) {
            const double* b0[3] = {betac[0].as<double>(), betac[1].as<double>(), betac[2].as<double>()};
            H.set_beta(b0, s);
            have_beta = true;
        }
        hb.alloc(gc.ncell() * sizeof(double));
        H.L[0].b = hb.as<T>();
    }

    // composite solve of the right-hand side in (bc, bf); result in (xc, xf)
    // with the covered coarse cells set to the mean of their fine cells and,
    // when singular, the volume-weighted mean removed
    void solve(double rtol, double atol, int mode, mgx_solve_info* info, hipStream_t s) {
        MGX_REQUIRE(beta_set, MGX_ERR_STATE, "AMR MAC projector: beta not set");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_UNSUPPORTED,
                    "AMR MAC projector: FP64 and mixed modes only");
        const bool mixed = mode == MGX_MODE_MIXED;
        const MGeomK kc = gc.k(), kf = gf.k();
        if (mixed) prepare(H32, h32_beta, s);
        else prepare(H64, h64_beta, s);
        MGX_HIP(hipMemsetAsync(ibuf.p, 0, sizeof(int), s));
        const bool sing = gc.singular();
        if (sing) remove_mean(bc.as<double>(), bf.as<double>(), 40, s);
        MGX_HIP(hipMemsetAsync(xc.p, 0, gc.ncell() * sizeof(double), s));
        MGX_HIP(hipMemsetAsync(xf.p, 0, gf.ncell() * sizeof(double), s));
        const double bnorm = residual<double>(s);  // x = 0: ||b||_inf over the unknowns
        const double tol = std::max(atol, rtol * bnorm);
        const double r0 = mixed ? residual<float>(s) : bnorm;
        double rn = r0;
        int it = 0;
        bool conv = std::isfinite(bnorm) && rn <= tol;
        const float* bf32[3] = {betaf32[0].as<float>(), betaf32[1].as<float>(), betaf32[2].as<float>()};
        const double* bf64[3] = {betaf[0].as<double>(), betaf[1].as<double>(), betaf[2].as<double>()};
        while (!conv && it < opts.maxiter && std::isfinite(rn)) {
            ++it;
            if (mixed) {
                vcycle<float>(H32, bf32, s);
                rn = residual<float>(s);
            } else {
                vcycle<double>(H64, bf64, s);
                rn = residual<double>(s);
            }
            if (opts.verbose)
                fprintf(stderr, "mgx AMR MAC MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                        bnorm > 0 ? rn / bnorm : rn);
            conv = rn <= tol;
        }
        if (sing) remove_mean(xc.as<double>(), xf.as<double>(), 41, s);
        amac_cover_mean_kernel<<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            kc, kf, cmask.as<unsigned char>(), xf.as<double>(), xc.as<double>());
        MGX_HIP(hipStreamSynchronize(s));
        if (info) {
            info->iters = it;
            info->vcycles = it;
            info->converged = conv ? 1 : 0;
            info->nlevels = nlev + 1;
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
    std::vector<int> cboxes, fboxes;
    int bclo[3] = {0, 0, 0}, bchi[3] = {0, 0, 0};
    bool bc_set = false;
    MGOpts opts;
    bool dirty = true;
    AmrMac solver;
    BoxLayout lay_c, lay_f;
    DevBuf uc[3], uf[3];
    hipStream_t stream() const { return ctx->stream; }

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
        solver.setup(g, fboxes, opts);
        int zero[3] = {0, 0, 0};
        double dxc[3] = {g.dx[0], g.dx[1], g.dx[2]};
        double dxf[3] = {0.5 * g.dx[0], 0.5 * g.dx[1], 0.5 * g.dx[2]};
        int nf[3] = {2 * n_cell[0], 2 * n_cell[1], 2 * n_cell[2]};
        lay_c.build(cboxes.data(), (int)cboxes.size() / 6, make_nodal_geom(n_cell, periodic, zero, zero, dxc));
        if (!fboxes.empty())
            lay_f.build(fboxes.data(), (int)fboxes.size() / 6, make_nodal_geom(nf, periodic, zero, zero, dxf),
                        true);
        for (int d = 0; d < 3; ++d) {
            uc[d].alloc(solver.gc.nface(d) * sizeof(double));
            uf[d].alloc(solver.gf.nface(d) * sizeof(double));
        }
        dirty = false;
    }

    // boxed face field of a level <-> canonical faces
    void faces(bool fine, int d, const double* boxed, int ng, double* canon, int to_box) {
        BoxLayout& L = fine ? lay_f : lay_c;
        const KGeom k = kgeom_of(fine ? solver.gf : solver.gc);
        face_copy_kernel<<<L.ntiles[REG_FACE0 + d], 256, 0, stream()>>>(
            k, L.dev_boxes(), L.tiles(REG_FACE0 + d), d, const_cast<double*>(boxed), ng, canon, to_box);
    }
    void cells_out(bool fine, double* boxed, int ng, double* canon) {
        BoxLayout& L = fine ? lay_f : lay_c;
        const KGeom k = kgeom_of(fine ? solver.gf : solver.gc);
        cell_copy_kernel<<<L.ntiles[REG_CELL], 256, 0, stream()>>>(k, L.dev_boxes(), L.tiles(REG_CELL),
                                                                   boxed, ng, canon, 1);
    }
    // fine level, periodic direction d: the seam faces take the held value
    void merge_periodic(double* u, int d) {
        const MacGeom& g = solver.gf;
        if (!g.per[d]) return;
        const long np = (long)g.n[d == 0 ? 1 : 0] * g.n[d == 2 ? 1 : 2];
        amac_periodic_merge_kernel<<<grid1d(np), 256, 0, stream()>>>(g.k(), d,
                                                                    solver.fmask.as<unsigned char>(), u);
    }
    // a periodic direction's face n_d equals face 0 on the canonical array
    void sync_periodic(const MacGeom& g, double* u, int d) {
        if (!g.per[d]) return;
        const long np = (long)g.n[d == 0 ? 1 : 0] * g.n[d == 2 ? 1 : 2];
        amac_periodic_face_kernel<<<grid1d(np), 256, 0, stream()>>>(g.k(), d, u);
    }
};

extern "C" {

MGX_API int mgx_amr_mac_create(mgx_ctx ctx, const mgx_level_desc* levels, int nlevels,
                               mgx_amr_mac* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    auto h = std::make_unique<mgx_amr_mac_s>();
    int status = guarded(ctx, [&] {
        MGX_REQUIRE(levels != nullptr && (nlevels == 1 || nlevels == 2), MGX_ERR_UNSUPPORTED,
                    "AMR MAC projector: one coarse level and at most one fine level (ratio 2)");
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
        if (nlevels == 2) {
            const mgx_level_desc& f = levels[1];
            for (int d = 0; d < 3; ++d) {
                MGX_REQUIRE(f.n_cell[d] == 2 * c.n_cell[d], MGX_ERR_UNSUPPORTED,
                            "AMR MAC projector: refinement ratio 2 only");
                MGX_REQUIRE(f.prob_lo[d] == c.prob_lo[d] && f.prob_hi[d] == c.prob_hi[d] &&
                                (f.is_periodic[d] != 0) == (c.is_periodic[d] != 0),
                            MGX_ERR_ARG, "AMR MAC projector: levels must share the domain");
            }
            MGX_REQUIRE(f.nboxes >= 0 && (f.nboxes == 0 || f.boxes), MGX_ERR_ARG,
                        "AMR MAC projector: bad fine box list");
            for (int b = 0; b < f.nboxes; ++b)
                for (int d = 0; d < 3; ++d) {
                    const int lo = f.boxes[6 * b + d], hi = f.boxes[6 * b + 3 + d];
                    MGX_REQUIRE(lo >= 0 && hi < f.n_cell[d] && lo <= hi, MGX_ERR_ARG,
                                "AMR MAC projector: fine box outside the domain");
                    MGX_REQUIRE(lo % 2 == 0 && hi % 2 == 1, MGX_ERR_ARG,
                                "AMR MAC projector: fine boxes must align with coarse cells");
                    MGX_REQUIRE(c.is_periodic[d] || (lo > 0 && hi < f.n_cell[d] - 1),
                                MGX_ERR_UNSUPPORTED,
                                "AMR MAC projector: fine boxes must stay off non-periodic domain faces");
                }
            h->fboxes.assign(f.boxes, f.boxes + 6 * f.nboxes);
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
        const bool two = !h->fboxes.empty();
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(beta[d] != nullptr && (!two || beta[3 + d] != nullptr), MGX_ERR_ARG,
                        "AMR MAC projector: beta needed on every level and direction");
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "beta ghost width out of range");
        h->ensure_setup();
        AmrMac& S = h->solver;
        for (int d = 0; d < 3; ++d) {
            h->faces(false, d, beta[d], ng, S.betac[d].as<double>(), 0);
            MGX_HIP(hipMemsetAsync(S.betaf[d].p, 0, S.gf.nface(d) * sizeof(double), h->stream()));
            if (two) h->faces(true, d, beta[3 + d], ng, S.betaf[d].as<double>(), 0);
            h->merge_periodic(S.betaf[d].as<double>(), d);
        }
        S.beta_changed(h->stream());
    });
}

MGX_API int mgx_amr_mac_project(mgx_amr_mac h, double* const* umac, int umac_ng, double* const* phi,
                                int phi_ng, double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h || !umac) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        const bool two = !h->fboxes.empty();
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(umac[d] != nullptr && (!two || umac[3 + d] != nullptr), MGX_ERR_ARG,
                        "AMR MAC projector: face velocity needed on every level and direction");
        MGX_REQUIRE(umac_ng >= 0 && umac_ng <= 8 && phi_ng >= 0, MGX_ERR_ARG, "bad ghost widths");
        h->ensure_setup();
        AmrMac& S = h->solver;
        MGX_REQUIRE(S.beta_set, MGX_ERR_STATE, "AMR MAC projector: beta not set");
        hipStream_t s = h->stream();
        const MGeomK kc = S.gc.k(), kf = S.gf.k();
        for (int d = 0; d < 3; ++d) {
            h->faces(false, d, umac[d], umac_ng, h->uc[d].as<double>(), 0);
            MGX_HIP(hipMemsetAsync(h->uf[d].p, 0, S.gf.nface(d) * sizeof(double), s));
            if (two) h->faces(true, d, umac[3 + d], umac_ng, h->uf[d].as<double>(), 0);
            h->merge_periodic(h->uf[d].as<double>(), d);
        }
        // right-hand side: fine cells from fine faces, coarse cells from coarse
        // faces with the fine faces averaged down onto the coarse-fine ones
        for (int d = 0; d < 3; ++d)
            amac_avgdown_faces_kernel<<<grid3(S.gc.n[0] + (d == 0), S.gc.n[1] + (d == 1), S.gc.n[2] + (d == 2)),
                                        block3(), 0, s>>>(kc, kf, d, S.cmask.as<unsigned char>(),
                                                          h->uf[d].as<double>(), h->uc[d].as<double>());
        mac_div_kernel<<<grid3(S.gf.n[0], S.gf.n[1], S.gf.n[2]), block3(), 0, s>>>(
            kf, h->uf[0].as<double>(), h->uf[1].as<double>(), h->uf[2].as<double>(), 1.0 / S.gf.dx[0],
            1.0 / S.gf.dx[1], 1.0 / S.gf.dx[2], S.bf.as<double>());
        mac_div_kernel<<<grid3(S.gc.n[0], S.gc.n[1], S.gc.n[2]), block3(), 0, s>>>(
            kc, h->uc[0].as<double>(), h->uc[1].as<double>(), h->uc[2].as<double>(), 1.0 / S.gc.dx[0],
            1.0 / S.gc.dx[1], 1.0 / S.gc.dx[2], S.bc.as<double>());
        S.solve(rtol, atol, mode, info, s);
        // u -= beta grad phi per level, then the fine faces averaged down
        for (int d = 0; d < 3; ++d) {
            mac_face_update_kernel<<<grid3(S.gc.n[0] + (d == 0), S.gc.n[1] + (d == 1), S.gc.n[2] + (d == 2)),
                                     block3(), 0, s>>>(kc, d, S.betac[d].as<double>(), S.xc.as<double>(),
                                                       1.0 / S.gc.dx[d], h->uc[d].as<double>(), 0);
            amac_fine_faces_kernel<<<grid3(S.gf.n[0] + (d == 0), S.gf.n[1] + (d == 1), S.gf.n[2] + (d == 2)),
                                     block3(), 0, s>>>(kf, kc, d, S.fmask.as<unsigned char>(),
                                                       S.betaf[d].as<double>(), S.xf.as<double>(),
                                                       S.xc.as<double>(), 1.0 / S.gf.dx[d],
                                                       h->uf[d].as<double>());
            h->sync_periodic(S.gf, h->uf[d].as<double>(), d);
            amac_avgdown_faces_kernel<<<grid3(S.gc.n[0] + (d == 0), S.gc.n[1] + (d == 1), S.gc.n[2] + (d == 2)),
                                        block3(), 0, s>>>(kc, kf, d, S.cmask.as<unsigned char>(),
                                                          h->uf[d].as<double>(), h->uc[d].as<double>());
            h->faces(false, d, umac[d], umac_ng, h->uc[d].as<double>(), 1);
            if (two) h->faces(true, d, umac[3 + d], umac_ng, h->uf[d].as<double>(), 1);
        }
        if (phi) {
            if (phi[0]) h->cells_out(false, phi[0], phi_ng, S.xc.as<double>());
            if (two && phi[1]) h->cells_out(true, phi[1], phi_ng, S.xf.as<double>());
        }
        MGX_HIP(hipStreamSynchronize(s));
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "AMR MAC projector: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

}  // extern "C"

// mac_kernels.h -- HIP kernels of the MAC projection's cell-centred multigrid.
//
// Operator (variable face coefficient beta, cells c, faces f):
//   F_f = beta_f (phi_R - phi_L)/h        interior and periodic faces
//   F_f = 0                               Neumann boundary faces (walls, inflow)
//   F_lo = 2 beta_f phi_0 / h,  F_hi = -2 beta_f phi_{n-1} / h   Dirichlet faces
//   (L phi)_c = sum_d (F_{c+1/2} - F_{c-1/2}) / h_d
// i.e. div(beta grad phi) for Hydro::MacProjector (AMReX-Hydro, called from
// src/convection/incflo_compute_MAC_projected_velocities.cpp:294-299).
// With a cell coefficient A (acoef, optional) the operator is
//   (L phi)_c = sum_d (F_{c+1/2} - F_{c-1/2}) / h_d - A_c phi_c,
// the Helmholtz form -(alpha a - beta div(b grad)) of MLABecLaplacian with
// A = alpha a and the face coefficients beta b (incflo's implicit viscous
// solve, src/diffusion/DiffusionScalarOp.cpp:396-478).
// Cells are stored x-fastest without ghosts; face arrays along d hold n_d+1
// faces (the periodic face n_d equals face 0).
#pragma once

#include <hip/hip_runtime.h>

#include "nodal_kernels.h"

namespace mgx {

struct MGeomK {
    int n[3];
    int per[3];
    int dlo[3], dhi[3];  // Dirichlet faces
};

template <class T>
struct MCoef {
    T ih2[3];  // 1/h_d^2
};

__device__ __forceinline__ long cidx(const MGeomK& g, int i, int j, int k) {
    return i + (long)g.n[0] * (j + (long)g.n[1] * k);
}

__device__ __forceinline__ long fidx(const MGeomK& g, int d, int i, int j, int k) {
    const long f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1);
    return i + f0 * (j + f1 * (long)k);
}

// (L phi)_c and the diagonal of L at cell (i,j,k); phi read with direct loads.
template <class T, class TB, class TP>
__device__ __forceinline__ T mac_apply_cell(const MGeomK& g, const MCoef<T>& c, const TB* bx,
                                            const TB* by, const TB* bz, const TB* acoef,
                                            const TP* phi, int i, int j, int k, T& diag) {
    const int p[3] = {i, j, k};
    const TB* beta[3] = {bx, by, bz};
    const T pc = (T)phi[cidx(g, i, j, k)];
    T acc = 0, dg = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // lo face
        {
            int q[3] = {p[0], p[1], p[2]};
            const T b = (T)beta[d][fidx(g, d, q[0], q[1], q[2])];
            if (p[d] > 0 || g.per[d]) {
                q[d] = p[d] > 0 ? p[d] - 1 : g.n[d] - 1;
                acc += c.ih2[d] * b * ((T)phi[cidx(g, q[0], q[1], q[2])] - pc);
                dg -= c.ih2[d] * b;
            } else if (g.dlo[d]) {
                acc -= 2 * c.ih2[d] * b * pc;
                dg -= 2 * c.ih2[d] * b;
            }
        }
        // hi face
        {
            int q[3] = {p[0], p[1], p[2]};
            q[d] = p[d] + 1;
            const T b = (T)beta[d][fidx(g, d, q[0], q[1], q[2])];
            if (p[d] + 1 < g.n[d] || g.per[d]) {
                q[d] = p[d] + 1 < g.n[d] ? p[d] + 1 : 0;
                acc += c.ih2[d] * b * ((T)phi[cidx(g, q[0], q[1], q[2])] - pc);
                dg -= c.ih2[d] * b;
            } else if (g.dhi[d]) {
                acc -= 2 * c.ih2[d] * b * pc;
                dg -= 2 * c.ih2[d] * b;
            }
        }
    }
    if (acoef) {
        const T av = (T)acoef[cidx(g, i, j, k)];
        acc -= av * pc;
        dg -= av;
    }
    diag = dg;
    return acc;
}

// One colour of red-black Gauss-Seidel, in place.  Threads cover pairs of
// cells along x; each updates the cell of the requested colour.
template <class T>
static __global__ void mac_gsrb_kernel(MGeomK g, MCoef<T> c, const T* __restrict__ bx,
                                       const T* __restrict__ by, const T* __restrict__ bz,
                                       const T* __restrict__ a, const T* __restrict__ b, T* x,
                                       int color) {
    int ih, j, k;
    if (!grid_point((g.n[0] + 1) / 2, g.n[1], ih, j, k)) return;
    const int i = 2 * ih + ((j + k + color) & 1);
    if (i >= g.n[0]) return;
    T dg;
    const T acc = mac_apply_cell<T, T, T>(g, c, bx, by, bz, a, x, i, j, k, dg);
    const long id = cidx(g, i, j, k);
    if (dg != (T)0) x[id] += (b[id] - acc) / dg;
}

enum { MK_RESID = 0, MK_UPDRES = 1, MK_APPLY = 2 };

// RESID: r = b - L x.  UPDRES: x_out = x + cor, r = b - L x_out.  APPLY: r = L x.
template <class T, class TC, int MODE>
static __global__ void mac_resid_kernel(MGeomK g, MCoef<T> c, const T* __restrict__ bx,
                                        const T* __restrict__ by, const T* __restrict__ bz,
                                        const T* __restrict__ a,
                                        const T* __restrict__ b, const T* __restrict__ x,
                                        const TC* __restrict__ cor, T* __restrict__ xout,
                                        TC* __restrict__ r, unsigned long long* norm) {
    __shared__ double red[4];
    int i, j, k;
    double m = 0.0;
    if (grid_point(g.n[0], g.n[1], i, j, k)) {
        const long id = cidx(g, i, j, k);
        T dg, acc;
        if (MODE == MK_UPDRES) {
            // L(x + cor) = L x + L cor (linear); both read directly
            T d2;
            acc = mac_apply_cell<T, T, T>(g, c, bx, by, bz, a, x, i, j, k, dg) +
                  mac_apply_cell<T, T, TC>(g, c, bx, by, bz, a, cor, i, j, k, d2);
            xout[id] = x[id] + (T)cor[id];
        } else {
            acc = mac_apply_cell<T, T, T>(g, c, bx, by, bz, a, x, i, j, k, dg);
        }
        if (MODE == MK_APPLY) {
            r[id] = (TC)acc;
        } else {
            const T rv = b[id] - acc;
            r[id] = (TC)rv;
            m = nabs((double)rv);
        }
    }
    if (MODE != MK_APPLY && norm) {
        const double bm = block_max<256>(m, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) atomic_max_nonneg(norm, bm);
    }
}

// coarse cell = mean of its 8 fine cells
template <class T>
static __global__ void mac_restrict_kernel(MGeomK gf, MGeomK gc, const T* __restrict__ rf,
                                           T* __restrict__ rc) {
    int I, J, K;
    if (!grid_point(gc.n[0], gc.n[1], I, J, K)) return;
    T s = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        s += rf[cidx(gf, 2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + ((q >> 2) & 1))];
    rc[cidx(gc, I, J, K)] = (T)0.125 * s;
}

// piecewise-constant prolongation added onto the fine correction
template <class T>
static __global__ void mac_prolong_add_kernel(MGeomK gf, MGeomK gc, const T* __restrict__ xc,
                                              T* __restrict__ xf) {
    int i, j, k;
    if (!grid_point(gf.n[0], gf.n[1], i, j, k)) return;
    xf[cidx(gf, i, j, k)] += xc[cidx(gc, i >> 1, j >> 1, k >> 1)];
}

// The coarse levels of a V-cycle in one workgroup (each level at most a few
// thousand cells, where separate launches would each cost more than their
// work): L[0..nlev-1] are smoothed, L[nlev] is the level below them (the
// bottom).  Down (up == 0): for each level x = 0, nu1 x (red, black)
// Gauss-Seidel, r = b - L x, then the next level's b = the mean of 8 r.  Up:
// from the deepest, x += the parent's x, nu2 x (red, black).  The per-cell
// arithmetic is that of mac_gsrb_kernel, mac_resid_kernel,
// mac_restrict_kernel and mac_prolong_add_kernel, with a workgroup barrier
// between passes, so the results are the same bit for bit.
constexpr int kMacCoarseMax = 8;

template <class T>
struct MacCoarseLevel {
    MGeomK g;
    MCoef<T> c;
    const T* beta[3];
    const T* a;
    T* b;
    T* x;
    T* r;
};

template <class T>
struct MacCoarseArgs {
    MacCoarseLevel<T> L[kMacCoarseMax + 1];
    int nlev, nu1, nu2, up;
};

template <class T>
__device__ __forceinline__ void mac_coarse_gs(const MacCoarseLevel<T>& v, int color) {
    const int hx = (v.g.n[0] + 1) / 2;
    const int nh = hx * v.g.n[1] * v.g.n[2];
    for (int h = threadIdx.x; h < nh; h += blockDim.x) {
        const int ih = h % hx, j = (h / hx) % v.g.n[1], k = h / (hx * v.g.n[1]);
        const int i = 2 * ih + ((j + k + color) & 1);
        if (i >= v.g.n[0]) continue;
        T dg;
        const T acc = mac_apply_cell<T, T, T>(v.g, v.c, v.beta[0], v.beta[1], v.beta[2], v.a, v.x, i, j, k, dg);
        const long id = cidx(v.g, i, j, k);
        if (dg != (T)0) v.x[id] += (v.b[id] - acc) / dg;
    }
    __syncthreads();
}

template <class T>
static __global__ void __launch_bounds__(1024) mac_coarse_cycle_kernel(MacCoarseArgs<T> a) {
    if (!a.up) {
        for (int q = 0; q < a.nlev; ++q) {
            const MacCoarseLevel<T>& v = a.L[q];
            const int n = v.g.n[0] * v.g.n[1] * v.g.n[2];
            for (int id = threadIdx.x; id < n; id += blockDim.x) v.x[id] = (T)0;
            __syncthreads();
            for (int s = 0; s < a.nu1; ++s) {
                mac_coarse_gs(v, 0);
                mac_coarse_gs(v, 1);
            }
            for (int id = threadIdx.x; id < n; id += blockDim.x) {
                const int i = id % v.g.n[0], j = (id / v.g.n[0]) % v.g.n[1], k = id / (v.g.n[0] * v.g.n[1]);
                T dg;
                const T acc = mac_apply_cell<T, T, T>(v.g, v.c, v.beta[0], v.beta[1], v.beta[2], v.a, v.x, i, j, k, dg);
                const long c = cidx(v.g, i, j, k);
                v.r[c] = v.b[c] - acc;
            }
            __syncthreads();
            const MacCoarseLevel<T>& w = a.L[q + 1];
            const int nc = w.g.n[0] * w.g.n[1] * w.g.n[2];
            for (int id = threadIdx.x; id < nc; id += blockDim.x) {
                const int I = id % w.g.n[0], J = (id / w.g.n[0]) % w.g.n[1], K = id / (w.g.n[0] * w.g.n[1]);
                T s = 0;
#pragma unroll
                for (int q8 = 0; q8 < 8; ++q8)
                    s += v.r[cidx(v.g, 2 * I + (q8 & 1), 2 * J + ((q8 >> 1) & 1), 2 * K + ((q8 >> 2) & 1))];
                w.b[cidx(w.g, I, J, K)] = (T)0.125 * s;
            }
            __syncthreads();
        }
    } else {
        for (int q = a.nlev - 1; q >= 0; --q) {
            const MacCoarseLevel<T>& v = a.L[q];
            const MacCoarseLevel<T>& w = a.L[q + 1];
            const int n = v.g.n[0] * v.g.n[1] * v.g.n[2];
            for (int id = threadIdx.x; id < n; id += blockDim.x) {
                const int i = id % v.g.n[0], j = (id / v.g.n[0]) % v.g.n[1], k = id / (v.g.n[0] * v.g.n[1]);
                v.x[cidx(v.g, i, j, k)] += w.x[cidx(w.g, i >> 1, j >> 1, k >> 1)];
            }
            __syncthreads();
            for (int s = 0; s < a.nu2; ++s) {
                mac_coarse_gs(v, 0);
                mac_coarse_gs(v, 1);
            }
        }
    }
}

// coarse face beta = mean of the 4 fine faces covering it
template <class T, class TF>
static __global__ void mac_beta_coarsen_kernel(MGeomK gf, MGeomK gc, int d,
                                               const TF* __restrict__ bf, T* __restrict__ bc) {
    const int f0 = gc.n[0] + (d == 0), f1 = gc.n[1] + (d == 1), f2 = gc.n[2] + (d == 2);
    int I, J, K;
    if (!grid_point(f0, f1, I, J, K)) return;
    if (K >= f2) return;
    const int C[3] = {I, J, K};
    T s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int f[3];
        int t = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (e == d) f[e] = 2 * C[e];
            else { f[e] = 2 * C[e] + ((q >> t) & 1); ++t; }
        }
        s += (T)bf[fidx(gf, d, f[0], f[1], f[2])];
    }
    bc[fidx(gc, d, I, J, K)] = (T)0.25 * s;
}

// Face flux F_f = beta grad(phi) . e_d at face (i,j,k) of direction d (see header).
template <class T>
__device__ __forceinline__ T mac_face_flux(const MGeomK& g, const T* beta, const T* phi, int d,
                                           int i, int j, int k, T ih) {
    const int p[3] = {i, j, k};
    const T b = beta[fidx(g, d, i, j, k)];
    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
    lo[d] = p[d] - 1;
    if (g.per[d]) {
        if (lo[d] < 0) lo[d] += g.n[d];
        if (hi[d] >= g.n[d]) hi[d] -= g.n[d];
    } else {
        if (p[d] == 0) return g.dlo[d] ? 2 * b * phi[cidx(g, hi[0], hi[1], hi[2])] * ih : (T)0;
        if (p[d] == g.n[d]) return g.dhi[d] ? -2 * b * phi[cidx(g, lo[0], lo[1], lo[2])] * ih : (T)0;
    }
    return b * (phi[cidx(g, hi[0], hi[1], hi[2])] - phi[cidx(g, lo[0], lo[1], lo[2])]) * ih;
}

// rhs = div(u_mac) on cells
static __global__ void mac_div_kernel(MGeomK g, const double* __restrict__ u,
                                      const double* __restrict__ v, const double* __restrict__ w,
                                      double ihx, double ihy, double ihz, double* __restrict__ rhs) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k)) return;
    rhs[cidx(g, i, j, k)] = (u[fidx(g, 0, i + 1, j, k)] - u[fidx(g, 0, i, j, k)]) * ihx +
                            (v[fidx(g, 1, i, j + 1, k)] - v[fidx(g, 1, i, j, k)]) * ihy +
                            (w[fidx(g, 2, i, j, k + 1)] - w[fidx(g, 2, i, j, k)]) * ihz;
}

// u_d -= F (projection) or out = -F (getFluxes), on faces of direction d
static __global__ void mac_face_update_kernel(MGeomK g, int d, const double* __restrict__ beta,
                                              const double* __restrict__ phi, double ih,
                                              double* __restrict__ u, int mode_fluxes) {
    const int f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1), f2 = g.n[2] + (d == 2);
    int i, j, k;
    if (!grid_point(f0, f1, i, j, k)) return;
    if (k >= f2) return;
    const double F = mac_face_flux<double>(g, beta, phi, d, i, j, k, ih);
    const long id = fidx(g, d, i, j, k);
    if (mode_fluxes) u[id] = -F;
    else u[id] -= F;
}

// beta_f = dt / (0.5 (rho_L + rho_R)); non-periodic boundary faces use the adjacent cell
static __global__ void mac_beta_from_rho_kernel(MGeomK g, int d, const double* __restrict__ rho,
                                                double dt, double* __restrict__ beta) {
    const int f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1), f2 = g.n[2] + (d == 2);
    int i, j, k;
    if (!grid_point(f0, f1, i, j, k)) return;
    if (k >= f2) return;
    int L[3] = {i, j, k}, R[3] = {i, j, k};
    L[d] -= 1;
    const int n = g.n[d];
    if (g.per[d]) {
        if (L[d] < 0) L[d] += n;
        if (R[d] >= n) R[d] -= n;
    } else {
        if (L[d] < 0) L[d] = 0;
        if (R[d] >= n) R[d] = n - 1;
    }
    beta[fidx(g, d, i, j, k)] =
        dt / (0.5 * (rho[cidx(g, L[0], L[1], L[2])] + rho[cidx(g, R[0], R[1], R[2])]));
}

// ---- bottom: BiCGStab in one workgroup on the MAC operator (FP64) ----
template <class T>
__device__ void mac_bottom_apply(const MGeomK& g, const MCoef<double>& c, const T* bx, const T* by,
                                 const T* bz, const T* a, const double* x, double* y, long n) {
    __syncthreads();
    for (long id = threadIdx.x; id < n; id += blockDim.x) {
        const int q = (int)(id / g.n[0]);
        const int i = (int)(id - (long)q * g.n[0]);
        const int j = q % g.n[1];
        const int k = q / g.n[1];
        double dg;
        y[id] = mac_apply_cell<double, T, double>(g, c, bx, by, bz, a, x, i, j, k, dg);
    }
    __syncthreads();
}

// The MAC operator in FP64 for the bottom solve; its smoother is the
// red-black sweep of the V-cycle, each colour updated from the values before
// it (cells of one colour meet across a periodic wrap of odd length).
template <class T>
struct MacBottomOp {
    MGeomK g;
    MCoef<double> c;
    const T *bx, *by, *bz, *a;
    __device__ long size() const { return (long)g.n[0] * g.n[1] * g.n[2]; }
    __device__ bool held(long) const { return false; }
    __device__ double nfree() const { return (double)size(); }
    __device__ void apply(const double* x, double* y, long n) const {
        mac_bottom_apply<T>(g, c, bx, by, bz, a, x, y, n);
    }
    __device__ void smooth(const BottomVecs& V, long n) const {
        for (int color = 0; color < 2; ++color) {
            __syncthreads();
            for (long id = threadIdx.x; id < n; id += blockDim.x) {
                const int q = (int)(id / g.n[0]);
                const int i = (int)(id - (long)q * g.n[0]);
                const int j = q % g.n[1];
                const int k = q / g.n[1];
                double u = V.x[id];
                if (((i + j + k) & 1) == color) {
                    double dg;
                    const double acc = mac_apply_cell<double, T, double>(g, c, bx, by, bz, a, V.x,
                                                                         i, j, k, dg);
                    if (dg != 0.0) u += (V.b[id] - acc) / dg;
                }
                V.t[id] = u;
            }
            __syncthreads();
            for (long id = threadIdx.x; id < n; id += blockDim.x) V.x[id] = V.t[id];
        }
        __syncthreads();
    }
};

template <class T>
static __global__ void __launch_bounds__(1024) mac_bottom_kernel(MGeomK g, MCoef<double> c,
                                                                 const T* bx, const T* by,
                                                                 const T* bz, const T* a,
                                                                 const T* b_in,
                                                                 T* x_out, double* work, int kind,
                                                                 int maxiter, double rtol,
                                                                 int singular, int* iters_acc) {
    __shared__ double red[16];
    const MacBottomOp<T> op{g, c, bx, by, bz, a};
    bottom_solve_body(op, b_in, x_out, work, kind, maxiter, rtol, singular, iters_acc, red);
}

// ---- Robin domain faces: a phi + b dphi/dn = f ---------------------------
// The face value of phi is the mean of the boundary cell and its ghost and
// the outward derivative their difference over h, so the relation fixes
//   dphi/dn = (f - a phi_0) / (a h/2 + b).
// The linear part is the Dirichlet-face flux 2 beta' phi_0 / h with the face
// coefficient scaled to beta' = beta a h / (a h + 2 b) (a Dirichlet face at
// a = 1, b = 0; a Neumann face at a = 0), so the operator kernels run
// unchanged with the side marked as a Dirichlet side.  The constant part
// beta c, c = f / (a h/2 + b), enters the right-hand side of the boundary
// cell (-beta c / h) and the face flux (+beta c on a high face, -beta c on a
// low one).  Planes hold one value per boundary face, x fastest over the two
// remaining directions.

// plane element t of side (d, side) -> the face of the face array of direction d
__device__ __forceinline__ bool robin_face(const MGeomK& g, int d, int side, long t, int& i, int& j,
                                           int& k) {
    const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;  // the two remaining directions
    const long m0 = g.n[a0];
    if (t >= m0 * g.n[a1]) return false;
    int p[3];
    p[d] = side ? g.n[d] : 0;
    p[a0] = (int)(t % m0);
    p[a1] = (int)(t / m0);
    i = p[0];
    j = p[1];
    k = p[2];
    return true;
}

// beta *= scale on the faces of side (d, side)
template <class T>
static __global__ void robin_beta_kernel(MGeomK g, int d, int side, T* __restrict__ beta,
                                         const double* __restrict__ scale) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!robin_face(g, d, side, t, i, j, k)) return;
    const long f = fidx(g, d, i, j, k);
    beta[f] = (T)((double)beta[f] * scale[t]);
}

// rhs of the boundary cell -= beta c / h  (beta: the unscaled face array)
static __global__ void robin_rhs_kernel(MGeomK g, int d, int side, const double* __restrict__ beta,
                                        const double* __restrict__ c, double ih,
                                        double* __restrict__ rhs) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!robin_face(g, d, side, t, i, j, k)) return;
    int p[3] = {i, j, k};
    if (side) p[d] -= 1;
    rhs[cidx(g, p[0], p[1], p[2])] -= beta[fidx(g, d, i, j, k)] * c[t] * ih;
}

// the constant part of the face flux: u -= (+-) beta c (projection), and the
// same for getFluxes' -F
static __global__ void robin_flux_kernel(MGeomK g, int d, int side, const double* __restrict__ beta,
                                         const double* __restrict__ c, double* __restrict__ u) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!robin_face(g, d, side, t, i, j, k)) return;
    const long f = fidx(g, d, i, j, k);
    const double F = beta[f] * c[t];
    u[f] -= side ? F : -F;
}

// ---- embedded boundary (cut cells) -------------------------------------
// With face apertures a_f (fluid fraction of each face) the EB operator in
// symmetric form is the regular operator with face coefficients a_f beta_f;
// its right-hand side is div(a u) = V div_EB(u); covered cells have only
// zero-aperture faces and drop out (zero rows).

// beta *= aperture on the faces of one direction
static __global__ void eb_face_scale_kernel(const double* __restrict__ ap, double* __restrict__ beta,
                                            long nf) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < nf) beta[id] *= ap[id];
}

// rhs = div(a u) on cells
static __global__ void eb_mac_div_kernel(MGeomK g, const double* __restrict__ u,
                                         const double* __restrict__ v, const double* __restrict__ w,
                                         const double* __restrict__ ax, const double* __restrict__ ay,
                                         const double* __restrict__ az, double ihx, double ihy,
                                         double ihz, double* __restrict__ rhs) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k)) return;
    const long x1 = fidx(g, 0, i + 1, j, k), x0 = fidx(g, 0, i, j, k);
    const long y1 = fidx(g, 1, i, j + 1, k), y0 = fidx(g, 1, i, j, k);
    const long z1 = fidx(g, 2, i, j, k + 1), z0 = fidx(g, 2, i, j, k);
    rhs[cidx(g, i, j, k)] = (ax[x1] * u[x1] - ax[x0] * u[x0]) * ihx +
                            (ay[y1] * v[y1] - ay[y0] * v[y0]) * ihy +
                            (az[z1] * w[z1] - az[z0] * w[z0]) * ihz;
}

// u -= beta grad(phi) on faces with fluid, u = 0 on fully covered faces
// (mode_fluxes: out = -beta grad(phi) there, 0 on covered faces)
static __global__ void eb_mac_face_update_kernel(MGeomK g, int d, const double* __restrict__ beta,
                                                 const double* __restrict__ ap,
                                                 const double* __restrict__ phi, double ih,
                                                 double* __restrict__ u, int mode_fluxes) {
    const int f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1), f2 = g.n[2] + (d == 2);
    int i, j, k;
    if (!grid_point(f0, f1, i, j, k)) return;
    if (k >= f2) return;
    const long id = fidx(g, d, i, j, k);
    if (ap[id] == 0.0) {
        u[id] = 0.0;
        return;
    }
    const double F = mac_face_flux<double>(g, beta, phi, d, i, j, k, ih);
    if (mode_fluxes) u[id] = -F;
    else u[id] -= F;
}

// EB inflow velocity (setEBInflowVelocity) in the MAC divergence: the body
// surface in a cell carries int u_eb . n dA with n out of the fluid, and by the
// divergence theorem on the fluid part of the cell int n_d dA = -(a_d+ - a_d-) h^2,
// so rhs += -sum_d (a_d+ - a_d-) u_eb,d / h_d (zero in regular cells)
static __global__ void eb_mac_inflow_kernel(MGeomK g, const double* __restrict__ ueb,
                                            const double* __restrict__ ax,
                                            const double* __restrict__ ay,
                                            const double* __restrict__ az, double ihx, double ihy,
                                            double ihz, double* __restrict__ rhs) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k)) return;
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    const long c = cidx(g, i, j, k);
    rhs[c] -= (ax[fidx(g, 0, i + 1, j, k)] - ax[fidx(g, 0, i, j, k)]) * ueb[c] * ihx +
              (ay[fidx(g, 1, i, j + 1, k)] - ay[fidx(g, 1, i, j, k)]) * ueb[nc + c] * ihy +
              (az[fidx(g, 2, i, j, k + 1)] - az[fidx(g, 2, i, j, k)]) * ueb[2 * nc + c] * ihz;
}

// phi = 0 in covered cells
static __global__ void eb_zero_covered_kernel(const unsigned char* __restrict__ covered,
                                              double* __restrict__ phi, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n && covered[id]) phi[id] = 0.0;
}

}  // namespace mgx

// nodal_kernels.h -- HIP kernels of the nodal (Q1, 27-point) multigrid.
//
// Operator (FD-normalised Q1 finite elements, variable cell coefficient sigma):
//   (L phi)_n = sum_{m != n} A_nm (phi_m - phi_n),  row sums zero, with
//   corner   A = sigma_c                      * (ihx2+ihy2+ihz2)/36
//   edge     A = (sum of the 2 sharing cells) * (2S - 3 ih2[d0])/36  (d0: the zero offset)
//   face     A = (sum of the 4 sharing cells) * (6 ih2[d] - 2S)/36  (0 when h is isotropic)
//   diag     = -4 (S/36) sum_{8 cells} sigma,   S = ihx2+ihy2+ihz2.
// This is the operator AMReX's MLNodeLaplacian applies for incflo's nodal
// projection (src/projection/incflo_apply_nodal_projection.cpp:187-215).
//
// Storage: nodes x-fastest without ghosts (periodic neighbours are wrapped,
// non-periodic ones clamped: their coefficients vanish because the cell
// coefficient is 0 outside the domain).  sigma has one ghost cell layer
// holding the periodic image or 0.
#pragma once

#include <hip/hip_runtime.h>

namespace mgx {

struct KGeom {
    int nn[3];
    int per[3];
    int dlo[3], dhi[3];
    int sg[3];
    int n[3];
};

template <class T>
struct Coef {
    T corner;
    T edge[3];
    T face[3];
    T diag4;
};

template <class T>
inline Coef<T> make_coef(const double dx[3]) {
    double ih2[3], S = 0.0;
    for (int d = 0; d < 3; ++d) { ih2[d] = 1.0 / (dx[d] * dx[d]); S += ih2[d]; }
    Coef<T> c;
    c.corner = (T)(S / 36.0);
    for (int d = 0; d < 3; ++d) {
        c.edge[d] = (T)((2.0 * S - 3.0 * ih2[d]) / 36.0);
        c.face[d] = (T)((6.0 * ih2[d] - 2.0 * S) / 36.0);
    }
    c.diag4 = (T)(4.0 * S / 36.0);
    return c;
}

// Neighbour index: periodic dims wrap (callers only need -1..n; lanes past the
// domain in a partial tile are clamped first so the result stays in [0,n)),
// non-periodic dims clamp.
__device__ __forceinline__ int wrapc(int i, int n, int per) {
    if (per) {
        i = i < -1 ? -1 : (i > n ? n : i);
        return i < 0 ? i + n : (i >= n ? i - n : i);
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ bool is_dir(const KGeom& g, int i, int j, int k) {
    return (g.dlo[0] && i == 0) || (g.dhi[0] && i == g.nn[0] - 1) || (g.dlo[1] && j == 0) ||
           (g.dhi[1] && j == g.nn[1] - 1) || (g.dlo[2] && k == 0) || (g.dhi[2] && k == g.nn[2] - 1);
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD owns a contiguous range of tiles and halo rows are shared in its L2.
__device__ __forceinline__ int xcd_remap(int h, int nblocks) {
    if ((nblocks & 7) != 0) return h;
    const int per = nblocks >> 3;
    return (h & 7) * per + (h >> 3);
}

// atomic max for non-negative doubles via their bit pattern
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double m = red[0];
        for (int w = 1; w < NT / 64; ++w) m = fmax(m, red[w]);
        red[0] = m;
    }
    __syncthreads();
    return red[0];
}

enum { KS_JACOBI = 0, KS_RESID = 1, KS_UPDRES = 2, KS_APPLY = 3 };

template <class T, class TC>
struct SArgs {
    KGeom g;
    Coef<T> c;
    const T* x;      // iterate (JACOBI, RESID) or outer phi (UPDRES)
    const TC* cor;   // UPDRES: correction added to x before the stencil
    const T* b;
    const T* sig;    // sigma with ghosts
    T* xout;         // JACOBI: new iterate; UPDRES: x + cor
    TC* rout;        // RESID/UPDRES: residual
    T omega;
    int kz;          // node planes per workgroup
    int gx, gy, gz;  // tile grid
    unsigned long long* norm;  // optional max |r|
};

// One 27-point sweep over a TX x TY column tile, streaming along z: each
// thread keeps its 3x3x3 phi neighbourhood and 2x2x2 sigma neighbourhood in
// registers and receives one new phi plane and one sigma plane per step
// through LDS (double buffered, one barrier per plane).
// FINE only separates the finest-level instantiation (a distinct kernel name
// in profiles); the code is identical.
template <class T, class TC, int MODE, bool ISO, int TX, int TY, bool FINE>
static __global__ void __launch_bounds__(TX* TY) nodal_stencil_kernel(SArgs<T, TC> a) {
    __shared__ T sphi[2][TY + 2][TX + 2];
    __shared__ T ssig[2][TY + 1][TX + 1];
    __shared__ double red[(TX * TY) / 64];
    const KGeom& g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int nb = a.gx * a.gy * a.gz;
    const int h = blockIdx.x;
    const int lb = xcd_remap(h, nb);
    const int bx = lb % a.gx, by = (lb / a.gx) % a.gy, bz = lb / (a.gx * a.gy);
    const int i0 = bx * TX, j0 = by * TY, k0 = bz * a.kz;
    const int k1 = min(k0 + a.kz, g.nn[2]);
    const int i = i0 + tx, j = j0 + ty;
    const bool active = (i < g.nn[0]) && (j < g.nn[1]);
    const long sxy = (long)g.nn[0] * g.nn[1];
    const long ssxy = (long)g.sg[0] * g.sg[1];

    T pm[3][3], p0[3][3], pp[3][3];
    T slo[2][2], shi[2][2];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int r = 0; r < 3; ++r) { pm[q][r] = 0; p0[q][r] = 0; pp[q][r] = 0; }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int r = 0; r < 2; ++r) { slo[q][r] = 0; shi[q][r] = 0; }
    double rmax = 0.0;

    for (int kk = k0 - 1; kk <= k1; ++kk) {
        const int buf = kk & 1;
        const int kw = wrapc(kk, g.nn[2], g.per[2]);
        for (int e = tid; e < (TX + 2) * (TY + 2); e += TX * TY) {
            const int lx = e % (TX + 2), ly = e / (TX + 2);
            const int gi = wrapc(i0 - 1 + lx, g.nn[0], g.per[0]);
            const int gj = wrapc(j0 - 1 + ly, g.nn[1], g.per[1]);
            const long id = gi + (long)g.nn[0] * gj + sxy * kw;
            T v = a.x[id];
            if (MODE == KS_UPDRES) v += (T)a.cor[id];
            sphi[buf][ly][lx] = v;
        }
        const int ks = min(max(kk, 0), g.sg[2] - 1);
        for (int e = tid; e < (TX + 1) * (TY + 1); e += TX * TY) {
            const int lx = e % (TX + 1), ly = e / (TX + 1);
            const int si = min(i0 + lx, g.sg[0] - 1), sj = min(j0 + ly, g.sg[1] - 1);
            ssig[buf][ly][lx] = a.sig[si + (long)g.sg[0] * sj + ssxy * ks];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                pm[q][r] = p0[q][r];
                p0[q][r] = pp[q][r];
                pp[q][r] = sphi[buf][ty + q][tx + r];
            }
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                slo[q][r] = shi[q][r];
                shi[q][r] = ssig[buf][ty + q][tx + r];
            }
        const int k = kk - 1;
        if (kk >= k0 + 1 && active) {
            const T c = p0[1][1];
            T acc = 0;
            // corners
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int aa = 0; aa < 2; ++aa)
                    acc += a.c.corner * (slo[bb][aa] * (pm[2 * bb][2 * aa] - c) +
                                         shi[bb][aa] * (pp[2 * bb][2 * aa] - c));
            // xy edges (dz = 0)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int aa = 0; aa < 2; ++aa)
                    acc += a.c.edge[2] * (slo[bb][aa] + shi[bb][aa]) * (p0[2 * bb][2 * aa] - c);
            // xz edges (dy = 0) and yz edges (dx = 0)
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                acc += a.c.edge[1] * ((slo[0][aa] + slo[1][aa]) * (pm[1][2 * aa] - c) +
                                      (shi[0][aa] + shi[1][aa]) * (pp[1][2 * aa] - c));
                acc += a.c.edge[0] * ((slo[aa][0] + slo[aa][1]) * (pm[2 * aa][1] - c) +
                                      (shi[aa][0] + shi[aa][1]) * (pp[2 * aa][1] - c));
            }
            if (!ISO) {
#pragma unroll
                for (int aa = 0; aa < 2; ++aa) {
                    acc += a.c.face[0] * (slo[0][aa] + slo[1][aa] + shi[0][aa] + shi[1][aa]) *
                           (p0[1][2 * aa] - c);
                    acc += a.c.face[1] * (slo[aa][0] + slo[aa][1] + shi[aa][0] + shi[aa][1]) *
                           (p0[2 * aa][1] - c);
                }
                acc += a.c.face[2] * ((slo[0][0] + slo[0][1] + slo[1][0] + slo[1][1]) * (pm[1][1] - c) +
                                      (shi[0][0] + shi[0][1] + shi[1][0] + shi[1][1]) * (pp[1][1] - c));
            }
            const long id = i + (long)g.nn[0] * j + sxy * k;
            const bool dir = is_dir(g, i, j, k);
            if (MODE == KS_JACOBI) {
                const T ssum = slo[0][0] + slo[0][1] + slo[1][0] + slo[1][1] + shi[0][0] +
                               shi[0][1] + shi[1][0] + shi[1][1];
                const T diag = -a.c.diag4 * ssum;
                T v = c;
                if (dir) v = 0;
                else if (ssum != (T)0) v = c + a.omega * (a.b[id] - acc) / diag;
                a.xout[id] = v;
            } else if (MODE == KS_APPLY) {
                a.rout[id] = (TC)acc;
            } else {
                const T r = dir ? (T)0 : (a.b[id] - acc);
                a.rout[id] = (TC)r;
                if (MODE == KS_UPDRES) a.xout[id] = dir ? (T)0 : c;
                rmax = fmax(rmax, (double)fabs(r));
            }
        }
    }
    if ((MODE == KS_RESID || MODE == KS_UPDRES) && a.norm) {
        const double m = block_max<TX * TY>(rmax, red);
        if (tid == 0) atomic_max_nonneg(a.norm, m);
    }
}

// x = omega * b / diag (first Jacobi sweep from a zero iterate)
template <class T>
static __global__ void nodal_jacobi_first_kernel(KGeom g, Coef<T> c, const T* __restrict__ b,
                                          const T* __restrict__ sig, T* __restrict__ x, T omega) {
    const long n = (long)g.nn[0] * g.nn[1] * g.nn[2];
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int i = id % g.nn[0];
    const int j = (id / g.nn[0]) % g.nn[1];
    const int k = id / ((long)g.nn[0] * g.nn[1]);
    if (is_dir(g, i, j, k)) { x[id] = 0; return; }
    T s = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int si = i + (q & 1), sj = j + ((q >> 1) & 1), sk = k + ((q >> 2) & 1);
        s += sig[si + (long)g.sg[0] * (sj + (long)g.sg[1] * sk)];
    }
    x[id] = (s != (T)0) ? omega * b[id] / (-c.diag4 * s) : (T)0;
}

// Full weighting: rc(I) = sum_{o in {-1,0,1}^3} prod_d w(o_d) rf(2I+o), w = (1/4,1/2,1/4).
template <class T>
static __global__ void nodal_restrict_kernel(KGeom gf, KGeom gc, const T* __restrict__ rf,
                                      T* __restrict__ rc) {
    const long n = (long)gc.nn[0] * gc.nn[1] * gc.nn[2];
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int I = id % gc.nn[0];
    const int J = (id / gc.nn[0]) % gc.nn[1];
    const int K = id / ((long)gc.nn[0] * gc.nn[1]);
    if (is_dir(gc, I, J, K)) { rc[id] = 0; return; }
    T acc = 0;
#pragma unroll
    for (int oz = -1; oz <= 1; ++oz) {
        int fk = 2 * K + oz;
        if (gf.per[2]) fk = fk < 0 ? fk + gf.nn[2] : (fk >= gf.nn[2] ? fk - gf.nn[2] : fk);
        else if (fk < 0 || fk >= gf.nn[2]) continue;
        const T wz = oz == 0 ? (T)0.5 : (T)0.25;
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy) {
            int fj = 2 * J + oy;
            if (gf.per[1]) fj = fj < 0 ? fj + gf.nn[1] : (fj >= gf.nn[1] ? fj - gf.nn[1] : fj);
            else if (fj < 0 || fj >= gf.nn[1]) continue;
            const T wy = oy == 0 ? (T)0.5 : (T)0.25;
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
                int fi = 2 * I + ox;
                if (gf.per[0]) fi = fi < 0 ? fi + gf.nn[0] : (fi >= gf.nn[0] ? fi - gf.nn[0] : fi);
                else if (fi < 0 || fi >= gf.nn[0]) continue;
                const T wx = ox == 0 ? (T)0.5 : (T)0.25;
                acc += wx * wy * wz * rf[fi + (long)gf.nn[0] * (fj + (long)gf.nn[1] * fk)];
            }
        }
    }
    rc[id] = acc;
}

// Trilinear interpolation of coarse nodes added onto fine nodes.
template <class T>
static __global__ void nodal_prolong_add_kernel(KGeom gf, KGeom gc, const T* __restrict__ xc,
                                         T* __restrict__ xf) {
    const long n = (long)gf.nn[0] * gf.nn[1] * gf.nn[2];
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int i = id % gf.nn[0];
    const int j = (id / gf.nn[0]) % gf.nn[1];
    const int k = id / ((long)gf.nn[0] * gf.nn[1]);
    if (is_dir(gf, i, j, k)) return;
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
    xf[id] += (T)0.125 * acc;
}

// Write a valid cell's value into a ghosted cell array and its periodic images.
template <class T>
__device__ __forceinline__ void write_cell_images(T* arr, const KGeom& g, int ci, int cj, int ck,
                                                  T v) {
    int xs[2] = {ci + 1, -1}, ys[2] = {cj + 1, -1}, zs[2] = {ck + 1, -1};
    if (g.per[0]) xs[1] = (ci == 0) ? g.n[0] + 1 : (ci == g.n[0] - 1 ? 0 : -1);
    if (g.per[1]) ys[1] = (cj == 0) ? g.n[1] + 1 : (cj == g.n[1] - 1 ? 0 : -1);
    if (g.per[2]) zs[1] = (ck == 0) ? g.n[2] + 1 : (ck == g.n[2] - 1 ? 0 : -1);
    for (int c = 0; c < 2; ++c) {
        if (zs[c] < 0) continue;
        for (int b = 0; b < 2; ++b) {
            if (ys[b] < 0) continue;
            for (int a = 0; a < 2; ++a) {
                if (xs[a] < 0) continue;
                arr[xs[a] + (long)g.sg[0] * (ys[b] + (long)g.sg[1] * zs[c])] = v;
            }
        }
    }
}

// coarse sigma = mean of the 8 fine cells
template <class T>
static __global__ void sigma_coarsen_kernel(KGeom gf, KGeom gc, const T* __restrict__ sf, T* sc) {
    const long n = (long)gc.n[0] * gc.n[1] * gc.n[2];
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int I = id % gc.n[0];
    const int J = (id / gc.n[0]) % gc.n[1];
    const int K = id / ((long)gc.n[0] * gc.n[1]);
    T s = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int fi = 2 * I + (q & 1) + 1, fj = 2 * J + ((q >> 1) & 1) + 1,
                  fk = 2 * K + ((q >> 2) & 1) + 1;
        s += sf[fi + (long)gf.sg[0] * (fj + (long)gf.sg[1] * fk)];
    }
    write_cell_images(sc, gc, I, J, K, (T)0.125 * s);
}

template <class TO, class TI>
static __global__ void cast_kernel(const TI* __restrict__ in, TO* __restrict__ out, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = (TO)in[id];
}

template <class T>
static __global__ void zero_kernel(T* __restrict__ a, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) a[id] = (T)0;
}

static __global__ void zero_u64_kernel(unsigned long long* p, int n) {
    const int id = threadIdx.x;
    if (id < n) p[id] = 0ull;
}

// ---- reductions (deterministic two-pass sums; max via atomic on bits) ----
template <class T>
static __global__ void __launch_bounds__(256) partial_sum_kernel(const T* __restrict__ a, long n,
                                                          double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long id = (long)blockIdx.x * 256 + threadIdx.x; id < n; id += (long)gridDim.x * 256)
        s += (double)a[id];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

static __global__ void __launch_bounds__(256) final_sum_kernel(const double* __restrict__ part, int np,
                                                        double* __restrict__ out, double scale) {
    __shared__ double red[256];
    double s = 0.0;
    for (int id = threadIdx.x; id < np; id += 256) s += part[id];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

template <class T>
static __global__ void subtract_scalar_kernel(T* __restrict__ a, long n, const double* __restrict__ s) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) a[id] = (T)((double)a[id] - s[0]);
}

template <class T>
static __global__ void __launch_bounds__(256) max_abs_kernel(const T* __restrict__ a, long n,
                                                      unsigned long long* out) {
    __shared__ double red[4];
    double m = 0.0;
    for (long id = (long)blockIdx.x * 256 + threadIdx.x; id < n; id += (long)gridDim.x * 256)
        m = fmax(m, fabs((double)a[id]));
    m = block_max<256>(m, red);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

template <class T>
static __global__ void zero_dirichlet_kernel(KGeom g, T* __restrict__ a) {
    const long n = (long)g.nn[0] * g.nn[1] * g.nn[2];
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int i = id % g.nn[0];
    const int j = (id / g.nn[0]) % g.nn[1];
    const int k = id / ((long)g.nn[0] * g.nn[1]);
    if (is_dir(g, i, j, k)) a[id] = 0;
}

// ---- bottom solve: BiCGStab (van der Vorst 1992) in one workgroup, FP64 ----
__device__ __forceinline__ double bsum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int tid = threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) s += red[w];
    return s;
}

__device__ __forceinline__ double bmax(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int tid = threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) s = fmax(s, red[w]);
    return s;
}

// (L x)_n with direct loads, in double; Dirichlet rows masked to 0.
template <class TS>
__device__ double apply_node_direct(const KGeom& g, const Coef<double>& c, const TS* sig,
                                    const double* x, int i, int j, int k) {
    if (is_dir(g, i, j, k)) return 0.0;
    double s[2][2][2];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int aa = q & 1, bb = (q >> 1) & 1, cc = (q >> 2) & 1;
        s[cc][bb][aa] = (double)sig[(i + aa) + (long)g.sg[0] * ((j + bb) + (long)g.sg[1] * (k + cc))];
    }
    const double xc = x[i + (long)g.nn[0] * (j + (long)g.nn[1] * k)];
    double acc = 0.0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                // sum sigma over the cells adjacent to both n and n+d
                double ss = 0.0;
                for (int cc = 0; cc < 2; ++cc) {
                    if (dz && cc != (dz + 1) / 2) continue;
                    for (int bb = 0; bb < 2; ++bb) {
                        if (dy && bb != (dy + 1) / 2) continue;
                        for (int aa = 0; aa < 2; ++aa) {
                            if (dx && aa != (dx + 1) / 2) continue;
                            ss += s[cc][bb][aa];
                        }
                    }
                }
                const int nz = (dx != 0) + (dy != 0) + (dz != 0);
                double w;
                if (nz == 3) w = c.corner;
                else if (nz == 2) w = c.edge[!dx ? 0 : (!dy ? 1 : 2)];
                else w = c.face[dx ? 0 : (dy ? 1 : 2)];
                const int ii = wrapc(i + dx, g.nn[0], g.per[0]);
                const int jj = wrapc(j + dy, g.nn[1], g.per[1]);
                const int kk = wrapc(k + dz, g.nn[2], g.per[2]);
                acc += w * ss * (x[ii + (long)g.nn[0] * (jj + (long)g.nn[1] * kk)] - xc);
            }
    return acc;
}

template <class TS>
__device__ void bottom_apply(const KGeom& g, const Coef<double>& c, const TS* sig, const double* x,
                             double* y, long n) {
    __syncthreads();
    for (long id = threadIdx.x; id < n; id += blockDim.x) {
        const int i = id % g.nn[0];
        const int j = (id / g.nn[0]) % g.nn[1];
        const int k = id / ((long)g.nn[0] * g.nn[1]);
        y[id] = apply_node_direct(g, c, sig, x, i, j, k);
    }
    __syncthreads();
}

template <class T>
static __global__ void __launch_bounds__(1024) bottom_bicgstab_kernel(KGeom g, Coef<double> c,
                                                               const T* sig, const T* b_in,
                                                               T* x_out, double* work, int maxiter,
                                                               double rtol, int singular,
                                                               int* iters_acc) {
    __shared__ double red[16];
    const long n = (long)g.nn[0] * g.nn[1] * g.nn[2];
    double* x = work;
    double* r = work + n;
    double* rh = work + 2 * n;
    double* p = work + 3 * n;
    double* v = work + 4 * n;
    double* s = work + 5 * n;
    double* t = work + 6 * n;
    const int tid = threadIdx.x, nt = blockDim.x;
    // b (mean-free when singular) into r; x = 0
    double bs = 0.0;
    for (long id = tid; id < n; id += nt) { r[id] = (double)b_in[id]; bs += r[id]; }
    bs = bsum(bs, red);
    const double bmean = singular ? bs / (double)n : 0.0;
    for (long id = tid; id < n; id += nt) {
        r[id] -= bmean;
        rh[id] = r[id];
        x[id] = 0.0; p[id] = 0.0; v[id] = 0.0;
    }
    double rn = 0.0;
    for (long id = tid; id < n; id += nt) rn = fmax(rn, fabs(r[id]));
    const double rnorm0 = bmax(rn, red);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    if (rnorm0 > 0.0) {
        for (it = 1; it <= maxiter; ++it) {
            double q = 0.0;
            for (long id = tid; id < n; id += nt) q += rh[id] * r[id];
            const double rho_new = bsum(q, red);
            if (rho_new == 0.0) break;
            const double beta = (rho_new / rho) * (alpha / omega);
            for (long id = tid; id < n; id += nt) p[id] = r[id] + beta * (p[id] - omega * v[id]);
            bottom_apply(g, c, sig, p, v, n);
            q = 0.0;
            for (long id = tid; id < n; id += nt) q += rh[id] * v[id];
            const double rhv = bsum(q, red);
            if (rhv == 0.0) break;
            alpha = rho_new / rhv;
            double sn = 0.0;
            for (long id = tid; id < n; id += nt) {
                x[id] += alpha * p[id];
                s[id] = r[id] - alpha * v[id];
                sn = fmax(sn, fabs(s[id]));
            }
            sn = bmax(sn, red);
            if (sn <= rtol * rnorm0) break;
            bottom_apply(g, c, sig, s, t, n);
            double tt = 0.0, ts = 0.0;
            for (long id = tid; id < n; id += nt) { tt += t[id] * t[id]; ts += t[id] * s[id]; }
            tt = bsum(tt, red);
            ts = bsum(ts, red);
            if (tt == 0.0) break;
            omega = ts / tt;
            double rr = 0.0;
            for (long id = tid; id < n; id += nt) {
                x[id] += omega * s[id];
                r[id] = s[id] - omega * t[id];
                rr = fmax(rr, fabs(r[id]));
            }
            rr = bmax(rr, red);
            if (rr <= rtol * rnorm0) break;
            if (omega == 0.0) break;
            rho = rho_new;
        }
        if (it > maxiter) it = maxiter;
    }
    __syncthreads();
    double xs = 0.0;
    for (long id = tid; id < n; id += nt) xs += x[id];
    xs = bsum(xs, red);
    const double xmean = singular ? xs / (double)n : 0.0;
    for (long id = tid; id < n; id += nt) x_out[id] = (T)(x[id] - xmean);
    if (tid == 0 && iters_acc) atomicAdd(iters_acc, it);
}

}  // namespace mgx

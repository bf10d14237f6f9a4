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

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "mgx.h"

#include <type_traits>

namespace mgx {

// Arithmetic type of a storage type: FP16-stored V-cycle arrays are computed
// in FP32; FP32 and FP64 compute in their own type.
template <class T> struct AccOf { using type = T; };
template <> struct AccOf<__half> { using type = float; };
template <class T> using acc_t = typename AccOf<T>::type;

// load / store through the arithmetic type (FP16 conversions round to nearest even)
template <class TA, class T>
__device__ __forceinline__ TA ld(const T& v) { return (TA)v; }
template <class TA>
__device__ __forceinline__ TA ld(const __half& v) { return (TA)__half2float(v); }
template <class T, class TA>
__device__ __forceinline__ T st(TA v) { return (T)v; }
template <>
__device__ __forceinline__ __half st<__half, float>(float v) { return __float2half(v); }
template <>
__device__ __forceinline__ __half st<__half, double>(double v) { return __float2half((float)v); }

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

// Point kernels run on a 3-D grid of 64 x 4 thread blocks (x, y) and one z
// plane per block row, so no thread decodes a linear index with divisions.
__device__ __forceinline__ bool grid_point(int nx, int ny, int& i, int& j, int& k) {
    i = blockIdx.x * 64 + threadIdx.x;
    j = blockIdx.y * 4 + threadIdx.y;
    k = blockIdx.z;
    return i < nx && j < ny;
}

// Blocks of blockDim.x x blockDim.y nodes of one plane taken from a list of
// (bx, by, k) entries (the blocks that hold active nodes of a partly used
// level); tl NULL: the full 64 x 4 grid.
__device__ __forceinline__ bool grid_point_tl(const int4* tl, int nx, int ny, int& i, int& j,
                                              int& k) {
    if (!tl) return grid_point(nx, ny, i, j, k);
    const int4 t = tl[blockIdx.x];
    i = t.x * (int)blockDim.x + threadIdx.x;
    j = t.y * (int)blockDim.y + threadIdx.y;
    k = t.z;
    return i < nx && j < ny;
}

inline dim3 grid3(int nx, int ny, int nz) { return dim3((nx + 63) / 64, (ny + 3) / 4, nz); }
inline dim3 block3() { return dim3(64, 4, 1); }

// |v| for max-norm reductions, with a NaN mapped to +inf: fmax drops NaN
// operands, so without this a non-finite residual would count as 0 and a
// diverged solve could pass its tolerance.
__host__ __device__ __forceinline__ double nabs(double v) {
    return v == v ? fabs(v) : __builtin_inf();
}

// atomic max for non-negative doubles via their bit pattern; the atomic is
// issued only when v exceeds the stored value (with many blocks on one
// address most skip it, where unconditional atomics serialise in L2)
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    if (bits > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, bits);
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

// JACOBI: x_out = x + w (b - Lx)/diag   RESID: r = b - Lx   UPDRES: x_out = x + cor,
// r = b - L x_out   APPLY: r = Lx   JFIRST: x_out = w b/diag (first sweep from zero)
enum { KS_JACOBI = 0, KS_RESID = 1, KS_UPDRES = 2, KS_APPLY = 3, KS_JFIRST = 4 };

template <class T, class TC>
struct SArgs {
    KGeom g;
    Coef<acc_t<T>> c;
    const T* x;      // iterate (JACOBI, RESID) or outer phi (UPDRES)
    const TC* cor;   // UPDRES: correction added to x before the stencil
    const T* b;
    const T* sig;    // sigma with ghosts
    T* xout;         // JACOBI: new iterate; UPDRES: x + cor
    TC* rout;        // RESID/UPDRES: residual
    acc_t<T> omega;
    int kz;          // node planes per workgroup
    int gx, gy, gz;  // tile grid
    unsigned long long* norm;  // optional max |r|
    TC* x1out;       // RESID/UPDRES, optional: omega1 r / diag, the first Jacobi
    double omega1;   //   iterate from zero of the next V-cycle (saves its pass)
    // optional node types (composite AMR levels): 1 unknown, 2 fixed (JACOBI
    // keeps the value), 0 outside (JACOBI / JFIRST write 0); RESID writes
    // r = 0 on every node that is not an unknown
    const unsigned char* nmask;
    // scaled reduced-precision V-cycle (0 means 1): UPDRES adds cor_scale*cor;
    // RESID/UPDRES store rscale*r and x1 from x1scale*r
    double cor_scale, rscale, x1scale;
    // optional list of the tiles to visit (linear tile indices bx + gx (by +
    // gy bz), one workgroup each): a level whose active nodes fill only part
    // of its array (an AMR fine level) is swept over those tiles alone
    const int* tiles;
};

// One 27-point pass over a 64 x 8 node tile, streaming along z.  A workgroup
// is 64 x 4 threads; thread (tx,ty) owns nodes (i0+tx, j0+2ty) and
// (i0+tx, j0+2ty+1).  Each thread keeps the 4 x 3 phi rows and 3 x 2 sigma
// rows of its two nodes for three (phi) / two (sigma) planes in registers and
// receives one new plane of each per step through LDS (double buffered, one
// barrier per plane).  The next plane's global loads (and the rhs of the
// plane being computed) are issued one step ahead, before the barrier, so
// every wave keeps a full plane of loads in flight while it computes.  All
// loads are unconditional (clamped indices) so the wait counts stay exact.
// FINE only separates the finest-level instantiation (a distinct kernel name
// in profiles); the code is identical.
// reciprocal for the smoother's diagonal scaling: the hardware reciprocal
// (1 ulp) in FP32, a division in FP64
template <class TA>
__device__ __forceinline__ TA diag_rcp(TA x) { return (TA)1 / x; }
template <>
__device__ __forceinline__ float diag_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// p advanced by a 32-bit byte offset: with p uniform across the wavefront the
// access compiles to a scalar-base + 32-bit vector-offset load or store, so a
// streaming kernel needs no per-thread 64-bit address arithmetic per plane
template <class V>
__device__ __forceinline__ V* byte_add(V* p, unsigned off) {
    return reinterpret_cast<V*>(reinterpret_cast<char*>(p) + off);
}
template <class V>
__device__ __forceinline__ const V* byte_add(const V* p, unsigned off) {
    return reinterpret_cast<const V*>(reinterpret_cast<const char*>(p) + off);
}

template <class T, bool ISO, int RO, int NPR, int NSR>
__device__ __forceinline__ T apply27(const T (&P0)[NPR][3], const T (&P1)[NPR][3],
                                     const T (&P2)[NPR][3], const T (&S0)[NSR][2],
                                     const T (&S1)[NSR][2], const Coef<T>& cf, T& ssum) {
    // planes: P0 below, P1 centre, P2 above; cell planes S0 below, S1 above
    const T c = P1[RO + 1][1];
    if constexpr (ISO && std::is_same<T, float>::value) {
        // FP32: weighted neighbour sums per weight class, minus the centre times
        // the total weight (every cell enters each class once, so the weights
        // sum to ssum (corner + edge0 + edge1 + edge2)).  This skips the 20
        // differences P - c; the rounding it adds is about eps32 (n / 2 pi)^2
        // relative on the smoothest modes (5e-4 at 256^3), far below what one
        // V-cycle removes, and the FP64 residual keeps the difference form.
        T ac = 0, a2 = 0, a1 = 0, a0 = 0;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                const T lo = S0[RO + bb][aa], hi = S1[RO + bb][aa];
                ac += lo * P0[RO + 2 * bb][2 * aa] + hi * P2[RO + 2 * bb][2 * aa];
                a2 += (lo + hi) * P1[RO + 2 * bb][2 * aa];
            }
#pragma unroll
        for (int aa = 0; aa < 2; ++aa) {
            a1 += (S0[RO][aa] + S0[RO + 1][aa]) * P0[RO + 1][2 * aa] +
                  (S1[RO][aa] + S1[RO + 1][aa]) * P2[RO + 1][2 * aa];
            a0 += (S0[RO + aa][0] + S0[RO + aa][1]) * P0[RO + 2 * aa][1] +
                  (S1[RO + aa][0] + S1[RO + aa][1]) * P2[RO + 2 * aa][1];
        }
        const T s0 = S0[RO][0] + S0[RO][1] + S0[RO + 1][0] + S0[RO + 1][1];
        const T s1 = S1[RO][0] + S1[RO][1] + S1[RO + 1][0] + S1[RO + 1][1];
        ssum = s0 + s1;
        const T wsum = cf.corner + cf.edge[0] + cf.edge[1] + cf.edge[2];
        return cf.corner * ac + cf.edge[2] * a2 + cf.edge[1] * a1 + cf.edge[0] * a0 - c * (ssum * wsum);
    }
    T acc = 0;
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
        for (int aa = 0; aa < 2; ++aa) {
            const T lo = S0[RO + bb][aa], hi = S1[RO + bb][aa];
            // corners (one cell each)
            acc += cf.corner * (lo * (P0[RO + 2 * bb][2 * aa] - c) + hi * (P2[RO + 2 * bb][2 * aa] - c));
            // xy edges (dz = 0): the cell below and above
            acc += cf.edge[2] * (lo + hi) * (P1[RO + 2 * bb][2 * aa] - c);
        }
#pragma unroll
    for (int aa = 0; aa < 2; ++aa) {
        // xz edges (dy = 0) and yz edges (dx = 0)
        acc += cf.edge[1] * ((S0[RO][aa] + S0[RO + 1][aa]) * (P0[RO + 1][2 * aa] - c) +
                             (S1[RO][aa] + S1[RO + 1][aa]) * (P2[RO + 1][2 * aa] - c));
        acc += cf.edge[0] * ((S0[RO + aa][0] + S0[RO + aa][1]) * (P0[RO + 2 * aa][1] - c) +
                             (S1[RO + aa][0] + S1[RO + aa][1]) * (P2[RO + 2 * aa][1] - c));
    }
    const T s0 = S0[RO][0] + S0[RO][1] + S0[RO + 1][0] + S0[RO + 1][1];
    const T s1 = S1[RO][0] + S1[RO][1] + S1[RO + 1][0] + S1[RO + 1][1];
    if (!ISO) {
#pragma unroll
        for (int aa = 0; aa < 2; ++aa) {
            acc += cf.face[0] * (S0[RO][aa] + S0[RO + 1][aa] + S1[RO][aa] + S1[RO + 1][aa]) *
                   (P1[RO + 1][2 * aa] - c);
            acc += cf.face[1] * (S0[RO + aa][0] + S0[RO + aa][1] + S1[RO + aa][0] + S1[RO + aa][1]) *
                   (P1[RO + 2 * aa][1] - c);
        }
        acc += cf.face[2] * (s0 * (P0[RO + 1][1] - c) + s1 * (P2[RO + 1][1] - c));
    }
    ssum = s0 + s1;
    return acc;
}

// TXP: tile width in x (64: one wave per node row; 16: a wave spans 4 rows of
// 16, 16 x 16 threads, for tile-listed launches over partly used levels,
// where narrow tiles skip more of the unused volume)
template <class T, class TC, int MODE, bool ISO, bool FINE, int NPT, bool MASK = false,
          int TXP = 64>
static __global__ void __launch_bounds__(TXP == 64 && NPT == 1 ? 512 : 256)
    nodal_stencil_kernel(SArgs<T, TC> a) {
    using TA = acc_t<T>;  // arithmetic type (T is the storage type of x, b, sigma)
    static_assert(TXP == 64 || (TXP == 16 && NPT == 1), "tile shapes: 64 x 4NPT or 16 x 16");
    // threads per row, tile rows, thread rows
    constexpr int TX = TXP, NY = TXP == 16 ? 16 : (NPT == 1 ? 8 : 4 * NPT), TYT = NY / NPT;
    constexpr int NT = TX * TYT;
    constexpr int PR = NY + 2, PC = TX + 2;          // phi tile incl. halo
    constexpr int SR = NY + 1, SC = TX + 1;          // sigma tile
    constexpr int NQP = (PR + TYT - 1) / TYT, NQS = (SR + TYT - 1) / TYT;
    __shared__ TA sphi[2][PR][PC];
    __shared__ TA ssig[2][SR][SC];
    __shared__ double red[NT / 64];
    const KGeom& g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int nb = a.gx * a.gy * a.gz;
    const int lb = a.tiles ? a.tiles[blockIdx.x] : xcd_remap(blockIdx.x, nb);
    const int bx = lb % a.gx, by = (lb / a.gx) % a.gy, bz = lb / (a.gx * a.gy);
    const int i0 = bx * TX, j0 = by * NY, k0 = bz * a.kz;
    const int k1 = min(k0 + a.kz, g.nn[2]);
    const int i = i0 + tx, jr = j0 + NPT * ty;   // first node row of this thread
    const long sxy = (long)g.nn[0] * g.nn[1];
    const long ssxy = (long)g.sg[0] * g.sg[1];

    // fixed per-thread load slots (indices clamped; stores to LDS predicated)
    const int gi = wrapc(i, g.nn[0], g.per[0]);
    int prow[NQP];
#pragma unroll
    for (int q = 0; q < NQP; ++q) prow[q] = wrapc(j0 - 1 + min(ty + TYT * q, PR - 1), g.nn[1], g.per[1]);
    const int hr = min(tid >> 1, PR - 1);
    const int hcol = wrapc((tid & 1) ? i0 + TX : i0 - 1, g.nn[0], g.per[0]);
    const int hrow = wrapc(j0 - 1 + hr, g.nn[1], g.per[1]);
    const bool hst = tid < 2 * PR;
    const int si = min(i0 + tx, g.sg[0] - 1);
    int srow[NQS];
#pragma unroll
    for (int q = 0; q < NQS; ++q) srow[q] = min(j0 + min(ty + TYT * q, SR - 1), g.sg[1] - 1);
    const int sxr = min(j0 + min(tid, SR - 1), g.sg[1] - 1);
    const int sxc = min(i0 + TX, g.sg[0] - 1);
    const bool sxst = tid < SR;
    // in-plane byte offsets of every load and store of this thread; the plane
    // bases (k * plane size) are uniform and added as pointers
    unsigned opx[NQP], opc[NQP], ohx, ohc, osg[NQS], osx, obb[NPT], oox[NPT], ooc[NPT], oel[NPT];
#pragma unroll
    for (int q = 0; q < NQP; ++q) {
        const unsigned e = (unsigned)(gi + g.nn[0] * prow[q]);
        opx[q] = e * (unsigned)sizeof(T);
        opc[q] = e * (unsigned)sizeof(TC);
    }
    {
        const unsigned e = (unsigned)(hcol + g.nn[0] * hrow);
        ohx = e * (unsigned)sizeof(T);
        ohc = e * (unsigned)sizeof(TC);
    }
#pragma unroll
    for (int q = 0; q < NQS; ++q) osg[q] = (unsigned)(si + g.sg[0] * srow[q]) * (unsigned)sizeof(T);
    osx = (unsigned)(sxc + g.sg[0] * sxr) * (unsigned)sizeof(T);
#pragma unroll
    for (int n = 0; n < NPT; ++n) {
        obb[n] = (unsigned)(min(i, g.nn[0] - 1) + g.nn[0] * min(jr + n, g.nn[1] - 1)) * (unsigned)sizeof(T);
        const unsigned e = (unsigned)(i + g.nn[0] * (jr + n));  // used only for nodes inside
        oel[n] = e;
        oox[n] = e * (unsigned)sizeof(T);
        ooc[n] = e * (unsigned)sizeof(TC);
    }

    const double cs = (MODE == KS_UPDRES && a.cor_scale != 0.0) ? a.cor_scale : 1.0;
    const double rsc = (a.rscale != 0.0) ? a.rscale : 1.0;
    const double x1s = (a.x1scale != 0.0) ? a.x1scale : 1.0;
    // registers of one plane's loads: phi rows + halo, sigma rows + column, rhs
    struct LoadSet {
        TA lp[NQP], lh, ls[NQS], lsx, lb[NPT];
    };
    LoadSet L0, L1;
#pragma unroll
    for (int n = 0; n < NPT; ++n) L0.lb[n] = L1.lb[n] = 0;
    auto issue = [&](LoadSet& ld_, int kk, int kb) {
        TA* lp = ld_.lp;
        TA& lh = ld_.lh;
        TA* ls = ld_.ls;
        TA& lsx = ld_.lsx;
        TA* lb_ = ld_.lb;
        const long kw = (long)wrapc(kk, g.nn[2], g.per[2]) * sxy;
        const long ks = (long)min(max(kk, 0), g.sg[2] - 1) * ssxy;
        if (MODE != KS_JFIRST) {
            const T* xk = a.x + kw;
            const TC* ck = (MODE == KS_UPDRES) ? a.cor + kw : nullptr;
#pragma unroll
            for (int q = 0; q < NQP; ++q) {
                TA v = ld<TA>(*byte_add(xk, opx[q]));
                if (MODE == KS_UPDRES) v += (TA)(cs * (double)ld<acc_t<TC>>(*byte_add(ck, opc[q])));
                lp[q] = v;
            }
            TA v = ld<TA>(*byte_add(xk, ohx));
            if (MODE == KS_UPDRES) v += (TA)(cs * (double)ld<acc_t<TC>>(*byte_add(ck, ohc)));
            lh = v;
        }
        const T* sk = a.sig + ks;
#pragma unroll
        for (int q = 0; q < NQS; ++q) ls[q] = ld<TA>(*byte_add(sk, osg[q]));
        lsx = ld<TA>(*byte_add(sk, osx));
        if (MODE != KS_APPLY) {
            const T* bk = a.b + (long)min(max(kb, 0), g.nn[2] - 1) * sxy;
#pragma unroll
            for (int n = 0; n < NPT; ++n) lb_[n] = ld<TA>(*byte_add(bk, obb[n]));
        }
    };

    TA P[3][NPT + 2][3], S[2][NPT + 1][2];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) P[p][r][c] = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[p][r][c] = 0;
    double rmax = 0.0;

    // One step per z-plane; the register planes rotate (P: 3, S: 2), so the
    // loop is unrolled by 6 and every step addresses its planes by constant
    // indices instead of copying them down.
    // cur: the loads of plane kk (issued `ahead` planes earlier); the same
    // registers then receive plane kk + ahead
    auto step = [&](auto JJ, int kk, LoadSet& cur, int ahead) {
        constexpr int J = decltype(JJ)::value;
        constexpr int PN = J % 3, PM = (J + 2) % 3, PB = (J + 1) % 3;  // above, centre, below
        constexpr int SN = J % 2, SB = (J + 1) % 2;                     // cell planes above, below
        const int buf = kk & 1;
        if (MODE != KS_JFIRST) {
#pragma unroll
            for (int q = 0; q < NQP; ++q)
                if (ty + TYT * q < PR) sphi[buf][ty + TYT * q][tx + 1] = cur.lp[q];
            if (hst) sphi[buf][hr][(tid & 1) ? PC - 1 : 0] = cur.lh;
        }
#pragma unroll
        for (int q = 0; q < NQS; ++q)
            if (ty + TYT * q < SR) ssig[buf][ty + TYT * q][tx] = cur.ls[q];
        if (sxst) ssig[buf][min(tid, SR - 1)][SC - 1] = cur.lsx;
        TA bcur[NPT];  // rhs of plane kk-1
#pragma unroll
        for (int n = 0; n < NPT; ++n) bcur[n] = cur.lb[n];
        if (kk + ahead <= k1) issue(cur, kk + ahead, kk + ahead - 1);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                P[PN][r][c] = (MODE == KS_JFIRST) ? (TA)0 : sphi[buf][NPT * ty + r][tx + c];
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[SN][r][c] = ssig[buf][NPT * ty + r][tx + c];
        if (kk >= k0 + 1) {
            const int k = kk - 1;
#pragma unroll
            for (int n = 0; n < NPT; ++n) {
                const int j = jr + n;
                if (!(i < g.nn[0] && j < g.nn[1])) continue;
                if (MODE == KS_JFIRST) {
                    TA ss = 0;
#pragma unroll
                    for (int r = 0; r < 2; ++r)
                        ss += S[SB][n + r][0] + S[SB][n + r][1] + S[SN][n + r][0] + S[SN][n + r][1];
                    const bool fixed = MASK && a.nmask[sxy * k + oel[n]] != 1;
                    *byte_add(a.xout + sxy * k, oox[n]) =
                        st<T>((is_dir(g, i, j, k) || ss == (TA)0 || fixed)
                                  ? (TA)0 : a.omega * bcur[n] * diag_rcp(-a.c.diag4 * ss));
                    continue;
                }
                TA ssum;
                // row offset n as a template argument (n is a constant after unrolling)
                const TA acc =
                    (n == 0) ? apply27<TA, ISO, 0>(P[PB], P[PM], P[PN], S[SB], S[SN], a.c, ssum)
                    : (n == 1) ? apply27<TA, ISO, (NPT > 1 ? 1 : 0)>(P[PB], P[PM], P[PN], S[SB], S[SN], a.c, ssum)
                    : (n == 2) ? apply27<TA, ISO, (NPT > 2 ? 2 : 0)>(P[PB], P[PM], P[PN], S[SB], S[SN], a.c, ssum)
                               : apply27<TA, ISO, (NPT > 3 ? 3 : 0)>(P[PB], P[PM], P[PN], S[SB], S[SN], a.c, ssum);
                const TA c = P[PM][n + 1][1];
                const long kp = sxy * k;  // uniform plane base
                const int nt = MASK ? a.nmask[kp + oel[n]] : 1;
                const bool dir = is_dir(g, i, j, k);
                const TA bv = bcur[n];
                if (MODE == KS_JACOBI) {
                    TA v = c;
                    if (dir || nt == 0) v = 0;
                    else if (nt == 1 && ssum != (TA)0)
                        v = c + (bv - acc) * (a.omega * diag_rcp(-a.c.diag4 * ssum));
                    *byte_add(a.xout + kp, oox[n]) = st<T>(v);
                } else if (MODE == KS_APPLY) {
                    *byte_add(a.rout + kp, ooc[n]) = st<TC>(acc);
                } else {
                    const TA r = (dir || nt != 1) ? (TA)0 : (bv - acc);
                    *byte_add(a.rout + kp, ooc[n]) = st<TC>((acc_t<TC>)(rsc * (double)r));
                    if (MODE == KS_UPDRES) *byte_add(a.xout + kp, oox[n]) = st<T>(dir ? (TA)0 : c);
                    if (a.x1out)
                        *byte_add(a.x1out + kp, ooc[n]) =
                            st<TC>((acc_t<TC>)((dir || ssum == (TA)0)
                                                   ? 0.0
                                                   : a.omega1 * x1s * (double)r /
                                                         (-(double)a.c.diag4 * (double)ssum)));
                    rmax = fmax(rmax, nabs((double)r));
                }
            }
        }
    };
    if (sizeof(TA) == 4) {
        // FP32 arithmetic: a plain loop with the planes copied down (cheaper than
        // the 6-fold unrolled rotation); the Jacobi sweeps issue their loads two
        // planes ahead through two register sets (the sweep is bound by load
        // latency, not by issue: 18% fewer VALU instructions changed nothing,
        // the second set cut the sweep by 9%, and it fits the 128-VGPR budget
        // of 4 waves per SIMD)
        auto shift = [&]() {
#pragma unroll
            for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    P[0][r][c] = P[1][r][c];
                    P[1][r][c] = P[2][r][c];
                }
#pragma unroll
            for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
                for (int c = 0; c < 2; ++c) S[1][r][c] = S[0][r][c];
        };
        // (the FP32 residual passes keep one plane: their extra registers would
        // cost a wave per SIMD, which measured slower)
        constexpr bool two = (MODE == KS_JACOBI || MODE == KS_JFIRST);
        issue(L0, k0 - 1, k0 - 2);
        if (two && k0 <= k1) issue(L1, k0, k0 - 1);
        for (int kk = k0 - 1; kk <= k1; kk += 2) {
            shift();
            step(std::integral_constant<int, 2>{}, kk, L0, two ? 2 : 1);
            if (kk + 1 > k1) break;
            shift();
            step(std::integral_constant<int, 2>{}, kk + 1, two ? L1 : L0, two ? 2 : 1);
        }
    } else {
        // FP64: loads two planes ahead as well (the second register set keeps
        // the kernels within 256 VGPRs, i.e. at the same 2 waves per SIMD)
        constexpr int A = 2;
        issue(L0, k0 - 1, k0 - 2);
        if (A == 2 && k0 <= k1) issue(L1, k0, k0 - 1);
        LoadSet& Lb = (A == 2) ? L1 : L0;
        for (int kk = k0 - 1; kk <= k1; kk += 6) {
            step(std::integral_constant<int, 0>{}, kk, L0, A);
            if (kk + 1 > k1) break;
            step(std::integral_constant<int, 1>{}, kk + 1, Lb, A);
            if (kk + 2 > k1) break;
            step(std::integral_constant<int, 2>{}, kk + 2, L0, A);
            if (kk + 3 > k1) break;
            step(std::integral_constant<int, 3>{}, kk + 3, Lb, A);
            if (kk + 4 > k1) break;
            step(std::integral_constant<int, 4>{}, kk + 4, L0, A);
            if (kk + 5 > k1) break;
            step(std::integral_constant<int, 5>{}, kk + 5, Lb, A);
        }
    }
    if ((MODE == KS_RESID || MODE == KS_UPDRES) && a.norm) {
        const double m = block_max<NT>(rmax, red);
        if (tid == 0) atomic_max_nonneg(a.norm, m);
    }
}

// Index wrap (periodic, any i) or clamp (non-periodic).
__device__ __forceinline__ int wrapg(int i, int n, int per) {
    if (per) return ((i % n) + n) % n;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
// ghosted sigma index whose real cell is c - 1, wrapped the same way
__device__ __forceinline__ int wrapsg(int c, int n, int sgn, int per) {
    if (per) return wrapg(c - 1, n, 1) + 1;
    return c < 0 ? 0 : (c >= sgn ? sgn - 1 : c);
}

// ---- the Jacobi sweep in wave strips (FP32, no LDS, no barriers) -----------
// Each wave owns a strip of 64 columns x NR node rows (lanes 0 and 63 are the
// halo columns, lanes 1..62 write) and streams it along z on its own: every
// lane loads its column's x, sigma and b rows, and the neighbours across
// columns come from the adjacent lanes by cross-lane shuffles, so waves never
// wait for each other.  The four waves of a workgroup hold four strips
// stacked in y.  Same formula at every node as nodal_stencil_kernel's
// Jacobi pass.
constexpr int kStripTX = 62;  // written columns per strip

template <class T, bool ISO, int NR>
static __global__ void __launch_bounds__(256) nodal_jacobi_strip_kernel(SArgs<T, T> a) {
    static_assert(std::is_same<T, float>::value, "strip sweep: FP32");
    using TA = T;
    const KGeom& g = a.g;
    const int lane = threadIdx.x, wv = threadIdx.y;
    const int nb = a.gx * a.gy * a.gz;
    const int lb = xcd_remap(blockIdx.x, nb);
    const int bx = lb % a.gx, by = (lb / a.gx) % a.gy, bz = lb / (a.gx * a.gy);
    const int i = bx * kStripTX - 1 + lane;     // this lane's column
    const int jw = (by * 4 + wv) * NR;           // first node row of the strip
    const int k0 = bz * a.kz, k1 = min(k0 + a.kz, g.nn[2]);
    const long sxy = (long)g.nn[0] * g.nn[1];
    const long ssxy = (long)g.sg[0] * g.sg[1];
    const bool writes = lane >= 1 && lane <= kStripTX && i < g.nn[0];
    const int gi = wrapg(i, g.nn[0], g.per[0]);
    const int sc = wrapsg(i, g.n[0], g.sg[0], g.per[0]);  // ghosted cell left of node i
    unsigned ox[NR + 2], os[NR + 1], ob[NR];
#pragma unroll
    for (int r = 0; r < NR + 2; ++r)
        ox[r] = (unsigned)(gi + g.nn[0] * wrapg(jw - 1 + r, g.nn[1], g.per[1])) * 4u;
#pragma unroll
    for (int r = 0; r < NR + 1; ++r)
        os[r] = (unsigned)(sc + g.sg[0] * wrapsg(jw + r, g.n[1], g.sg[1], g.per[1])) * 4u;
#pragma unroll
    for (int r = 0; r < NR; ++r)
        ob[r] = (unsigned)(gi + g.nn[0] * wrapg(jw + r, g.nn[1], g.per[1])) * 4u;

    struct LoadSet {
        TA x[NR + 2], s[NR + 1], b[NR];
    };
    LoadSet L0, L1;
    auto issue = [&](LoadSet& ld_, int kk) {
        const T* xk = a.x + (long)wrapg(kk, g.nn[2], g.per[2]) * sxy;
        const T* sk = a.sig + (long)wrapsg(kk, g.n[2], g.sg[2], g.per[2]) * ssxy;
        const T* bk = a.b + (long)wrapg(kk - 1, g.nn[2], g.per[2]) * sxy;
#pragma unroll
        for (int r = 0; r < NR + 2; ++r) ld_.x[r] = *byte_add(xk, ox[r]);
#pragma unroll
        for (int r = 0; r < NR + 1; ++r) ld_.s[r] = *byte_add(sk, os[r]);
#pragma unroll
        for (int r = 0; r < NR; ++r) ld_.b[r] = *byte_add(bk, ob[r]);
    };
    TA P[3][NR + 2][3], S[2][NR + 1][2];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < NR + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) P[p][r][c] = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < NR + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[p][r][c] = 0;
    auto step = [&](int kk, LoadSet& cur) {
        TA bcur[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) bcur[r] = cur.b[r];
#pragma unroll
        for (int r = 0; r < NR + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                P[0][r][c] = P[1][r][c];
                P[1][r][c] = P[2][r][c];
            }
#pragma unroll
        for (int r = 0; r < NR + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[1][r][c] = S[0][r][c];
        // the new plane: own column, left and right neighbours from the lanes beside
#pragma unroll
        for (int r = 0; r < NR + 2; ++r) {
            const TA v = cur.x[r];
            P[2][r][0] = __shfl_up(v, 1, 64);
            P[2][r][1] = v;
            P[2][r][2] = __shfl_down(v, 1, 64);
        }
#pragma unroll
        for (int r = 0; r < NR + 1; ++r) {
            const TA v = cur.s[r];
            S[0][r][0] = v;
            S[0][r][1] = __shfl_down(v, 1, 64);
        }
        if (kk + 2 <= k1) issue(cur, kk + 2);
        if (kk >= k0 + 1) {
            const int k = kk - 1;
            const long kp = sxy * k;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int j = jw + r;
                if (!(writes && j < g.nn[1])) continue;
                TA ssum;
                TA acc;
                if (r == 0) acc = apply27<TA, ISO, 0>(P[0], P[1], P[2], S[1], S[0], a.c, ssum);
                else if (r == 1) acc = apply27<TA, ISO, (NR > 1 ? 1 : 0)>(P[0], P[1], P[2], S[1], S[0], a.c, ssum);
                else if (r == 2) acc = apply27<TA, ISO, (NR > 2 ? 2 : 0)>(P[0], P[1], P[2], S[1], S[0], a.c, ssum);
                else acc = apply27<TA, ISO, (NR > 3 ? 3 : 0)>(P[0], P[1], P[2], S[1], S[0], a.c, ssum);
                const TA c = P[1][r + 1][1];
                TA v = c;
                if (is_dir(g, i, j, k)) v = 0;
                else if (ssum != (TA)0) v = c + (bcur[r] - acc) * (a.omega * diag_rcp(-a.c.diag4 * ssum));
                a.xout[kp + i + (long)g.nn[0] * j] = v;
            }
        }
    };
    issue(L0, k0 - 1);
    if (k0 <= k1) issue(L1, k0);
    for (int kk = k0 - 1; kk <= k1; kk += 2) {
        step(kk, L0);
        if (kk + 1 > k1) break;
        step(kk + 1, L1);
    }
}

// ---- two Jacobi sweeps in one pass (FP32) ----------------------------------
// x2 = J(J(x)): the first sweep runs on a 64 x 8-node tile (the register
// layout of nodal_stencil_kernel<..., KS_JACOBI, ..., 2>), its values go into
// a ring of three LDS planes, and the second sweep, one plane behind, writes
// the 62 x 6 interior nodes.  Tiles overlap by one node on each side, so every
// node of the second sweep sees first-sweep values computed from the same
// inputs as the unfused sweep; the pair reads x, b and sigma once (plus the
// overlap) instead of twice and writes x once.

constexpr int kPairTX = 62, kPairTY = 6;  // output nodes per tile

template <class T, bool ISO>
static __global__ void __launch_bounds__(256) nodal_jacobi_pair_kernel(SArgs<T, T> a) {
    static_assert(std::is_same<T, float>::value, "the fused pair runs in FP32");
    using TA = T;
    constexpr int TX = 64, NY = 8, NPT = 2, TYT = 4, NT = 256;
    constexpr int PR = NY + 2, PC = TX + 2, SR = NY + 1, SC = TX + 1;
    constexpr int NQP = (PR + TYT - 1) / TYT, NQS = (SR + TYT - 1) / TYT;
    __shared__ TA sphi[2][PR][PC];
    __shared__ TA ssig[2][SR][SC];
    __shared__ TA s1[3][NY][TX];  // first-sweep planes
    const KGeom& g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int nb = a.gx * a.gy * a.gz;
    const int lb = xcd_remap(blockIdx.x, nb);
    const int bx = lb % a.gx, by = (lb / a.gx) % a.gy, bz = lb / (a.gx * a.gy);
    const int i0 = bx * kPairTX - 1, j0 = by * kPairTY - 1, k0 = bz * a.kz;
    const int k1 = min(k0 + a.kz, g.nn[2]);  // output planes k0 .. k1-1
    const int i = i0 + tx, jr = j0 + NPT * ty;
    const long sxy = (long)g.nn[0] * g.nn[1];
    const long ssxy = (long)g.sg[0] * g.sg[1];
    const bool pout_x = tx >= 1 && tx <= kPairTX && i < g.nn[0];

    // load slots (all indices wrapped / clamped into the arrays)
    const int gi = wrapg(i, g.nn[0], g.per[0]);
    int prow[NQP];
#pragma unroll
    for (int q = 0; q < NQP; ++q)
        prow[q] = wrapg(j0 - 1 + min(ty + TYT * q, PR - 1), g.nn[1], g.per[1]);
    const int hr = min(tid >> 1, PR - 1);
    const int hcol = wrapg((tid & 1) ? i0 + TX : i0 - 1, g.nn[0], g.per[0]);
    const int hrow = wrapg(j0 - 1 + hr, g.nn[1], g.per[1]);
    const bool hst = tid < 2 * PR;
    const int si = wrapsg(i0 + tx, g.n[0], g.sg[0], g.per[0]);
    int srow[NQS];
#pragma unroll
    for (int q = 0; q < NQS; ++q)
        srow[q] = wrapsg(j0 + min(ty + TYT * q, SR - 1), g.n[1], g.sg[1], g.per[1]);
    const int sxr = wrapsg(j0 + min(tid, SR - 1), g.n[1], g.sg[1], g.per[1]);
    const int sxc = wrapsg(i0 + TX, g.n[0], g.sg[0], g.per[0]);
    const bool sxst = tid < SR;
    unsigned opx[NQP], ohx, osg[NQS], osx, obb[NPT];
#pragma unroll
    for (int q = 0; q < NQP; ++q) opx[q] = (unsigned)(gi + g.nn[0] * prow[q]) * 4u;
    ohx = (unsigned)(hcol + g.nn[0] * hrow) * 4u;
#pragma unroll
    for (int q = 0; q < NQS; ++q) osg[q] = (unsigned)(si + g.sg[0] * srow[q]) * 4u;
    osx = (unsigned)(sxc + g.sg[0] * sxr) * 4u;
#pragma unroll
    for (int n = 0; n < NPT; ++n)
        obb[n] = (unsigned)(gi + g.nn[0] * wrapg(jr + n, g.nn[1], g.per[1])) * 4u;

    struct LoadSet {
        TA lp[NQP], lh, ls[NQS], lsx, lb[NPT];
    };
    LoadSet L0, L1;
    auto issue = [&](LoadSet& ld_, int kk) {
        const T* xk = a.x + (long)wrapg(kk, g.nn[2], g.per[2]) * sxy;
#pragma unroll
        for (int q = 0; q < NQP; ++q) ld_.lp[q] = *byte_add(xk, opx[q]);
        ld_.lh = *byte_add(xk, ohx);
        const T* sk = a.sig + (long)wrapsg(kk, g.n[2], g.sg[2], g.per[2]) * ssxy;
#pragma unroll
        for (int q = 0; q < NQS; ++q) ld_.ls[q] = *byte_add(sk, osg[q]);
        ld_.lsx = *byte_add(sk, osx);
        const T* bk = a.b + (long)wrapg(kk - 1, g.nn[2], g.per[2]) * sxy;
#pragma unroll
        for (int n = 0; n < NPT; ++n) ld_.lb[n] = *byte_add(bk, obb[n]);
    };

    TA P[3][NPT + 2][3], S[3][NPT + 1][2], bprev[NPT];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) P[p][r][c] = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[p][r][c] = 0;
#pragma unroll
    for (int n = 0; n < NPT; ++n) bprev[n] = 0;
    const TA wdg = a.omega;

    // step kk: x / sigma plane kk and b plane kk-1 arrive; the first sweep
    // computes node plane kk-1, the second node plane kk-2
    auto step = [&](int kk, LoadSet& cur) {
        const int buf = kk & 1;
#pragma unroll
        for (int q = 0; q < NQP; ++q)
            if (ty + TYT * q < PR) sphi[buf][ty + TYT * q][tx + 1] = cur.lp[q];
        if (hst) sphi[buf][hr][(tid & 1) ? PC - 1 : 0] = cur.lh;
#pragma unroll
        for (int q = 0; q < NQS; ++q)
            if (ty + TYT * q < SR) ssig[buf][ty + TYT * q][tx] = cur.ls[q];
        if (sxst) ssig[buf][min(tid, SR - 1)][SC - 1] = cur.lsx;
        TA bcur[NPT];
#pragma unroll
        for (int n = 0; n < NPT; ++n) bcur[n] = cur.lb[n];
        if (kk + 2 <= k1 + 1) issue(cur, kk + 2);
        __syncthreads();
        // rotate: planes below <- centre <- above; sigma: older <- ... <- newest
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                P[0][r][c] = P[1][r][c];
                P[1][r][c] = P[2][r][c];
                P[2][r][c] = sphi[buf][NPT * ty + r][tx + c];
            }
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                S[2][r][c] = S[1][r][c];
                S[1][r][c] = S[0][r][c];
                S[0][r][c] = ssig[buf][NPT * ty + r][tx + c];
            }
        // first sweep, node plane kk-1 (the whole 64 x 8 tile)
        const int ka = kk - 1;
        if (ka >= k0 - 1) {
            const int kaw = wrapg(ka, g.nn[2], g.per[2]);
#pragma unroll
            for (int n = 0; n < NPT; ++n) {
                TA ssum;
                const TA acc = (n == 0) ? apply27<TA, ISO, 0>(P[0], P[1], P[2], S[1], S[0], a.c, ssum)
                                        : apply27<TA, ISO, 1>(P[0], P[1], P[2], S[1], S[0], a.c, ssum);
                const TA c = P[1][n + 1][1];
                const int jw = wrapg(jr + n, g.nn[1], g.per[1]);
                TA v = c;
                if (is_dir(g, gi, jw, kaw)) v = 0;
                else if (ssum != (TA)0) v = c + (bcur[n] - acc) * (wdg * diag_rcp(-a.c.diag4 * ssum));
                s1[(ka + 3) % 3][NPT * ty + n][tx] = v;
            }
        }
        __syncthreads();
        // second sweep, node plane kk-2 (the 62 x 6 interior)
        const int kb = kk - 2;
        if (kb >= k0 && kb < k1) {
            const int pa = (kb + 2) % 3, pm = (kb + 3) % 3, pn = (kb + 4) % 3;  // below, centre, above
#pragma unroll
            for (int n = 0; n < NPT; ++n) {
                const int rr = NPT * ty + n;  // tile row of this node
                const int j = jr + n;
                if (!(pout_x && rr >= 1 && rr <= kPairTY && j < g.nn[1])) continue;
                TA Q0[3][3], Q1[3][3], Q2[3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        Q0[r][c] = s1[pa][rr - 1 + r][tx - 1 + c];
                        Q1[r][c] = s1[pm][rr - 1 + r][tx - 1 + c];
                        Q2[r][c] = s1[pn][rr - 1 + r][tx - 1 + c];
                    }
                TA S0[2][2], S1[2][2];
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        S0[r][c] = S[2][n + r][c];
                        S1[r][c] = S[1][n + r][c];
                    }
                TA ssum;
                const TA acc = apply27<TA, ISO, 0>(Q0, Q1, Q2, S0, S1, a.c, ssum);
                const TA c = Q1[1][1];
                TA v = c;
                if (is_dir(g, i, j, kb)) v = 0;
                else if (ssum != (TA)0) v = c + (bprev[n] - acc) * (wdg * diag_rcp(-a.c.diag4 * ssum));
                a.xout[sxy * kb + i + (long)g.nn[0] * j] = v;
            }
        }
#pragma unroll
        for (int n = 0; n < NPT; ++n) bprev[n] = bcur[n];
    };
    issue(L0, k0 - 2);
    issue(L1, k0 - 1);
    for (int kk = k0 - 2; kk <= k1 + 1; kk += 2) {
        step(kk, L0);
        if (kk + 1 > k1 + 1) break;
        step(kk + 1, L1);
    }
}

// ---- the Jacobi sweep with LDS-direct plane loads (4-byte storage) ---------
// Same tile, z-streaming and arithmetic as nodal_stencil_kernel<..., KS_JACOBI,
// ..., 2> (64 x 8 nodes, 64 x 4 threads, 2 nodes per thread), but every plane's
// x, sigma and b rows go from global memory straight into a ring of R plane
// slots in LDS (global_load_lds_dword: 64 lanes fill 64 consecutive words),
// R - 1 planes ahead, with no staging registers and no LDS stores.  A plane
// is 29 row loads (10 x rows, their two halo columns, 9 sigma rows and their
// extra column, 8 b rows), dealt round-robin to the 4 waves; each wave waits
// for its own loads of a plane with s_waitcnt vmcnt (loads return in order,
// so the count of loads issued after it is the bound) and a barrier then
// publishes the plane to the workgroup.  A slot is refilled only after the
// barrier that follows the last reads of it.
constexpr int kDmaRowLoads = 29;
constexpr int kDmaMinLoadsPerWave = kDmaRowLoads / 4;  // 7 (wave 0 issues 8)

template <int N>
__device__ __forceinline__ void wait_vmcnt_lgkm0() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    // gfx9 encoding: vmcnt[3:0], expcnt[6:4] (7: no wait), lgkmcnt[11:8], vmcnt[15:14]
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (0 << 8) | ((N >> 4) << 14));
}

// workgroup barrier without a memory fence (a fence would wait for every
// outstanding load, including the planes in flight); the "memory" clobber
// keeps the compiler from moving LDS accesses across it
__device__ __forceinline__ void barrier_nofence() { asm volatile("s_barrier" ::: "memory"); }

template <class T, bool ISO, bool FINE, int R>
static __global__ void __launch_bounds__(256) nodal_jacobi_dma_kernel(SArgs<T, T> a) {
    static_assert(sizeof(T) == 4, "LDS-direct sweep: 4-byte storage");
    using TA = acc_t<T>;
    constexpr int TX = 64, NY = 8, NPT = 2, TYT = 4, D = R - 1;
    constexpr int PR = NY + 2, SR = NY + 1;
    // slot layout (words): x rows | x halo columns (left rows 0..9, right rows
    // 0..9, rest unused) | sigma rows | sigma extra column | b rows
    constexpr int O_PH = 0, O_PHH = O_PH + PR * TX, O_SG = O_PHH + TX, O_SGX = O_SG + SR * TX,
                  O_B = O_SGX + TX, SLOT = O_B + NY * TX;
    __shared__ T lds[R][SLOT];
    const KGeom& g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int nb = a.gx * a.gy * a.gz;
    const int lb = a.tiles ? a.tiles[blockIdx.x] : xcd_remap(blockIdx.x, nb);
    const int bx = lb % a.gx, by = (lb / a.gx) % a.gy, bz = lb / (a.gx * a.gy);
    const int i0 = bx * TX, j0 = by * NY, k0 = bz * a.kz;
    const int k1 = min(k0 + a.kz, g.nn[2]);
    const int i = i0 + tx, jr = j0 + NPT * ty;
    const long sxy = (long)g.nn[0] * g.nn[1];
    const long ssxy = (long)g.sg[0] * g.sg[1];

    // this lane's in-plane byte offset for each of the wave's row loads
    // (indices wrapped / clamped so every address is valid)
    constexpr int NQ = (kDmaRowLoads + 3) / 4;  // 8
    unsigned off[NQ];
    int kind[NQ], lds_off[NQ];  // kind: 0 x, 1 sigma, 2 b, -1 none
#pragma unroll
    for (int s = 0; s < NQ; ++s) {
        const int q = ty + 4 * s;
        kind[s] = -1;
        off[s] = 0;
        lds_off[s] = 0;
        if (q >= kDmaRowLoads) continue;
        if (q < PR) {  // x row j0-1+q, interior columns
            const int row = wrapc(j0 - 1 + q, g.nn[1], g.per[1]);
            const int col = wrapc(i0 + tx, g.nn[0], g.per[0]);
            kind[s] = 0;
            off[s] = (unsigned)(col + g.nn[0] * row) * 4u;
            lds_off[s] = O_PH + q * TX;
        } else if (q == PR) {  // x halo columns: lanes 0..9 left, 10..19 right
            const int r = tx < PR ? tx : (tx < 2 * PR ? tx - PR : 0);
            const int row = wrapc(j0 - 1 + r, g.nn[1], g.per[1]);
            const int col = wrapc(tx >= PR && tx < 2 * PR ? i0 + TX : i0 - 1, g.nn[0], g.per[0]);
            kind[s] = 0;
            off[s] = (unsigned)(col + g.nn[0] * row) * 4u;
            lds_off[s] = O_PHH;
        } else if (q < PR + 1 + SR) {  // sigma row j0+r (ghosted indexing)
            const int r = q - PR - 1;
            const int row = min(j0 + r, g.sg[1] - 1);
            const int col = min(i0 + tx, g.sg[0] - 1);
            kind[s] = 1;
            off[s] = (unsigned)(col + g.sg[0] * row) * 4u;
            lds_off[s] = O_SG + r * TX;
        } else if (q == PR + 1 + SR) {  // sigma extra column i0+64, lanes 0..8
            const int row = min(j0 + min(tx, SR - 1), g.sg[1] - 1);
            const int col = min(i0 + TX, g.sg[0] - 1);
            kind[s] = 1;
            off[s] = (unsigned)(col + g.sg[0] * row) * 4u;
            lds_off[s] = O_SGX;
        } else {  // b row j0+r
            const int r = q - PR - 2 - SR;
            const int row = min(j0 + r, g.nn[1] - 1);
            const int col = min(i0 + tx, g.nn[0] - 1);
            kind[s] = 2;
            off[s] = (unsigned)(col + g.nn[0] * row) * 4u;
            lds_off[s] = O_B + r * TX;
        }
    }
    // issue plane kk's row loads of this wave into its slot
    auto issue = [&](int kk) {
        T* slot = lds[(kk + R) % R];
        const T* xk = a.x + (long)wrapc(kk, g.nn[2], g.per[2]) * sxy;
        const T* sk = a.sig + (long)min(max(kk, 0), g.sg[2] - 1) * ssxy;
        const T* bk = a.b + (long)min(max(kk - 1, 0), g.nn[2] - 1) * sxy;
#pragma unroll
        for (int s = 0; s < NQ; ++s) {
            // kind and LDS offset depend on the wave only: made scalar, the
            // choice is a scalar branch and the load takes a scalar base
            const int kd = __builtin_amdgcn_readfirstlane(kind[s]);
            if (kd < 0) continue;
            const T* base = kd == 0 ? xk : (kd == 1 ? sk : bk);
            __builtin_amdgcn_global_load_lds(byte_add(base, off[s]),
                                             slot + __builtin_amdgcn_readfirstlane(lds_off[s]), 4,
                                             0, 0);
        }
    };
    // wait until this wave's loads of the plane `later` planes before the
    // newest issued one have landed, then publish the slot
    auto wait_plane = [&](int later) {
        if (later >= 2) wait_vmcnt_lgkm0<2 * kDmaMinLoadsPerWave>();
        else if (later == 1) wait_vmcnt_lgkm0<kDmaMinLoadsPerWave>();
        else wait_vmcnt_lgkm0<0>();
        barrier_nofence();
    };
    static_assert(D <= 3, "wait_plane covers up to 2 planes in flight behind the one waited for");

    TA P[3][NPT + 2][3], S[2][NPT + 1][2];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) P[p][r][c] = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[p][r][c] = 0;

    for (int p = 0; p < D; ++p)
        if (k0 - 1 + p <= k1) issue(k0 - 1 + p);
    for (int kk = k0 - 1; kk <= k1; ++kk) {
        wait_plane(min(D - 1, k1 - kk));
        if (kk + D <= k1) issue(kk + D);
        const T* slot = lds[(kk + R) % R];
        // rotate the register planes: below <- centre <- above, sigma below <- above
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                P[0][r][c] = P[1][r][c];
                P[1][r][c] = P[2][r][c];
            }
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) S[1][r][c] = S[0][r][c];
#pragma unroll
        for (int r = 0; r < NPT + 2; ++r) {
            const int row = NPT * ty + r;
            P[2][r][0] = tx == 0 ? slot[O_PHH + row] : slot[O_PH + row * TX + tx - 1];
            P[2][r][1] = slot[O_PH + row * TX + tx];
            P[2][r][2] = tx == TX - 1 ? slot[O_PHH + PR + row] : slot[O_PH + row * TX + tx + 1];
        }
#pragma unroll
        for (int r = 0; r < NPT + 1; ++r) {
            const int row = NPT * ty + r;
            S[0][r][0] = slot[O_SG + row * TX + tx];
            S[0][r][1] = tx == TX - 1 ? slot[O_SGX + row] : slot[O_SG + row * TX + tx + 1];
        }
        if (kk >= k0 + 1) {
            const int k = kk - 1;
            const long kp = sxy * k;
#pragma unroll
            for (int n = 0; n < NPT; ++n) {
                const int j = jr + n;
                if (!(i < g.nn[0] && j < g.nn[1])) continue;
                TA ssum;
                const TA acc = (n == 0)
                                   ? apply27<TA, ISO, 0>(P[0], P[1], P[2], S[1], S[0], a.c, ssum)
                                   : apply27<TA, ISO, 1>(P[0], P[1], P[2], S[1], S[0], a.c, ssum);
                const TA c = P[1][n + 1][1];
                const TA bv = slot[O_B + (NPT * ty + n) * TX + tx];
                TA v = c;
                if (is_dir(g, i, j, k)) v = 0;
                else if (ssum != (TA)0) v = c + (bv - acc) * (a.omega * diag_rcp(-a.c.diag4 * ssum));
                a.xout[kp + i + (long)g.nn[0] * j] = st<T>(v);
            }
        }
    }
}

// x = omega * b / diag (first Jacobi sweep from a zero iterate)
template <class T>
static __global__ void nodal_jacobi_first_kernel(KGeom g, Coef<T> c, const T* __restrict__ b,
                                          const T* __restrict__ sig, T* __restrict__ x, T omega) {
    int i, j, k;
    if (!grid_point(g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
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
                                      T* __restrict__ rc, const T* __restrict__ sigc = nullptr,
                                      T diag4c = 0, T omegac = 0, T* __restrict__ xc1 = nullptr) {
    int I, J, K;
    if (!grid_point(gc.nn[0], gc.nn[1], I, J, K)) return;
    const long id = I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K);
    if (is_dir(gc, I, J, K)) {
        rc[id] = 0;
        if (xc1) xc1[id] = 0;
        return;
    }
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
    if (xc1) {
        // first coarse Jacobi sweep from zero: omega rc / diag (same as KS_JFIRST)
        T ss = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            ss += sigc[(I + (q & 1)) + (long)gc.sg[0] * ((J + ((q >> 1) & 1)) + (long)gc.sg[1] * (K + ((q >> 2) & 1)))];
        xc1[id] = (ss != (T)0) ? omegac * acc / (-diag4c * ss) : (T)0;
    }
}

// Trilinear interpolation of coarse nodes added onto fine nodes.
template <class T>
static __global__ void nodal_prolong_add_kernel(KGeom gf, KGeom gc, const T* __restrict__ xc,
                                         T* __restrict__ xf) {
    int i, j, k;
    if (!grid_point(gf.nn[0], gf.nn[1], i, j, k)) return;
    const long id = i + (long)gf.nn[0] * (j + (long)gf.nn[1] * k);
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

// The same prolongation with one thread per coarse node: the 8 coarse values
// of the cell above (I, J, K) are loaded once and the 8 fine nodes
// (2I + a, 2J + b, 2K + c), a, b, c in {0, 1}, are updated (same sums as
// nodal_prolong_add_kernel, so the results are bitwise equal).
template <class T, class TF = T>
static __global__ void __launch_bounds__(256) nodal_prolong_add_coarse_kernel(
    KGeom gf, KGeom gc, const T* __restrict__ xc, TF* __restrict__ xf) {
    int I, J, K;
    if (!grid_point(gc.nn[0], gc.nn[1], I, J, K)) return;
    int up[3] = {I + 1, J + 1, K + 1};
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (up[d] >= gc.nn[d]) up[d] = gc.per[d] ? up[d] - gc.nn[d] : gc.nn[d] - 1;
    T c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int II = (q & 1) ? up[0] : I, JJ = (q & 2) ? up[1] : J, KK = (q & 4) ? up[2] : K;
        c[q] = xc[II + (long)gc.nn[0] * (JJ + (long)gc.nn[1] * KK)];
    }
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        const int a = f & 1, b = (f >> 1) & 1, cc = (f >> 2) & 1;
        const int i = 2 * I + a, j = 2 * J + b, k = 2 * K + cc;
        if (i >= gf.nn[0] || j >= gf.nn[1] || k >= gf.nn[2] || is_dir(gf, i, j, k)) continue;
        T acc = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            // corner q of the fine node's interpolation stencil: the upper
            // coarse node along d only when the fine index is odd there
            const int sx = (q & 1) & a, sy = ((q >> 1) & 1) & b, sz = ((q >> 2) & 1) & cc;
            acc += c[sx | (sy << 1) | (sz << 2)];
        }
        TF& dst = xf[i + (long)gf.nn[0] * (j + (long)gf.nn[1] * (long)k)];
        dst = st<TF>(ld<T>(dst) + (T)0.125 * acc);
    }
}

// Full-weighting restriction streamed along z with the fine plane staged in
// LDS: a 64 x 4 workgroup owns a 64 x 4 column of coarse nodes over kz coarse
// planes; every fine plane of its 129 x 9 footprint is read from HBM once,
// restricted in-plane from LDS and combined along z (1/4, 1/2, 1/4) in
// registers.  Same weights as nodal_restrict_kernel; optional first coarse
// Jacobi sweep from zero (xc1) as there.
template <class T, class TF = T>
static __global__ void __launch_bounds__(256) nodal_restrict_tiled_kernel(
    KGeom gf, KGeom gc, const TF* __restrict__ rf, T* __restrict__ rc, const T* __restrict__ sigc,
    T diag4c, T omegac, T* __restrict__ xc1, int kz, int gx, int gy, int gz) {
    constexpr int TX = 64, TY = 4, FC = 2 * TX + 1, FR = 2 * TY + 1;
    __shared__ T tile[FR][FC];
    const int nb = gx * gy * gz;
    const int lb = xcd_remap(blockIdx.x, nb);
    const int bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
    const int I0 = bx * TX, J0 = by * TY, K0 = bz * kz;
    const int K1 = min(K0 + kz, gc.nn[2]);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
    const int I = I0 + tx, J = J0 + ty;
    const long sxy = (long)gf.nn[0] * gf.nn[1];
    // fixed per-thread tile slots: fine (column, row) offsets and validity
    constexpr int NQ = (FR * FC + 255) / 256;
    long off[NQ];
    bool ok[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + 256 * q;
        const int r = e / FC, c = e - (e / FC) * FC;
        int fi = 2 * I0 - 1 + c, fj = 2 * J0 - 1 + r;
        bool v = e < FR * FC;
        if (gf.per[0]) { fi = fi < 0 ? fi + gf.nn[0] : (fi >= gf.nn[0] ? fi - gf.nn[0] : fi); }
        else v = v && fi >= 0 && fi < gf.nn[0];
        if (gf.per[1]) { fj = fj < 0 ? fj + gf.nn[1] : (fj >= gf.nn[1] ? fj - gf.nn[1] : fj); }
        else v = v && fj >= 0 && fj < gf.nn[1];
        fi = min(max(fi, 0), gf.nn[0] - 1);
        fj = min(max(fj, 0), gf.nn[1] - 1);
        off[q] = fi + (long)gf.nn[0] * fj;
        ok[q] = v;
    }
    T acc = 0;
    for (int p = 2 * K0 - 1; p <= 2 * K1 - 1; ++p) {
        int fk = p;
        bool zv = true;
        if (gf.per[2]) fk = fk < 0 ? fk + gf.nn[2] : (fk >= gf.nn[2] ? fk - gf.nn[2] : fk);
        else zv = fk >= 0 && fk < gf.nn[2];
        fk = min(max(fk, 0), gf.nn[2] - 1);
        T v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] = (ok[q] && zv) ? ld<T>(rf[off[q] + sxy * fk]) : (T)0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q;
            if (e < FR * FC) tile[e / FC][e - (e / FC) * FC] = v[q];
        }
        __syncthreads();
        T s2 = 0;
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy) {
            const T wy = oy == 0 ? (T)0.5 : (T)0.25;
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
                const T wx = ox == 0 ? (T)0.5 : (T)0.25;
                s2 += wx * wy * tile[2 * ty + 1 + oy][2 * tx + 1 + ox];
            }
        }
        __syncthreads();
        if ((p & 1) == 0) {
            acc += (T)0.5 * s2;
            continue;
        }
        // odd plane p = 2K + 1 closes coarse plane K = (p - 1) / 2 and opens K + 1
        if (p > 2 * K0 - 1) {
            acc += (T)0.25 * s2;
            const int K = (p - 1) >> 1;
            if (I < gc.nn[0] && J < gc.nn[1]) {
                const long id = I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K);
                const bool dir = is_dir(gc, I, J, K);
                rc[id] = dir ? (T)0 : acc;
                if (xc1) {
                    T ss = 0;
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        ss += sigc[(I + (q & 1)) + (long)gc.sg[0] * ((J + ((q >> 1) & 1)) +
                                                                   (long)gc.sg[1] * (K + ((q >> 2) & 1)))];
                    xc1[id] = (dir || ss == (T)0) ? (T)0 : omegac * acc / (-diag4c * ss);
                }
            }
        }
        acc = (T)0.25 * s2;
    }
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
    int I, J, K;
    if (!grid_point(gc.n[0], gc.n[1], I, J, K)) return;
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

// per-block sum / min / max (part[0..np), part[np..2np), part[2np..3np))
static __global__ void __launch_bounds__(256) stats_kernel(const double* __restrict__ a, long n,
                                                           double* __restrict__ part, int np) {
    __shared__ double rs[256], rmn[256], rmx[256];
    double s = 0.0, mn = 1e300, mx = -1e300;
    for (long id = (long)blockIdx.x * 256 + threadIdx.x; id < n; id += (long)gridDim.x * 256) {
        const double v = a[id];
        s += v;
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    rs[threadIdx.x] = s;
    rmn[threadIdx.x] = mn;
    rmx[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rs[threadIdx.x] += rs[threadIdx.x + o];
            rmn[threadIdx.x] = fmin(rmn[threadIdx.x], rmn[threadIdx.x + o]);
            rmx[threadIdx.x] = fmax(rmx[threadIdx.x], rmx[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rs[0];
        part[np + blockIdx.x] = rmn[0];
        part[2 * np + blockIdx.x] = rmx[0];
    }
}

// mean (written to *mean when singular, else 0) and ||a - mean||_inf (as bits in *norm)
static __global__ void __launch_bounds__(256) final_stats_kernel(const double* __restrict__ part,
                                                                 int np, long n, int singular,
                                                                 double* mean,
                                                                 unsigned long long* norm) {
    __shared__ double rs[256], rmn[256], rmx[256];
    double s = 0.0, mn = 1e300, mx = -1e300;
    for (int id = threadIdx.x; id < np; id += 256) {
        s += part[id];
        mn = fmin(mn, part[np + id]);
        mx = fmax(mx, part[2 * np + id]);
    }
    rs[threadIdx.x] = s;
    rmn[threadIdx.x] = mn;
    rmx[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rs[threadIdx.x] += rs[threadIdx.x + o];
            rmn[threadIdx.x] = fmin(rmn[threadIdx.x], rmn[threadIdx.x + o]);
            rmx[threadIdx.x] = fmax(rmx[threadIdx.x], rmx[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = singular ? rs[0] / (double)n : 0.0;
        *mean = m;
        const double d1 = rmx[0] - m, d2 = m - rmn[0];
        double bn = d1 > d2 ? d1 : d2;
        // a non-finite entry makes the norm +inf (fmax would drop a NaN)
        if (!(rs[0] == rs[0]) || !(d1 == d1) || !(d2 == d2)) bn = __builtin_inf();
        *norm = (unsigned long long)__double_as_longlong(bn > 0.0 ? bn : 0.0);
    }
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
        m = fmax(m, nabs((double)a[id]));
    m = block_max<256>(m, red);
    if (threadIdx.x == 0) atomic_max_nonneg(out, m);
}

template <class T>
static __global__ void zero_dirichlet_kernel(KGeom g, T* __restrict__ a) {
    int i, j, k;
    if (!grid_point(g.nn[0], g.nn[1], i, j, k)) return;
    if (is_dir(g, i, j, k)) a[i + (long)g.nn[0] * (j + (long)g.nn[1] * k)] = 0;
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
// (L x)_n with direct loads of the 3x3x3 neighbourhood, in TO precision
// (same unrolled arithmetic as the streaming kernel).
template <class TO, class TS, class TX>
__device__ __forceinline__ TO apply_node_nbhd(const KGeom& g, const Coef<TO>& c, const TS* sig,
                                              const TX* x, int i, int j, int k, TO& ssum) {
    TO P[3][3][3], S[2][2][2];
    int ii[3], jj[3], kk[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ii[d] = wrapc(i - 1 + d, g.nn[0], g.per[0]);
        jj[d] = wrapc(j - 1 + d, g.nn[1], g.per[1]);
        kk[d] = wrapc(k - 1 + d, g.nn[2], g.per[2]);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                P[p][r][q] = (TO)x[ii[q] + (long)g.nn[0] * (jj[r] + (long)g.nn[1] * kk[p])];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                S[p][r][q] = (TO)sig[(i + q) + (long)g.sg[0] * ((j + r) + (long)g.sg[1] * (k + p))];
    const bool iso = c.face[0] == (TO)0 && c.face[1] == (TO)0 && c.face[2] == (TO)0;
    return iso ? apply27<TO, true, 0>(P[0], P[1], P[2], S[0], S[1], c, ssum)
               : apply27<TO, false, 0>(P[0], P[1], P[2], S[0], S[1], c, ssum);
}

// (L x)_n in double; Dirichlet rows masked to 0.
template <class TS>
__device__ double apply_node_direct(const KGeom& g, const Coef<double>& c, const TS* sig,
                                    const double* x, int i, int j, int k) {
    if (is_dir(g, i, j, k)) return 0.0;
    double ssum;
    return apply_node_nbhd<double, TS, double>(g, c, sig, x, i, j, k, ssum);
}

// A node held at zero: a Dirichlet node, or mask 0 of an overset mask (m may be null).
__device__ __forceinline__ bool node_fixed(const KGeom& g, const unsigned char* m, long id, int i,
                                           int j, int k) {
    return is_dir(g, i, j, k) || (m && !m[id]);
}

template <class TS>
__device__ void bottom_apply(const KGeom& g, const Coef<double>& c, const TS* sig,
                             const unsigned char* m, const double* x, double* y, long n) {
    __syncthreads();
    for (long id = threadIdx.x; id < n; id += blockDim.x) {
        const int i = id % g.nn[0];
        const int j = (id / g.nn[0]) % g.nn[1];
        const int k = id / ((long)g.nn[0] * g.nn[1]);
        y[id] = (m && !m[id]) ? 0.0 : apply_node_direct(g, c, sig, x, i, j, k);
    }
    __syncthreads();
}

// The bottom_solver choices of MLMG (MGX_BOTTOM_*): BiCGStab (van der Vorst
// 1992), CG (Hestenes and Stiefel 1952), BiCGStab and then CG from zero when
// the first one fails ("bicgcg"), the reverse ("cgbicg"), or smoothing sweeps
// only ("smoother").  A Krylov solve that does not reach rtol is followed by
// kBottomSmooths smoothing sweeps, from zero when it broke down (a zero
// inner product).  Both operators are negative (semi-)definite and
// symmetric, so CG applies as it stands.
constexpr int kBottomSmooths = 8;

// the 7 FP64 work vectors of the bottom solve
struct BottomVecs {
    double *x, *r, *b, *p, *v, *s, *t;
};

__device__ __forceinline__ BottomVecs bottom_vecs(double* w, long n) {
    return BottomVecs{w, w + n, w + 2 * n, w + 3 * n, w + 4 * n, w + 5 * n, w + 6 * n};
}

// 0: converged, 1: broke down, 2: iteration limit.  Starts from x = 0.
template <class Op>
__device__ int bottom_bicgstab(const Op& op, const BottomVecs& V, long n, int maxiter,
                               double rtol, double rnorm0, int& iters, double* red) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double *x = V.x, *r = V.r, *rh = V.b, *p = V.p, *v = V.v, *s = V.s, *t = V.t;
    for (long id = tid; id < n; id += nt) {
        r[id] = rh[id];
        x[id] = 0.0; p[id] = 0.0; v[id] = 0.0;
    }
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    for (int it = 1; it <= maxiter; ++it) {
        iters = it;
        double q = 0.0;
        for (long id = tid; id < n; id += nt) q += rh[id] * r[id];
        const double rho_new = bsum(q, red);
        if (rho_new == 0.0) return 1;
        const double beta = (rho_new / rho) * (alpha / omega);
        for (long id = tid; id < n; id += nt) p[id] = r[id] + beta * (p[id] - omega * v[id]);
        op.apply(p, v, n);
        q = 0.0;
        for (long id = tid; id < n; id += nt) q += rh[id] * v[id];
        const double rhv = bsum(q, red);
        if (rhv == 0.0) return 1;
        alpha = rho_new / rhv;
        double sn = 0.0;
        for (long id = tid; id < n; id += nt) {
            x[id] += alpha * p[id];
            s[id] = r[id] - alpha * v[id];
            sn = fmax(sn, nabs(s[id]));
        }
        sn = bmax(sn, red);
        if (sn <= rtol * rnorm0) return 0;
        op.apply(s, t, n);
        double tt = 0.0, ts = 0.0;
        for (long id = tid; id < n; id += nt) { tt += t[id] * t[id]; ts += t[id] * s[id]; }
        tt = bsum(tt, red);
        ts = bsum(ts, red);
        if (tt == 0.0) return 1;
        omega = ts / tt;
        double rr = 0.0;
        for (long id = tid; id < n; id += nt) {
            x[id] += omega * s[id];
            r[id] = s[id] - omega * t[id];
            rr = fmax(rr, nabs(r[id]));
        }
        rr = bmax(rr, red);
        if (rr <= rtol * rnorm0) return 0;
        if (omega == 0.0) return 1;
        rho = rho_new;
    }
    return 2;
}

// Conjugate gradients from x = 0; same return codes.
template <class Op>
__device__ int bottom_cg(const Op& op, const BottomVecs& V, long n, int maxiter, double rtol,
                         double rnorm0, int& iters, double* red) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double *x = V.x, *r = V.r, *p = V.p, *v = V.v;
    for (long id = tid; id < n; id += nt) {
        r[id] = V.b[id];
        x[id] = 0.0; p[id] = 0.0;
    }
    double rho_old = 1.0;
    for (int it = 1; it <= maxiter; ++it) {
        iters = it;
        double q = 0.0;
        for (long id = tid; id < n; id += nt) q += r[id] * r[id];
        const double rho = bsum(q, red);
        if (rho == 0.0) return 0;  // the residual vanished
        const double beta = it == 1 ? 0.0 : rho / rho_old;
        for (long id = tid; id < n; id += nt) p[id] = r[id] + beta * p[id];
        op.apply(p, v, n);
        q = 0.0;
        for (long id = tid; id < n; id += nt) q += p[id] * v[id];
        const double pv = bsum(q, red);
        if (pv == 0.0) return 1;
        const double alpha = rho / pv;
        double rn = 0.0;
        for (long id = tid; id < n; id += nt) {
            x[id] += alpha * p[id];
            r[id] -= alpha * v[id];
            rn = fmax(rn, nabs(r[id]));
        }
        rn = bmax(rn, red);
        if (rn <= rtol * rnorm0) return 0;
        rho_old = rho;
    }
    return 2;
}

// The bottom solve of op's level: b_in (mean-free over the free unknowns when
// singular, zero at held ones) -> x_out (likewise; zero at held unknowns).
// Adds the Krylov iterations (or smoothing sweeps) to *iters_acc.
template <class Op, class T>
__device__ void bottom_solve_body(const Op& op, const T* b_in, T* x_out, double* work, int kind,
                                  int maxiter, double rtol, int singular, int* iters_acc,
                                  double* red) {
    const long n = op.size();
    const BottomVecs V = bottom_vecs(work, n);
    const int tid = threadIdx.x, nt = blockDim.x;
    double bs = 0.0;
    for (long id = tid; id < n; id += nt) {
        V.b[id] = op.held(id) ? 0.0 : (double)b_in[id];
        bs += V.b[id];
    }
    bs = bsum(bs, red);
    const double nfree = op.nfree();
    const double bmean = (singular && nfree > 0.0) ? bs / nfree : 0.0;
    double rn = 0.0;
    for (long id = tid; id < n; id += nt) {
        if (!op.held(id)) V.b[id] -= bmean;
        V.x[id] = 0.0;
        rn = fmax(rn, nabs(V.b[id]));
    }
    const double rnorm0 = bmax(rn, red);
    int it = 0;
    if (rnorm0 > 0.0) {
        int st = 2, smooths = 0;
        if (kind == MGX_BOTTOM_SMOOTHER) {
            smooths = kBottomSmooths;
        } else {
            const bool cg_first = kind == MGX_BOTTOM_CG || kind == MGX_BOTTOM_CGBICG;
            st = cg_first ? bottom_cg(op, V, n, maxiter, rtol, rnorm0, it, red)
                          : bottom_bicgstab(op, V, n, maxiter, rtol, rnorm0, it, red);
            if (st != 0 && (kind == MGX_BOTTOM_BICGCG || kind == MGX_BOTTOM_CGBICG)) {
                int it2 = 0;
                st = cg_first ? bottom_bicgstab(op, V, n, maxiter, rtol, rnorm0, it2, red)
                              : bottom_cg(op, V, n, maxiter, rtol, rnorm0, it2, red);
                it += it2;
            }
            if (st != 0) smooths = kBottomSmooths;
        }
        if (st == 1)
            for (long id = tid; id < n; id += nt) V.x[id] = 0.0;
        for (int q = 0; q < smooths; ++q) op.smooth(V, n);
        if (kind == MGX_BOTTOM_SMOOTHER) it = smooths;
    }
    __syncthreads();
    double xs = 0.0;
    for (long id = tid; id < n; id += nt) xs += op.held(id) ? 0.0 : V.x[id];
    xs = bsum(xs, red);
    const double xmean = (singular && nfree > 0.0) ? xs / nfree : 0.0;
    for (long id = tid; id < n; id += nt) x_out[id] = op.held(id) ? (T)0 : (T)(V.x[id] - xmean);
    if (tid == 0 && iters_acc) atomicAdd(iters_acc, it);
}

// The nodal operator in FP64 for the bottom solve; its smoother is the
// weighted Jacobi sweep of the V-cycle.
template <class TS>
struct NodalBottomOp {
    KGeom g;
    Coef<double> c;
    const TS* sig;
    const unsigned char* m;  // overset mask, may be null
    double omega;
    __device__ long size() const { return (long)g.nn[0] * g.nn[1] * g.nn[2]; }
    __device__ bool held(long id) const { return m && !m[id]; }
    // a singular problem has no held nodes (a mask that holds one makes it regular)
    __device__ double nfree() const { return (double)size(); }
    __device__ void apply(const double* x, double* y, long n) const {
        bottom_apply(g, c, sig, m, x, y, n);
    }
    __device__ void smooth(const BottomVecs& V, long n) const {
        __syncthreads();
        for (long id = threadIdx.x; id < n; id += blockDim.x) {
            const int i = id % g.nn[0];
            const int j = (id / g.nn[0]) % g.nn[1];
            const int k = id / ((long)g.nn[0] * g.nn[1]);
            double ssum;
            const double acc = apply_node_nbhd<double, TS, double>(g, c, sig, V.x, i, j, k, ssum);
            double u = V.x[id];
            if (node_fixed(g, m, id, i, j, k)) u = 0.0;
            else if (ssum != 0.0) u += omega * (V.b[id] - acc) / (-c.diag4 * ssum);
            V.t[id] = u;
        }
        __syncthreads();
        for (long id = threadIdx.x; id < n; id += blockDim.x) V.x[id] = V.t[id];
        __syncthreads();
    }
};

template <class T>
static __global__ void __launch_bounds__(1024) bottom_solve_kernel(KGeom g, Coef<double> c,
                                                                   const T* sig,
                                                                   const unsigned char* m,
                                                                   const T* b_in, T* x_out,
                                                                   double* work, int kind,
                                                                   int maxiter, double rtol,
                                                                   double omega, int singular,
                                                                   int* iters_acc) {
    __shared__ double red[16];
    const NodalBottomOp<T> op{g, c, sig, m, omega};
    bottom_solve_body(op, b_in, x_out, work, kind, maxiter, rtol, singular, iters_acc, red);
}

// ---- coarse levels: the whole V-cycle below a size threshold in one workgroup ----
// The levels are small enough that one workgroup sweeps each of them in a few
// microseconds, so running smoothing, residual, restriction, prolongation and
// the BiCGStab bottom solve back to back with workgroup barriers replaces
// about 7 launches per level.  Same operations, same order, as the per-level
// kernels.
constexpr int kMaxCoarse = 8;

template <class T>
struct CoarseLevel {
    KGeom g;
    Coef<T> c;
    T omega;
    const T* sig;
    T* b;
    T* x;
    T* x2;
    T* r;
    const unsigned char* m;  // overset mask (global memory), null without one
};

template <class T>
struct CoarseArgs {
    CoarseLevel<T> L[kMaxCoarse];
    int nl;           // levels in this launch (the last one is the bottom)
    int nu1, nu2;
    Coef<double> cb;  // bottom operator coefficients (double)
    double* work;
    int bottom_kind;  // MGX_BOTTOM_*
    double bottom_omega;
    int maxiter;
    double rtol;
    int singular;
    int* iters_acc;
};

template <class T>
__device__ __forceinline__ void node_of(const KGeom& g, int id, int& i, int& j, int& k) {
    const int q = id / g.nn[0];
    i = id - q * g.nn[0];
    j = q % g.nn[1];
    k = q / g.nn[1];
}

// L x at node (i,j,k) with direct loads; also returns the sum of the 8 cell coefficients.
template <class T>
__device__ __forceinline__ T apply_node_T(const KGeom& g, const Coef<T>& c, const T* sig,
                                          const T* x, int i, int j, int k, T& ssum) {
    return apply_node_nbhd<T, T, T>(g, c, sig, x, i, j, k, ssum);
}

// Bytes of dynamic LDS the coarse kernel needs for levels L[0..nl-1]: every
// level's sigma, b, x, x2 and r (r shares x2's storage) and the bottom
// solver's 7 FP64 work vectors.
template <class T>
inline size_t coarse_lds_bytes(const KGeom* g, int nl) {
    size_t b = 0;
    for (int l = 0; l < nl; ++l) {
        const size_t nn = (size_t)g[l].nn[0] * g[l].nn[1] * g[l].nn[2];
        const size_t ns = (size_t)g[l].sg[0] * g[l].sg[1] * g[l].sg[2];
        b += ((ns + 3 * nn) * sizeof(T) + 15) / 16 * 16;
    }
    const size_t nb = (size_t)g[nl - 1].nn[0] * g[nl - 1].nn[1] * g[nl - 1].nn[2];
    return b + 7 * nb * sizeof(double);
}

// The whole coarse hierarchy lives in LDS for the duration of the launch:
// sigma and the top level's rhs are copied in, the top level's correction is
// copied out; everything in between never touches global memory.
template <class T>
static __global__ void __launch_bounds__(512) coarse_vcycle_kernel(const CoarseArgs<T> a) {
    __shared__ double red[16];
    __shared__ CoarseLevel<T> sl[kMaxCoarse];  // level descriptors with LDS pointers
    __shared__ double* swork;
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        size_t off = 0;
        for (int l = 0; l < a.nl; ++l) {
            CoarseLevel<T> lv = a.L[l];
            const int nn = lv.g.nn[0] * lv.g.nn[1] * lv.g.nn[2];
            const int ns = lv.g.sg[0] * lv.g.sg[1] * lv.g.sg[2];
            T* base = reinterpret_cast<T*>(lds + off);
            lv.sig = base;
            lv.b = base + ns;
            lv.x = base + ns + nn;
            lv.x2 = base + ns + 2 * nn;
            lv.r = lv.x2;  // the residual is formed after smoothing, when x2 is free
            sl[l] = lv;
            off += ((size_t)(ns + 3 * nn) * sizeof(T) + 15) / 16 * 16;
        }
        swork = reinterpret_cast<double*>(lds + off);
    }
    __syncthreads();
    for (int l = 0; l < a.nl; ++l) {
        const int ns = sl[l].g.sg[0] * sl[l].g.sg[1] * sl[l].g.sg[2];
        T* dst = const_cast<T*>(sl[l].sig);
        const T* src = a.L[l].sig;
        for (int id = tid; id < ns; id += nt) dst[id] = src[id];
    }
    {
        const int nn = sl[0].g.nn[0] * sl[0].g.nn[1] * sl[0].g.nn[2];
        for (int id = tid; id < nn; id += nt) sl[0].b[id] = a.L[0].b[id];
    }
    __syncthreads();
    T* top_x_global = a.L[0].x;
    unsigned swmask = 0u;  // bit l: level l's current iterate is in x2
    auto cur = [&](int l) { return ((swmask >> l) & 1u) ? sl[l].x2 : sl[l].x; };
    auto oth = [&](int l) { return ((swmask >> l) & 1u) ? sl[l].x : sl[l].x2; };
    auto jacobi = [&](int l) {
        const CoarseLevel<T>& lv = sl[l];
        const int n = lv.g.nn[0] * lv.g.nn[1] * lv.g.nn[2];
        const T* xin = cur(l);
        T* xout = oth(l);
        for (int id = tid; id < n; id += nt) {
            int i, j, k;
            node_of<T>(lv.g, id, i, j, k);
            T ssum;
            const T acc = apply_node_T<T>(lv.g, lv.c, lv.sig, xin, i, j, k, ssum);
            T v = xin[id];
            if (node_fixed(lv.g, lv.m, id, i, j, k)) v = 0;
            else if (ssum != (T)0) v = v + lv.omega * (lv.b[id] - acc) / (-lv.c.diag4 * ssum);
            xout[id] = v;
        }
        swmask ^= (1u << l);
        __syncthreads();
    };
    // down
    for (int l = 0; l < a.nl - 1; ++l) {
        const CoarseLevel<T>& lv = sl[l];
        const int n = lv.g.nn[0] * lv.g.nn[1] * lv.g.nn[2];
        if (a.nu1 > 0) {
            T* x = cur(l);
            for (int id = tid; id < n; id += nt) {
                int i, j, k;
                node_of<T>(lv.g, id, i, j, k);
                T ssum = 0;
                for (int q = 0; q < 8; ++q)
                    ssum += lv.sig[(i + (q & 1)) + (long)lv.g.sg[0] * ((j + ((q >> 1) & 1)) +
                                                                    (long)lv.g.sg[1] * (k + ((q >> 2) & 1)))];
                x[id] = (node_fixed(lv.g, lv.m, id, i, j, k) || ssum == (T)0)
                            ? (T)0
                            : lv.omega * lv.b[id] / (-lv.c.diag4 * ssum);
            }
            __syncthreads();
            for (int s = 1; s < a.nu1; ++s) jacobi(l);
        } else {
            T* x = cur(l);
            for (int id = tid; id < n; id += nt) x[id] = 0;
            __syncthreads();
        }
        {
            const T* x = cur(l);
            for (int id = tid; id < n; id += nt) {
                int i, j, k;
                node_of<T>(lv.g, id, i, j, k);
                T ssum;
                const T acc = apply_node_T<T>(lv.g, lv.c, lv.sig, x, i, j, k, ssum);
                oth(l)[id] = node_fixed(lv.g, lv.m, id, i, j, k) ? (T)0 : lv.b[id] - acc;
            }
            __syncthreads();
        }
        {
            const CoarseLevel<T>& cv = sl[l + 1];
            const T* rf = oth(l);  // the residual just formed
            const KGeom& gf = lv.g;
            const KGeom& gc = cv.g;
            const int nc = gc.nn[0] * gc.nn[1] * gc.nn[2];
            for (int id = tid; id < nc; id += nt) {
                int I, J, K;
                node_of<T>(gc, id, I, J, K);
                T acc = 0;
                if (!node_fixed(gc, cv.m, id, I, J, K)) {
                    for (int oz = -1; oz <= 1; ++oz) {
                        int fk = 2 * K + oz;
                        if (gf.per[2]) fk = fk < 0 ? fk + gf.nn[2] : (fk >= gf.nn[2] ? fk - gf.nn[2] : fk);
                        else if (fk < 0 || fk >= gf.nn[2]) continue;
                        for (int oy = -1; oy <= 1; ++oy) {
                            int fj = 2 * J + oy;
                            if (gf.per[1]) fj = fj < 0 ? fj + gf.nn[1] : (fj >= gf.nn[1] ? fj - gf.nn[1] : fj);
                            else if (fj < 0 || fj >= gf.nn[1]) continue;
                            for (int ox = -1; ox <= 1; ++ox) {
                                int fi = 2 * I + ox;
                                if (gf.per[0]) fi = fi < 0 ? fi + gf.nn[0] : (fi >= gf.nn[0] ? fi - gf.nn[0] : fi);
                                else if (fi < 0 || fi >= gf.nn[0]) continue;
                                const T w = (ox ? (T)0.25 : (T)0.5) * (oy ? (T)0.25 : (T)0.5) *
                                            (oz ? (T)0.25 : (T)0.5);
                                acc += w * rf[fi + (long)gf.nn[0] * (fj + (long)gf.nn[1] * fk)];
                            }
                        }
                    }
                }
                cv.b[id] = acc;
            }
            __syncthreads();
        }
    }
    // bottom
    {
        const int lb = a.nl - 1;
        const NodalBottomOp<T> op{sl[lb].g, a.cb, sl[lb].sig, sl[lb].m, a.bottom_omega};
        bottom_solve_body(op, sl[lb].b, cur(lb), swork, a.bottom_kind, a.maxiter, a.rtol,
                          a.singular, a.iters_acc, red);
        __syncthreads();
    }
    // up
    for (int l = a.nl - 2; l >= 0; --l) {
        const CoarseLevel<T>& lv = sl[l];
        const KGeom& gf = lv.g;
        const KGeom& gc = sl[l + 1].g;
        const T* xc = cur(l + 1);
        T* xf = cur(l);
        const int nf = gf.nn[0] * gf.nn[1] * gf.nn[2];
        for (int id = tid; id < nf; id += nt) {
            int i, j, k;
            node_of<T>(gf, id, i, j, k);
            if (node_fixed(gf, lv.m, id, i, j, k)) continue;
            const int f[3] = {i, j, k};
            int lo[3], hi[3];
            for (int d = 0; d < 3; ++d) {
                lo[d] = f[d] >> 1;
                hi[d] = (f[d] & 1) ? lo[d] + 1 : lo[d];
                if (gc.per[d] && hi[d] >= gc.nn[d]) hi[d] -= gc.nn[d];
            }
            T acc = 0;
            for (int q = 0; q < 8; ++q) {
                const int I = (q & 1) ? hi[0] : lo[0];
                const int J = (q & 2) ? hi[1] : lo[1];
                const int K = (q & 4) ? hi[2] : lo[2];
                acc += xc[I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K)];
            }
            xf[id] += (T)0.125 * acc;
        }
        __syncthreads();
        for (int s = 0; s < a.nu2; ++s) jacobi(l);
    }
    // the top level's correction goes back to global memory
    {
        const CoarseLevel<T>& lv = sl[0];
        const int n = lv.g.nn[0] * lv.g.nn[1] * lv.g.nn[2];
        const T* xs = cur(0);
        for (int id = tid; id < n; id += nt) top_x_global[id] = xs[id];
    }
}

}  // namespace mgx

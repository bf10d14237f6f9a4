// nodal_mg.hip -- host side of the nodal multigrid: hierarchy set-up, V-cycle
// and the outer MLMG iteration in FP64 or as an FP32 correction inside FP64
// iterative refinement.
//
// Algorithm (the same steps as the oracle, oracle/orc_nodal.c):
//   b <- b with Dirichlet rows zeroed; singular problems: b <- b - mean(b)
//   r <- b - L phi;  tol = max(atol, rtol ||b||_inf)
//   repeat: cor <- Vcycle(r); phi <- phi + cor; r <- b - L phi   (FP64)
//   until ||r||_inf <= tol  (incflo: project(rtol=1e-11, atol=1e-14),
//   src/projection/incflo_apply_nodal_projection.cpp:215, src/incflo.H:453-454)
// The V-cycle: weighted Jacobi (nu1 sweeps, the first from zero), residual,
// full-weighting restriction, recursion, trilinear prolongation, nu2 sweeps;
// the coarsest level is solved by a one-workgroup FP64 BiCGStab.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "eb_nodal.h"
#include "learned_kernels.h"
#include "mgx_ctx.h"
#include "nodal_mg.h"

namespace mgx {

// ---- launch shapes of the stencil kernel ------------------------------------
// A shape is the nodes per thread (NPT: 1 -> 64x8 threads, 2 -> 64x4, 4 -> 64x2
// threads per workgroup of 64 x 8 nodes), a target workgroup count that sets
// the z-run length, and the shortest z-run.  Defaults: NPT 2, 1024 workgroups
// for 4-byte storage and 512 for FP64 (measured best at 256^3 on MI355X), z-runs
// of at least 2 planes.  mgx_nodal_autotune times candidates on the handle's
// own levels and stores the fastest per (level shape, storage, pass) in this
// process-wide table (mgx_tuning_get / mgx_tuning_set read and write it).
struct StencilShape {
    int npt = 2;
    long target = 1024;
    int kzmin = 2;
    // Jacobi with 4-byte storage and npt 2: 0 LDS-staged registers, 1 LDS-direct
    // plane loads, 2 / 3 wave strips of 2 / 4 rows (cross-lane shuffles, no LDS)
    int dma = 0;
};
struct ShapeKey {
    int nn0, nn1, nn2, word, cword, mode, fine;
    bool operator<(const ShapeKey& o) const {
        const int a[7] = {nn0, nn1, nn2, word, cword, mode, fine};
        const int b[7] = {o.nn0, o.nn1, o.nn2, o.word, o.cword, o.mode, o.fine};
        return std::lexicographical_compare(a, a + 7, b, b + 7);
    }
};
static std::mutex g_tune_mtx;
static std::map<ShapeKey, StencilShape> g_shapes;
// levels with at most this many nodes run inside the single-workgroup launch
// that holds the bottom solve (1: only the bottom level; measured at 256^3:
// per-level launches for 16^3 and 8^3 beat folding them into one workgroup,
// which is then instruction-latency bound on one CU)
static long g_coarse_nodes = 1;

static StencilShape stencil_shape(const NodalGeom& g, int word, int cword, int mode, bool fine) {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    auto it = g_shapes.find(ShapeKey{g.nn[0], g.nn[1], g.nn[2], word, cword, mode, fine ? 1 : 0});
    if (it != g_shapes.end()) return it->second;
    StencilShape d;
    d.target = word == 8 ? 512 : 1024;
    return d;
}

static long coarse_threshold() {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    return g_coarse_nodes;
}

// two consecutive FP32 Jacobi sweeps of a level run as one fused pass
// (nodal_jacobi_pair_kernel) when set in the tuning table; off by default:
// at 256^3 the fused pass measured slower than two single sweeps (mixed
// projection 8.85 against 6.85 ms)
static int g_jacobi_pairs = 0;
// cell-centred V-cycles: levels of at most this many cells (below the first
// level) run in one workgroup (mac_coarse_cycle_kernel); 0: separate launches.
// Taylor-Green 64^3 step: 6.00 ms with 0, 5.64 with 512, 5.84 with 4096,
// 10.5 with 32768 (one workgroup is too slow for a 32^3 level)
static long g_mac_coarse_cells = 512;
// stencil tiles of launches over a TileSet: width (16 or 64 nodes) for FP64
// and for FP32 storage, planes per tile (z-run)
static int g_amr_tile_width = 16;
static int g_amr_tile_width32 = 64;
static int g_amr_tile_kz = TileSet::kKZ;

long mac_coarse_cells() {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    return g_mac_coarse_cells;
}

static bool jacobi_pairs_on() {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    return g_jacobi_pairs != 0;
}

template <class T>
static bool launch_jacobi_pair(const NodalGeom&, SArgs<T, T>, hipStream_t) {
    return false;
}

template <>
bool launch_jacobi_pair<float>(const NodalGeom& g, SArgs<float, float> a, hipStream_t s) {
    if (a.nmask || !jacobi_pairs_on()) return false;
    a.gx = (g.nn[0] + kPairTX - 1) / kPairTX;
    a.gy = (g.nn[1] + kPairTY - 1) / kPairTY;
    const long gxy = (long)a.gx * a.gy;
    int kz = (int)std::min<long>(256, std::max<long>(2, (g.nn[2] * gxy) / 1024));
    kz = std::min(kz, g.nn[2]);
    a.kz = kz;
    a.gz = (g.nn[2] + kz - 1) / kz;
    const unsigned nb = (unsigned)(gxy * a.gz);
    if (g.iso())
        nodal_jacobi_pair_kernel<float, true><<<nb, dim3(64, 4), 0, s>>>(a);
    else
        nodal_jacobi_pair_kernel<float, false><<<nb, dim3(64, 4), 0, s>>>(a);
    return true;
}

template <class T, class TC, int MODE, int NPT>
static void launch_stencil_npt(const NodalGeom& g, SArgs<T, TC> a, hipStream_t s, bool fine,
                               const StencilShape& sh) {
    constexpr int TX = 64, NY = (NPT == 1 ? 8 : 4 * NPT);
    a.gx = (g.nn[0] + TX - 1) / TX;
    a.gy = (g.nn[1] + NY - 1) / NY;
    const long gxy = (long)a.gx * a.gy;
    // z-planes per workgroup from the target workgroup count: long z-runs
    // amortise the two halo planes and the pipeline start on big levels;
    // small levels are latency bound and want short runs / more workgroups
    int kz = (int)std::min<long>(256, std::max<long>(sh.kzmin, (g.nn[2] * gxy) / sh.target));
    kz = std::min(kz, g.nn[2]);
    a.kz = kz;
    a.gz = (g.nn[2] + kz - 1) / kz;
    const unsigned nb = (unsigned)(gxy * a.gz);
    dim3 blk(TX, NY / NPT);
    if constexpr (MODE == KS_JACOBI && NPT == 2 && std::is_same<T, float>::value &&
                  std::is_same<TC, float>::value) {
        if ((sh.dma == 2 || sh.dma == 3) && !a.nmask) {
            // wave strips of 64 columns x NR rows, four per workgroup
            const int nr = sh.dma == 2 ? 2 : 4;
            SArgs<T, TC> b = a;
            b.gx = (g.nn[0] + kStripTX - 1) / kStripTX;
            b.gy = (g.nn[1] + 4 * nr - 1) / (4 * nr);
            const long sxy_t = (long)b.gx * b.gy;
            int kzs = (int)std::min<long>(256, std::max<long>(sh.kzmin, (g.nn[2] * sxy_t) / sh.target));
            kzs = std::min(kzs, g.nn[2]);
            b.kz = kzs;
            b.gz = (g.nn[2] + kzs - 1) / kzs;
            const unsigned nbs = (unsigned)(sxy_t * b.gz);
            if (nr == 2) {
                if (g.iso()) nodal_jacobi_strip_kernel<T, true, 2><<<nbs, blk, 0, s>>>(b);
                else nodal_jacobi_strip_kernel<T, false, 2><<<nbs, blk, 0, s>>>(b);
            } else {
                if (g.iso()) nodal_jacobi_strip_kernel<T, true, 4><<<nbs, blk, 0, s>>>(b);
                else nodal_jacobi_strip_kernel<T, false, 4><<<nbs, blk, 0, s>>>(b);
            }
            return;
        }
        if (sh.dma == 1 && !a.nmask) {
            if (g.iso()) {
                if (fine)
                    nodal_jacobi_dma_kernel<T, true, true, 4><<<nb, blk, 0, s>>>(a);
                else
                    nodal_jacobi_dma_kernel<T, true, false, 4><<<nb, blk, 0, s>>>(a);
            } else {
                nodal_jacobi_dma_kernel<T, false, false, 4><<<nb, blk, 0, s>>>(a);
            }
            return;
        }
    }
    if (a.nmask) {
        // composite-level sweeps (node types); kept out of the plain kernels
        if (g.iso())
            nodal_stencil_kernel<T, TC, MODE, true, false, NPT, true><<<nb, blk, 0, s>>>(a);
        else
            nodal_stencil_kernel<T, TC, MODE, false, false, NPT, true><<<nb, blk, 0, s>>>(a);
        return;
    }
    if (g.iso()) {
        if (fine)
            nodal_stencil_kernel<T, TC, MODE, true, true, NPT><<<nb, blk, 0, s>>>(a);
        else
            nodal_stencil_kernel<T, TC, MODE, true, false, NPT><<<nb, blk, 0, s>>>(a);
    } else {
        nodal_stencil_kernel<T, TC, MODE, false, false, NPT><<<nb, blk, 0, s>>>(a);
    }
}

template <class T, class TC, int MODE>
static void launch_stencil_shape(const NodalGeom& g, SArgs<T, TC> a, hipStream_t s, bool fine,
                                 const StencilShape& sh) {
    if (sh.npt == 1)
        launch_stencil_npt<T, TC, MODE, 1>(g, a, s, fine, sh);
    else if (sh.npt == 4)
        launch_stencil_npt<T, TC, MODE, 4>(g, a, s, fine, sh);
    else
        launch_stencil_npt<T, TC, MODE, 2>(g, a, s, fine, sh);
}

// a pass over the active tiles of a partly used level: 16 x 16-node tiles of
// 8 planes, one workgroup each
template <class T, class TC, int MODE>
static void launch_stencil_tiles(const NodalGeom& g, SArgs<T, TC> a, hipStream_t s, TileSet* ts) {
    int width, kz;
    {
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        width = sizeof(T) == 4 ? g_amr_tile_width32 : g_amr_tile_width;
        kz = g_amr_tile_kz;
    }
    a.kz = std::min(kz, g.nn[2]);
    a.gz = (g.nn[2] + a.kz - 1) / a.kz;
    int cnt = 0;
    if (width == 64) {
        // the dense kernel's 64 x 8 tiles (2 nodes per thread)
        a.gx = (g.nn[0] + 63) / 64;
        a.gy = (g.nn[1] + 7) / 8;
        a.tiles = ts->tiles(64, 8, a.kz, a.gx, a.gy, a.gz, cnt);
        if (cnt == 0) return;
        const dim3 blk(64, 4);
        if (a.nmask) {
            if (g.iso()) nodal_stencil_kernel<T, TC, MODE, true, false, 2, true><<<cnt, blk, 0, s>>>(a);
            else nodal_stencil_kernel<T, TC, MODE, false, false, 2, true><<<cnt, blk, 0, s>>>(a);
        } else {
            if (g.iso()) nodal_stencil_kernel<T, TC, MODE, true, false, 2><<<cnt, blk, 0, s>>>(a);
            else nodal_stencil_kernel<T, TC, MODE, false, false, 2><<<cnt, blk, 0, s>>>(a);
        }
        return;
    }
    constexpr int TX = TileSet::kTX, TY = TileSet::kTY;
    a.gx = (g.nn[0] + TX - 1) / TX;
    a.gy = (g.nn[1] + TY - 1) / TY;
    a.tiles = ts->tiles(TX, TY, a.kz, a.gx, a.gy, a.gz, cnt);
    if (cnt == 0) return;
    const dim3 blk(TX, TY);
    if (a.nmask) {
        if (g.iso())
            nodal_stencil_kernel<T, TC, MODE, true, false, 1, true, TX><<<cnt, blk, 0, s>>>(a);
        else
            nodal_stencil_kernel<T, TC, MODE, false, false, 1, true, TX><<<cnt, blk, 0, s>>>(a);
    } else {
        if (g.iso())
            nodal_stencil_kernel<T, TC, MODE, true, false, 1, false, TX><<<cnt, blk, 0, s>>>(a);
        else
            nodal_stencil_kernel<T, TC, MODE, false, false, 1, false, TX><<<cnt, blk, 0, s>>>(a);
    }
}

template <class T, class TC, int MODE>
void launch_stencil(const NodalGeom& g, SArgs<T, TC> a, hipStream_t s, bool fine, TileSet* ts) {
    if (ts) {
        launch_stencil_tiles<T, TC, MODE>(g, a, s, ts);
        return;
    }
    launch_stencil_shape<T, TC, MODE>(
        g, a, s, fine, stencil_shape(g, (int)sizeof(T), (int)sizeof(TC), MODE, fine));
}

void TileSet::build(const NodalGeom& g, const std::vector<unsigned char>& active) { build(g.nn, active); }

void TileSet::build(const int dims[3], const std::vector<unsigned char>& active) {
    MGX_REQUIRE((long)active.size() == (long)dims[0] * dims[1] * dims[2], MGX_ERR_ARG,
                "tile set: one flag per point");
    for (int d = 0; d < 3; ++d) nn_[d] = dims[d];
    g16_ = (nn_[0] + 15) / 16;
    col_.assign((size_t)g16_ * nn_[1] * nn_[2], 0);
    for (int k = 0; k < nn_[2]; ++k)
        for (int j = 0; j < nn_[1]; ++j)
            for (int i = 0; i < nn_[0]; ++i)
                if (active[i + (size_t)nn_[0] * (j + (size_t)nn_[1] * k)])
                    col_[i / 16 + (size_t)g16_ * (j + (size_t)nn_[1] * k)] = 1;
    lists_.clear();
    nblk_ = -1;
}

const int* TileSet::tiles(int tx, int ty, int kz, int gx, int gy, int gz, int& count) {
    MGX_REQUIRE(built() && tx % 16 == 0 && gx == (nn_[0] + tx - 1) / tx &&
                    gy == (nn_[1] + ty - 1) / ty,
                MGX_ERR_STATE, "tile set: not built for this level");
    auto it = lists_.find({tx, ty, kz});
    if (it == lists_.end()) {
        const int c16 = tx / 16;
        std::vector<int> t;
        for (int bz = 0; bz < gz; ++bz)
            for (int by = 0; by < gy; ++by)
                for (int bx = 0; bx < gx; ++bx) {
                    bool any = false;
                    for (int k = bz * kz; k < std::min((bz + 1) * kz, nn_[2]) && !any; ++k)
                        for (int j = by * ty; j < std::min((by + 1) * ty, nn_[1]) && !any; ++j)
                            for (int c = bx * c16; c < std::min((bx + 1) * c16, g16_) && !any; ++c)
                                any = col_[c + (size_t)g16_ * (j + (size_t)nn_[1] * k)] != 0;
                    if (any) t.push_back(bx + gx * (by + gy * bz));
                }
        DevBuf buf;
        if (!t.empty()) {
            buf.alloc(t.size() * sizeof(int));
            MGX_HIP(hipMemcpy(buf.p, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        it = lists_.emplace(std::array<int, 3>{tx, ty, kz},
                            std::make_pair(std::move(buf), (int)t.size()))
                 .first;
    }
    count = it->second.second;
    return it->second.first.as<int>();
}

const int4* TileSet::blocks(int& count) {
    MGX_REQUIRE(built(), MGX_ERR_STATE, "tile set: not built");
    if (nblk_ < 0) {
        std::vector<int4> t;
        const int gy = (nn_[1] + kTY - 1) / kTY;
        for (int k = 0; k < nn_[2]; ++k)
            for (int by = 0; by < gy; ++by)
                for (int bx = 0; bx < g16_; ++bx) {
                    bool any = false;
                    for (int j = by * kTY; j < std::min((by + 1) * kTY, nn_[1]) && !any; ++j)
                        any = col_[bx + (size_t)g16_ * (j + (size_t)nn_[1] * k)] != 0;
                    if (any) t.push_back(make_int4(bx, by, k, 0));
                }
        nblk_ = (int)t.size();
        if (!t.empty()) {
            blk_.alloc(t.size() * sizeof(int4));
            MGX_HIP(hipMemcpy(blk_.p, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
        }
    }
    count = nblk_;
    return blk_.as<int4>();
}

// Full-weighting restriction onto coarse level gc (z-streamed, LDS-staged);
// xc1 (optional) receives the coarse first Jacobi sweep from zero.
template <class T, class TF = T>
static void launch_restrict(const KGeom& kf, const NodalGeom& gc, const KGeom& kc, const TF* rf, T* rc,
                            const T* sigc, T diag4c, T omegac, T* xc1, hipStream_t s) {
    const int gx = (gc.nn[0] + 63) / 64, gy = (gc.nn[1] + 3) / 4;
    const long gxy = (long)gx * gy;
    int kz = (int)std::min<long>(64, std::max<long>(1, (gc.nn[2] * gxy) / 2048));
    kz = std::min(kz, gc.nn[2]);
    const int gz = (gc.nn[2] + kz - 1) / kz;
    nodal_restrict_tiled_kernel<T, TF><<<(unsigned)(gxy * gz), dim3(64, 4), 0, s>>>(
        kf, kc, rf, rc, sigc, diag4c, omegac, xc1, kz, gx, gy, gz);
}

template <class T>
void restrict_full(const NodalGeom& gf, const NodalGeom& gc, const T* rf, T* rc, hipStream_t s) {
    launch_restrict<T, T>(kgeom(gf), gc, kgeom(gc), rf, rc, nullptr, (T)0, (T)0, nullptr, s);
}
template void restrict_full<float>(const NodalGeom&, const NodalGeom&, const float*, float*, hipStream_t);
template void restrict_full<double>(const NodalGeom&, const NodalGeom&, const double*, double*, hipStream_t);

void LaunchTimer::begin(hipStream_t s) {
    if (!on) return;
    while (ev.size() < used + 2) {
        hipEvent_t e;
        MGX_HIP(hipEventCreate(&e));
        ev.push_back(e);
    }
    MGX_HIP(hipEventRecord(ev[used], s));
}

void LaunchTimer::end(hipStream_t s) {
    if (!on) return;
    MGX_HIP(hipEventRecord(ev[used + 1], s));
    used += 2;
}

void LaunchTimer::reserve(size_t pairs) {
    while (ev.size() < used + 2 * pairs) {
        hipEvent_t e;
        MGX_HIP(hipEventCreate(&e));
        ev.push_back(e);
    }
}

void LaunchTimer::collect() {
    for (size_t q = 0; q + 1 < used; q += 2) {
        float ms = 0.f;
        MGX_HIP(hipEventSynchronize(ev[q + 1]));
        MGX_HIP(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
        total_ms += ms;
        ++count;
    }
    used = 0;
}

LaunchTimer::~LaunchTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
}

template <class T>
void NodalHierarchy<T>::drop_graphs() {
    for (auto& c : cache_)
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
    cache_.clear();
}

template <class T>
void NodalHierarchy<T>::build(const NodalGeom& g0, int nlev) {
    drop_graphs();
    static bool attr_set = false;
    if (!attr_set) {
        // the coarse-level kernel may use up to 150 KiB of LDS (advisory on
        // AMD GPUs, which grant up to the CU's LDS without it); set outside
        // any stream capture
        (void)hipFuncSetAttribute((const void*)coarse_vcycle_kernel<T>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        attr_set = true;
    }
    L.clear();
    L.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        MGLevel<T>& lv = L[l];
        lv.g = (l == 0) ? g0 : coarsen(L[l - 1].g);
        lv.kg = kgeom(lv.g);
        lv.coef = make_coef<T>(lv.g.dx);
        lv.iso = lv.g.iso();
        const long nn = lv.g.nnode(), ns = lv.g.nsig();
        lv.sig_buf.alloc(ns * sizeof(T));
        MGX_HIP(hipMemset(lv.sig_buf.p, 0, ns * sizeof(T)));
        if (l > 0) lv.b_buf.alloc(nn * sizeof(T));
        lv.x_buf.alloc(nn * sizeof(T));
        lv.x2_buf.alloc(nn * sizeof(T));
        if (l < nlev - 1) lv.r_buf.alloc(nn * sizeof(T));
        lv.sig = lv.sig_buf.template as<T>();
        lv.b = lv.b_buf.template as<T>();
        lv.x = lv.x_buf.template as<T>();
        lv.x2 = lv.x2_buf.template as<T>();
        lv.r = lv.r_buf.template as<T>();
    }
    bottom_work.alloc(7 * L.back().g.nnode() * sizeof(double));
    built = true;
}

template <class T>
void NodalHierarchy<T>::set_sigma(const double* sig0, hipStream_t s) {
    const long ns = L[0].g.nsig();
    if (std::is_same<T, double>::value) {
        if ((const void*)sig0 != (const void*)L[0].sig)
            MGX_HIP(hipMemcpyAsync(L[0].sig, sig0, ns * sizeof(double), hipMemcpyDeviceToDevice, s));
    } else {
        cast_kernel<T, double><<<grid1d(ns), 256, 0, s>>>(sig0, L[0].sig, ns);
    }
    for (size_t l = 1; l < L.size(); ++l) {
        const long nc = (long)L[l].g.n[0] * L[l].g.n[1] * L[l].g.n[2];
        sigma_coarsen_kernel<T><<<grid3(L[l].g.n[0], L[l].g.n[1], L[l].g.n[2]), block3(), 0, s>>>(L[l - 1].kg, L[l].kg, L[l - 1].sig,
                                                           L[l].sig);
    }
}

template <class T>
void NodalHierarchy<T>::smooth(int l, int nsweeps, bool from_zero, const MGOpts& o, hipStream_t s,
                               bool x_ready) {
    MGLevel<T>& lv = L[l];
    const long nn = lv.g.nnode();
    int sweeps = nsweeps;
    int q0 = 0;  // index of the next sweep among the nsweeps
    if (from_zero && x_ready && sweeps > 0) {
        --sweeps;
        q0 = 1;
    } else if (from_zero) {
        if (sweeps == 0) {
            zero_kernel<T><<<grid1d(nn), 256, 0, s>>>(lv.x, nn);
            return;
        }
        SArgs<T, T> a{};
        a.g = lv.kg;
        a.c = lv.coef;
        a.b = lv.b;
        a.sig = lv.sig;
        a.xout = lv.x;
        a.omega = (T)sweep_omega(lv.g, o, cheb, 0, nsweeps);
        a.nmask = mask(l);
        launch_stencil<T, T, KS_JFIRST>(lv.g, a, s, false);
        --sweeps;
        q0 = 1;
    }
    for (int q = 0; q < sweeps;) {
        SArgs<T, T> a{};
        a.g = lv.kg;
        a.c = lv.coef;
        a.x = lv.x;
        a.b = lv.b;
        a.sig = lv.sig;
        a.xout = lv.x2;
        a.omega = (T)sweep_omega(lv.g, o, cheb, q0 + q, nsweeps);
        a.nmask = mask(l);
        // the fused pair applies one factor to both sweeps: Jacobi only
        if (!cheb && sweeps - q >= 2 && launch_jacobi_pair<T>(lv.g, a, s)) {
            q += 2;
        } else {
            if (l == 0 && timer) timer->begin(s);
            launch_stencil<T, T, KS_JACOBI>(lv.g, a, s, l == 0);
            if (l == 0 && timer) timer->end(s);
            ++q;
        }
        std::swap(lv.x, lv.x2);
    }
}

static bool same_opts(const MGOpts& a, const MGOpts& b) {
    return a.nu1 == b.nu1 && a.nu2 == b.nu2 && a.omega == b.omega &&
           a.bottom_maxiter == b.bottom_maxiter && a.bottom_rtol == b.bottom_rtol &&
           a.bottom_solver == b.bottom_solver && a.smoother == b.smoother;
}

template <class T>
void NodalHierarchy<T>::vcycle(const MGOpts& o, hipStream_t s, int* bottom_iters) {
    cheb = use_chebyshev(o, sizeof(T) == 4);
    // Levels below 0 are recomputed from scratch every cycle: start each from
    // its own buffer pair, so every cycle launches the same sequence.
    for (size_t l = 1; l < L.size(); ++l) {
        L[l].x = L[l].x_buf.template as<T>();
        L[l].x2 = L[l].x2_buf.template as<T>();
    }
    if (!x0_ready && !x0_guess) {
        L[0].x = L[0].x_buf.template as<T>();
        L[0].x2 = L[0].x2_buf.template as<T>();
    }
    // Timed cycles (HIP events around the finest-level sweeps) run with plain
    // launches: events recorded inside a graph cannot be timed.
    const bool timed = timer && timer->on;
    if (!graphs || timed) {
        vcycle_body(o, s, bottom_iters);
        return;
    }
    const bool ready0 = x0_ready && o.nu1 > 0;
    const bool guess0 = x0_guess;
    CycleGraph* hit = nullptr;
    for (auto& c : cache_)
        if (c.x0 == L[0].x && c.b0 == L[0].b && c.ready0 == ready0 && c.guess0 == guess0 &&
            c.timed == timed &&
            c.bottom_iters == bottom_iters && same_opts(c.o, o))
            hit = &c;
    if (!hit) {
        CycleGraph c;
        c.x0 = L[0].x;
        c.b0 = L[0].b;
        c.ready0 = ready0;
        c.guess0 = guess0;
        c.timed = timed;
        c.bottom_iters = bottom_iters;
        c.o = o;
        const bool saved_ready = x0_ready;
        const bool saved_guess = x0_guess;
        if (!cap_stream_ && hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            cap_stream_ = nullptr;
            graphs = false;
            vcycle_body(o, s, bottom_iters);
            return;
        }
        hipStream_t cs = cap_stream_;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            graphs = false;
            vcycle_body(o, s, bottom_iters);
            return;
        }
        bool body_ok = true;
        try {
            vcycle_body(o, cs, bottom_iters);
        } catch (...) {
            body_ok = false;  // e.g. a call the runtime does not capture
        }
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &g);
        if (e == hipSuccess && body_ok) e = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess || !body_ok) {
            // capture not available here: run the cycle with plain launches
            if (o.verbose)
                fprintf(stderr, "mgx: V-cycle graph capture failed (%s); plain launches\n",
                        hipGetErrorString(e));
            (void)hipGetLastError();
            graphs = false;
            x0_ready = saved_ready;
            x0_guess = saved_guess;
            for (size_t l = 1; l < L.size(); ++l) {
                L[l].x = L[l].x_buf.template as<T>();
                L[l].x2 = L[l].x2_buf.template as<T>();
            }
            L[0].x = c.x0;
            L[0].x2 = (c.x0 == L[0].x_buf.template as<T>()) ? L[0].x2_buf.template as<T>()
                                                            : L[0].x_buf.template as<T>();
            vcycle_body(o, s, bottom_iters);
            return;
        }
        for (auto& lv : L) {
            c.xs.push_back(lv.x);
            c.x2s.push_back(lv.x2);
        }
        cache_.push_back(std::move(c));
        hit = &cache_.back();
    }
    MGX_HIP(hipGraphLaunch(hit->exec, s));
    for (size_t l = 0; l < L.size(); ++l) {
        L[l].x = hit->xs[l];
        L[l].x2 = hit->x2s[l];
    }
    x0_ready = false;
    x0_guess = false;
}

template <class T>
void NodalHierarchy<T>::vcycle_body(const MGOpts& o, hipStream_t s, int* bottom_iters) {
    const int nl = (int)L.size();
    // levels lc..nl-1 run in the LDS-resident coarse kernel (when they fit)
    constexpr size_t kLdsBudget = 150 * 1024;
    auto lds_for = [&](int first) {
        KGeom gs[kMaxCoarse];
        for (int q = 0; q < nl - first; ++q) gs[q] = L[first + q].kg;
        return coarse_lds_bytes<T>(gs, nl - first);
    };
    int lc = nl - 1;
    const bool bottom_in_lds = lds_for(nl - 1) <= kLdsBudget;
    if (bottom_in_lds)
        while (lc > 0 && nl - (lc - 1) <= kMaxCoarse && L[lc - 1].g.nnode() <= coarse_threshold() &&
               lds_for(lc - 1) <= kLdsBudget)
            --lc;
    bool ready = x0_ready && o.nu1 > 0;
    const bool guess = x0_guess;
    x0_ready = false;
    x0_guess = false;
    for (int l = 0; l < lc; ++l) {
        MGLevel<T>& lv = L[l];
        // level 0 with an initial guess pre-smooths from it
        smooth(l, o.nu1, !(l == 0 && guess), o, s, ready);
        SArgs<T, T> a{};
        a.g = lv.kg;
        a.c = lv.coef;
        a.x = lv.x;
        a.b = lv.b;
        a.sig = lv.sig;
        a.rout = lv.r;
        a.nmask = mask(l);
        launch_stencil<T, T, KS_RESID>(lv.g, a, s, false);
        // a coarse level smoothed by its own launches gets its first sweep
        // from zero out of the restriction (not with an overset mask: the
        // fused sweep does not know it)
        MGLevel<T>& lc1 = L[l + 1];
        ready = (l + 1 < lc) && o.nu1 > 0 && masks.empty();
        launch_restrict<T>(lv.kg, lc1.g, lc1.kg, lv.r, lc1.b, lc1.sig, lc1.coef.diag4,
                           (T)sweep_omega(lc1.g, o, cheb, 0, o.nu1), ready ? lc1.x : nullptr, s);
    }
    if (!bottom_in_lds) {
        MGLevel<T>& lb = L[nl - 1];
        const long nb = lb.g.nnode();
        const int threads = nb >= 1024 ? 1024 : (nb >= 256 ? 256 : 64);
        bottom_solve_kernel<T><<<1, threads, 0, s>>>(
            lb.kg, make_coef<double>(lb.g.dx), lb.sig, mask(nl - 1), lb.b, lb.x,
            bottom_work.as<double>(), o.bottom_solver, o.bottom_maxiter, o.bottom_rtol,
            nodal_omega(lb.g, o.omega), bottom_singular() ? 1 : 0, bottom_iters);
    } else {
        CoarseArgs<T> ca{};
        ca.nl = nl - lc;
        long nmax = 0;
        for (int q = 0; q < ca.nl; ++q) {
            MGLevel<T>& lv = L[lc + q];
            ca.L[q].g = lv.kg;
            ca.L[q].c = lv.coef;
            ca.L[q].omega = (T)nodal_omega(lv.g, o.omega);
            ca.L[q].sig = lv.sig;
            ca.L[q].b = lv.b;
            ca.L[q].x = lv.x;
            ca.L[q].x2 = lv.x2;
            ca.L[q].r = lv.r;
            ca.L[q].m = mask(lc + q);
            nmax = std::max(nmax, lv.g.nnode());
        }
        const MGLevel<T>& lb = L[nl - 1];
        ca.nu1 = o.nu1;
        ca.nu2 = o.nu2;
        ca.cb = make_coef<double>(lb.g.dx);
        ca.work = bottom_work.as<double>();
        ca.bottom_kind = o.bottom_solver;
        ca.bottom_omega = nodal_omega(lb.g, o.omega);
        ca.maxiter = o.bottom_maxiter;
        ca.rtol = o.bottom_rtol;
        ca.singular = bottom_singular() ? 1 : 0;
        ca.iters_acc = bottom_iters;
        const int threads = nmax >= 512 ? 512 : (nmax >= 256 ? 256 : 64);
        coarse_vcycle_kernel<T><<<1, threads, lds_for(lc), s>>>(ca);
    }
    for (int l = lc - 1; l >= 0; --l) {
        MGLevel<T>& lv = L[l];
        const long nf = lv.g.nnode();
        const NodalGeom& gcl = L[l + 1].g;
        nodal_prolong_add_coarse_kernel<T><<<grid3(gcl.nn[0], gcl.nn[1], gcl.nn[2]), block3(), 0, s>>>(
            lv.kg, L[l + 1].kg, L[l + 1].x, lv.x);
        smooth(l, o.nu2, false, o, s);
    }
}

template void launch_stencil<double, double, KS_APPLY>(const NodalGeom&, SArgs<double, double>, hipStream_t, bool, TileSet*);
template void launch_stencil<double, double, KS_RESID>(const NodalGeom&, SArgs<double, double>, hipStream_t, bool, TileSet*);
template void launch_stencil<double, double, KS_JACOBI>(const NodalGeom&, SArgs<double, double>, hipStream_t, bool, TileSet*);
template void launch_stencil<double, double, KS_JFIRST>(const NodalGeom&, SArgs<double, double>, hipStream_t, bool, TileSet*);
template void launch_stencil<float, float, KS_APPLY>(const NodalGeom&, SArgs<float, float>, hipStream_t, bool, TileSet*);
template void launch_stencil<float, float, KS_JACOBI>(const NodalGeom&, SArgs<float, float>, hipStream_t, bool, TileSet*);
template void launch_stencil<float, float, KS_JFIRST>(const NodalGeom&, SArgs<float, float>, hipStream_t, bool, TileSet*);

template class NodalHierarchy<double>;
template class NodalHierarchy<float>;

// ---------------------------------------------------------------------------

NodalSolver::~NodalSolver() {
    if (h_pinned) (void)hipHostFree(h_pinned);
}

void NodalSolver::setup(const NodalGeom& g, const MGOpts& o) {
    // sigma (cells with one ghost layer: periodic images or 0) does not depend
    // on the boundary types or the multigrid options, so a re-setup for new
    // ones (e.g. setDomainBC after the coefficients, as incflo calls them)
    // keeps it
    const bool keep_sigma = sigma_set && g.n[0] == g0.n[0] && g.n[1] == g0.n[1] &&
                            g.n[2] == g0.n[2] && g.per[0] == g0.per[0] && g.per[1] == g0.per[1] &&
                            g.per[2] == g0.per[2];
    g0 = g;
    opts = o;
    nlev = nodal_nlevels(g, o.max_coarsening_level);
    const long nn = g.nnode();
    sig0_buf.alloc(g.nsig() * sizeof(double));
    if (!keep_sigma) MGX_HIP(hipMemset(sig0_buf.p, 0, g.nsig() * sizeof(double)));
    phi_a.alloc(nn * sizeof(double));
    phi_b.alloc(nn * sizeof(double));
    MGX_HIP(hipMemset(phi_a.p, 0, nn * sizeof(double)));
    phi = phi_a.as<double>();
    scal.alloc(64 * sizeof(double));
    part.alloc(4096 * sizeof(double));
    ibuf.alloc(16 * sizeof(int));
    if (!h_pinned) MGX_HIP(hipHostMalloc((void**)&h_pinned, 64 * sizeof(unsigned long long), 0));
    H64.built = false;
    H16.built = false;
    mask_on = mask_fixed = false;  // the handle re-applies its mask after a set-up
    mask_lv.clear();
    H32.built = false;
    sigma_set = keep_sigma;
    h32_sigma = false;
    logsig_lvl = -1;
}

void NodalSolver::zero_norm(int slot, hipStream_t s) {
    zero_u64_kernel<<<1, 64, 0, s>>>(scal.as<unsigned long long>() + slot, 1);
}

double NodalSolver::read_norm(int slot, hipStream_t s) {
    MGX_HIP(hipMemcpyAsync(h_pinned + slot, scal.as<unsigned long long>() + slot,
                           sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    if (fine_timer.on) fine_timer.collect();
    unsigned long long v = h_pinned[slot];
    double d;
    std::memcpy(&d, &v, sizeof(d));
    return d;
}

double NodalSolver::sum(const double* a, long n, hipStream_t s) {
    const int np = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
    partial_sum_kernel<double><<<np, 256, 0, s>>>(a, n, part.as<double>());
    final_sum_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, scal.as<double>() + 32, 1.0);
    double h;
    MGX_HIP(hipMemcpyAsync(&h, scal.as<double>() + 32, sizeof(double), hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    return h;
}

static void remove_mean(NodalSolver& S, double* a, long n, hipStream_t s) {
    const int np = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
    partial_sum_kernel<double><<<np, 256, 0, s>>>(a, n, S.part.as<double>());
    final_sum_kernel<<<1, 256, 0, s>>>(S.part.as<double>(), np, S.scal.as<double>() + 33,
                                       1.0 / (double)n);
    subtract_scalar_kernel<double><<<grid1d(n), 256, 0, s>>>(a, n, S.scal.as<double>() + 33);
}

void NodalSolver::prepare_h32(hipStream_t s) {
    if (!H32.built || (int)H32.L.size() != nlev) { H32.build(g0, nlev); h32_sigma = false; }
    if (!h32_sigma) { H32.set_sigma(sigma0(), s); h32_sigma = true; }
    res32.alloc(g0.nnode() * sizeof(float));
    H32.L[0].b = res32.as<float>();
}

// out = (TO)(in * scale)
template <class TO>
static __global__ void scale_cast_kernel(const double* __restrict__ in, TO* __restrict__ out, long n,
                                         double scale) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] = st<TO>((acc_t<TO>)(in[id] * scale));
}

// coarse overset mask by injection: node I takes the fine node 2I
static __global__ void mask_inject_kernel(KGeom gf, KGeom gc, const unsigned char* __restrict__ mf,
                                          unsigned char* __restrict__ mc) {
    int I, J, K;
    if (!grid_point(gc.nn[0], gc.nn[1], I, J, K)) return;
    mc[I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K)] =
        mf[2 * I + (long)gf.nn[0] * (2 * J + (long)gf.nn[1] * 2 * K)];
}

// b = 0 where the mask is 0
static __global__ void mask_zero_kernel(const unsigned char* __restrict__ m, double* __restrict__ b, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n && !m[id]) b[id] = 0.0;
}

void NodalSolver::set_mask(const unsigned char* m0, hipStream_t s) {
    // captured V-cycles hold the old mask pointers
    H64.drop_graphs();
    H32.drop_graphs();
    H64.masks.clear();
    H32.masks.clear();
    mask_on = false;
    mask_fixed = false;
    mask_lv.clear();
    if (!m0) return;
    mask_lv.resize(nlev);
    NodalGeom gl = g0, gprev = g0;
    for (int l = 0; l < nlev; ++l) {
        mask_lv[l].alloc(gl.nnode());
        if (l == 0) {
            MGX_HIP(hipMemcpyAsync(mask_lv[0].p, m0, gl.nnode(), hipMemcpyDeviceToDevice, s));
        } else {
            mask_inject_kernel<<<grid3(gl.nn[0], gl.nn[1], gl.nn[2]), block3(), 0, s>>>(
                kgeom(gprev), kgeom(gl), mask_lv[l - 1].as<unsigned char>(), mask_lv[l].as<unsigned char>());
        }
        gprev = gl;
        if (l + 1 < nlev) gl = coarsen(gl);
    }
    std::vector<unsigned char> h(g0.nnode());
    MGX_HIP(hipMemcpyAsync(h.data(), m0, g0.nnode(), hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    for (unsigned char v : h) if (!v) { mask_fixed = true; break; }
    // the bottom problem is singular unless its own (injected) mask fixes a node
    const long nb = gl.nnode();  // gl is the bottom level here
    std::vector<unsigned char> hb(nb);
    MGX_HIP(hipMemcpyAsync(hb.data(), mask_lv[nlev - 1].p, nb, hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    mask_bottom_fixed = false;
    for (unsigned char v : hb) if (!v) { mask_bottom_fixed = true; break; }
    mask_on = true;
}

void NodalSolver::prepare_h16(hipStream_t s) {
    HalfTop& T = H16;
    if (!T.built) {
        T.dx0 = g0.dx[0];
        T.g = g0;
        for (int d = 0; d < 3; ++d) T.g.dx[d] = g0.dx[d] / T.dx0;
        T.kg = kgeom(T.g);
        T.coef = make_coef<float>(T.g.dx);
        const long nn = g0.nnode();
        T.sig.alloc(g0.nsig() * sizeof(__half));
        for (DevBuf* b : {&T.xa, &T.xb, &T.b, &T.r}) b->alloc(nn * sizeof(__half));
        T.x = T.xa.as<__half>();
        T.x2 = T.xb.as<__half>();
        const NodalGeom gs = coarsen(T.g);
        T.sub.build(gs, nlev - 1);
        T.sub.timer = nullptr;
        T.sub_b.alloc(gs.nnode() * sizeof(float));
        T.sub.L[0].b = T.sub_b.as<float>();
        T.sub_sig.alloc(gs.nsig() * sizeof(double));
        T.built = true;
        T.sigma_ok = false;
        T.ready = false;
    }
    if (!T.sigma_ok) {
        zero_norm(2, s);
        max_abs_kernel<double><<<256, 256, 0, s>>>(sigma0(), g0.nsig(), scal.as<unsigned long long>() + 2);
        const double smax = read_norm(2, s);
        MGX_REQUIRE(smax > 0.0, MGX_ERR_STATE, "nodal solve: sigma is zero everywhere");
        T.sigmax = smax;
        scale_cast_kernel<__half><<<grid1d(g0.nsig()), 256, 0, s>>>(sigma0(), T.sig.as<__half>(),
                                                                    g0.nsig(), 1.0 / smax);
        const NodalGeom gs = coarsen(T.g);
        MGX_HIP(hipMemsetAsync(T.sub_sig.p, 0, gs.nsig() * sizeof(double), s));
        sigma_coarsen_kernel<double><<<grid3(gs.n[0], gs.n[1], gs.n[2]), block3(), 0, s>>>(
            kgeom(T.g), kgeom(gs), sigma0(), T.sub_sig.as<double>());
        scale_cast_kernel<double><<<grid1d(gs.nsig()), 256, 0, s>>>(
            T.sub_sig.as<double>(), T.sub_sig.as<double>(), gs.nsig(), 1.0 / smax);
        T.sub.set_sigma(T.sub_sig.as<double>(), s);
        T.sigma_ok = true;
    }
}

// One V-cycle on the FP16 level-0 residual H16.b; the correction (scaled) ends in H16.x.
void NodalSolver::vcycle16(hipStream_t s) {
    HalfTop& T = H16;
    const MGOpts& o = opts;
    const long nn = g0.nnode();
    SArgs<__half, __half> a{};
    a.g = T.kg;
    a.c = T.coef;
    a.b = T.b.as<__half>();
    a.sig = T.sig.as<__half>();
    a.omega = (float)nodal_omega(T.g, o.omega);
    int sweeps = o.nu1;
    if (sweeps > 0) {
        if (!T.ready) {
            a.xout = T.x;
            launch_stencil<__half, __half, KS_JFIRST>(T.g, a, s, false);
        }
        --sweeps;
    } else {
        zero_kernel<__half><<<grid1d(nn), 256, 0, s>>>(T.x, nn);
    }
    T.ready = false;
    auto jacobi = [&]() {
        a.x = T.x;
        a.xout = T.x2;
        fine_timer.begin(s);
        launch_stencil<__half, __half, KS_JACOBI>(T.g, a, s, true);
        fine_timer.end(s);
        std::swap(T.x, T.x2);
    };
    for (int q = 0; q < sweeps; ++q) jacobi();
    {
        SArgs<__half, __half> r = a;
        r.x = T.x;
        r.xout = nullptr;
        r.rout = T.r.as<__half>();
        launch_stencil<__half, __half, KS_RESID>(T.g, r, s, false);
    }
    MGLevel<float>& c1 = T.sub.L[0];
    c1.x = c1.x_buf.as<float>();
    c1.x2 = c1.x2_buf.as<float>();
    const bool x1 = o.nu1 > 0 && T.sub.L.size() > 1;
    launch_restrict<float, __half>(T.kg, c1.g, c1.kg, T.r.as<__half>(), c1.b, c1.sig, c1.coef.diag4,
                                   (float)nodal_omega(c1.g, o.omega), x1 ? c1.x : nullptr, s);
    T.sub.x0_ready = x1;
    T.sub.vcycle(o, s, ibuf.as<int>());
    nodal_prolong_add_coarse_kernel<float, __half>
        <<<grid3(c1.g.nn[0], c1.g.nn[1], c1.g.nn[2]), block3(), 0, s>>>(T.kg, c1.kg, c1.x, T.x);
    for (int q = 0; q < o.nu2; ++q) jacobi();
}

static void prepare_h64(NodalSolver& S, hipStream_t s) {
    if (!S.H64.built || (int)S.H64.L.size() != S.nlev) S.H64.build(S.g0, S.nlev);
    // level-0 sigma of the FP64 hierarchy is the solver's own array
    if (S.H64.L[0].sig != S.sigma0()) {
        S.H64.L[0].sig_buf.release();
        S.H64.L[0].sig = S.sigma0();
    }
    S.res64.alloc(S.g0.nnode() * sizeof(double));
    S.H64.L[0].b = S.res64.as<double>();
    S.H64.set_sigma(S.sigma0(), s);  // coarse levels; level 0 is shared
}

// ---- autotuning of the launch shapes ----------------------------------------

namespace {

struct TuneTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TuneTimer() {
        MGX_HIP(hipEventCreate(&e0));
        MGX_HIP(hipEventCreate(&e1));
    }
    ~TuneTimer() {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    // microseconds per call of f, after two warm-up calls
    template <class F>
    double time(F&& f, hipStream_t s, int reps = 5) {
        f();
        f();
        MGX_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) f();
        MGX_HIP(hipEventRecord(e1, s));
        MGX_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MGX_HIP(hipEventElapsedTime(&ms, e0, e1));
        return 1e3 * ms / reps;
    }
};

// Candidate shapes, one per distinct (NPT, z-run) they give on this level.
// dma: also the LDS-direct variants (Jacobi, 4-byte storage, npt 2)
std::vector<StencilShape> shape_candidates(const NodalGeom& g, bool dma) {
    std::vector<StencilShape> out;
    std::vector<std::pair<int, int>> seen;
    for (int npt : {1, 2, 4}) {
        const int ny = npt == 1 ? 8 : 4 * npt;
        const long gxy = (long)((g.nn[0] + 63) / 64) * ((g.nn[1] + ny - 1) / ny);
        for (long target : {128L, 256L, 512L, 1024L, 2048L, 4096L, 8192L, 16384L})
            for (int kzmin : {1, 2}) {
                int kz = (int)std::min<long>(256, std::max<long>(kzmin, (g.nn[2] * gxy) / target));
                kz = std::min(kz, g.nn[2]);
                const std::pair<int, int> key{npt, kz};
                if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
                seen.push_back(key);
                StencilShape sh;
                sh.npt = npt;
                sh.target = target;
                sh.kzmin = kzmin;
                out.push_back(sh);
                if (dma && npt == 2)
                    for (int v : {1, 2, 3}) {  // LDS-direct, wave strips of 2 / 4 rows
                        sh.dma = v;
                        out.push_back(sh);
                    }
            }
    }
    return out;
}

template <class T, class TC, int MODE>
void tune_pass(const NodalGeom& g, const SArgs<T, TC>& a, hipStream_t s, bool fine, TuneTimer& tt,
               std::ostringstream& rep) {
    const StencilShape def = stencil_shape(g, (int)sizeof(T), (int)sizeof(TC), MODE, fine);
    const double t_def = tt.time([&] { launch_stencil_shape<T, TC, MODE>(g, a, s, fine, def); }, s);
    StencilShape best = def;
    double t_best = t_def;
    constexpr bool dma = MODE == KS_JACOBI && std::is_same<T, float>::value &&
                         std::is_same<TC, float>::value;
    for (const StencilShape& c : shape_candidates(g, dma)) {
        const double t = tt.time([&] { launch_stencil_shape<T, TC, MODE>(g, a, s, fine, c); }, s);
        if (t < 0.97 * t_best) {  // keep the incumbent unless clearly faster
            t_best = t;
            best = c;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mtx);
        g_shapes[ShapeKey{g.nn[0], g.nn[1], g.nn[2], (int)sizeof(T), (int)sizeof(TC), MODE,
                          fine ? 1 : 0}] = best;
    }
    rep << "stencil " << g.nn[0] << " " << g.nn[1] << " " << g.nn[2] << " " << sizeof(T) << " "
        << sizeof(TC) << " " << MODE << " " << (fine ? 1 : 0) << " npt " << best.npt << " target "
        << best.target << " kzmin " << best.kzmin << " dma " << best.dma << " us " << t_best
        << " default_us " << t_def << "\n";
}

template <class T>
void tune_levels(NodalHierarchy<T>& H, const MGOpts& o, hipStream_t s, TuneTimer& tt,
                 std::ostringstream& rep) {
    for (size_t l = 0; l < H.L.size(); ++l) {
        MGLevel<T>& lv = H.L[l];
        const long nn = lv.g.nnode();
        zero_kernel<T><<<grid1d(nn), 256, 0, s>>>(lv.x_buf.template as<T>(), nn);
        zero_kernel<T><<<grid1d(nn), 256, 0, s>>>(lv.x2_buf.template as<T>(), nn);
        zero_kernel<T><<<grid1d(nn), 256, 0, s>>>(lv.b, nn);
        if (lv.r) zero_kernel<T><<<grid1d(nn), 256, 0, s>>>(lv.r, nn);
        SArgs<T, T> a{};
        a.g = lv.kg;
        a.c = lv.coef;
        a.x = lv.x_buf.template as<T>();
        a.b = lv.b;
        a.sig = lv.sig;
        a.xout = lv.x2_buf.template as<T>();
        a.rout = lv.r ? lv.r : lv.x2_buf.template as<T>();
        a.omega = (T)nodal_omega(lv.g, o.omega);
        tune_pass<T, T, KS_JACOBI>(lv.g, a, s, l == 0, tt, rep);
        if (l + 1 < H.L.size()) tune_pass<T, T, KS_RESID>(lv.g, a, s, false, tt, rep);
        tune_pass<T, T, KS_JFIRST>(lv.g, a, s, false, tt, rep);
    }
}

}  // namespace

std::string NodalSolver::autotune(int mode, hipStream_t s) {
    MGX_REQUIRE(sigma_set, MGX_ERR_STATE, "autotune: sigma has not been set (project once first)");
    MGX_REQUIRE(!eb && !mask_on, MGX_ERR_UNSUPPORTED,
                "autotune: plain (no EB, no overset mask) levels only");
    std::ostringstream rep;
    rep.precision(4);
    TuneTimer tt;
    const bool mixed = mode != MGX_MODE_FP64;
    const long nn = g0.nnode();
    DevBuf tb, tx, ty;
    tb.alloc(nn * sizeof(double));
    tx.alloc(nn * sizeof(double));
    ty.alloc(nn * sizeof(double));
    zero_kernel<double><<<grid1d(nn), 256, 0, s>>>(tb.as<double>(), nn);
    zero_kernel<double><<<grid1d(nn), 256, 0, s>>>(tx.as<double>(), nn);
    const Coef<double> c64 = make_coef<double>(g0.dx);
    auto outer = [&](auto& H, auto zero_cor) {
        using TC = typename std::remove_pointer<decltype(zero_cor)>::type;
        SArgs<double, TC> a{};
        a.g = kgeom(g0);
        a.c = c64;
        a.x = tx.as<double>();
        a.b = tb.as<double>();
        a.sig = sigma0();
        a.cor = H.L[0].x_buf.template as<TC>();
        a.xout = ty.as<double>();
        a.rout = H.L[0].b;
        a.norm = scal.as<unsigned long long>() + 3;
        if (opts.nu1 > 0 && H.L.size() > 1) {
            a.x1out = H.L[0].x2_buf.template as<TC>();
            a.omega1 = nodal_omega(g0, opts.omega);
        }
        tune_pass<double, TC, KS_RESID>(g0, a, s, false, tt, rep);
        tune_pass<double, TC, KS_UPDRES>(g0, a, s, false, tt, rep);
    };
    auto coarse = [&](auto& H) {
        // the level count folded into the single-workgroup coarse launch
        std::vector<long> cands{1};
        for (auto& lv : H.L)
            if (lv.g.nnode() <= 40000 && lv.g.nnode() > 1) cands.push_back(lv.g.nnode());
        long best = coarse_threshold();
        double t_best = 1e300, t_def = 0.0;
        const long def = best;
        for (long c : cands) {
            {
                std::lock_guard<std::mutex> lk(g_tune_mtx);
                g_coarse_nodes = c;
            }
            H.drop_graphs();
            const double t = tt.time([&] {
                H.x0_ready = false;
                H.x0_guess = false;
                H.vcycle(opts, s, ibuf.as<int>());
            }, s);
            if (c == def) t_def = t;
            if (t < t_best) {
                t_best = t;
                best = c;
            }
        }
        {
            std::lock_guard<std::mutex> lk(g_tune_mtx);
            g_coarse_nodes = best;
        }
        H.drop_graphs();
        rep << "coarse_nodes " << best << " vcycle_us " << t_best << " default_us " << t_def << "\n";
    };
    if (mixed) {
        prepare_h32(s);
        tune_levels(H32, opts, s, tt, rep);
        outer(H32, (float*)nullptr);
        coarse(H32);
        H32.x0_ready = H32.x0_guess = false;
    } else {
        prepare_h64(*this, s);
        tune_levels(H64, opts, s, tt, rep);
        outer(H64, (double*)nullptr);
        coarse(H64);
        H64.x0_ready = H64.x0_guess = false;
    }
    MGX_HIP(hipStreamSynchronize(s));
    return rep.str();
}

std::string tuning_table() {
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    std::ostringstream o;
    for (auto& kv : g_shapes) {
        const ShapeKey& k = kv.first;
        o << "stencil " << k.nn0 << " " << k.nn1 << " " << k.nn2 << " " << k.word << " " << k.cword
          << " " << k.mode << " " << k.fine << " " << kv.second.npt << " " << kv.second.target << " "
          << kv.second.kzmin << " " << kv.second.dma << "\n";
    }
    o << "coarse_nodes " << g_coarse_nodes << "\n";
    o << "jacobi_pairs " << g_jacobi_pairs << "\n";
    o << "mac_coarse_cells " << g_mac_coarse_cells << "\n";
    o << "amr_tile_width " << g_amr_tile_width << "\n";
    o << "amr_tile_width32 " << g_amr_tile_width32 << "\n";
    o << "amr_tile_kz " << g_amr_tile_kz << "\n";
    return o.str();
}

void set_tuning_table(const std::string& text) {
    std::istringstream in(text);
    std::string tag;
    std::map<ShapeKey, StencilShape> shapes;
    long coarse = 1, mac_coarse = 512;
    int pairs = 0, tilew = 16, tilew32 = 64, tilekz = TileSet::kKZ;
    while (in >> tag) {
        if (tag == "stencil") {
            ShapeKey k{};
            StencilShape sh;
            in >> k.nn0 >> k.nn1 >> k.nn2 >> k.word >> k.cword >> k.mode >> k.fine >> sh.npt >>
                sh.target >> sh.kzmin >> sh.dma;
            MGX_REQUIRE(!in.fail() && (sh.npt == 1 || sh.npt == 2 || sh.npt == 4) && sh.target > 0 &&
                            (sh.kzmin == 1 || sh.kzmin == 2) && sh.dma >= 0 && sh.dma <= 3,
                        MGX_ERR_ARG, "tuning table: bad stencil line");
            // the LDS-direct and strip sweeps exist for the 4-byte Jacobi pass with npt 2 only
            if (!(k.mode == KS_JACOBI && k.word == 4 && k.cword == 4 && sh.npt == 2)) sh.dma = 0;
            shapes[k] = sh;
        } else if (tag == "coarse_nodes") {
            in >> coarse;
            MGX_REQUIRE(!in.fail() && coarse >= 1, MGX_ERR_ARG, "tuning table: bad coarse_nodes");
        } else if (tag == "amr_tile_width") {
            in >> tilew;
            MGX_REQUIRE(!in.fail() && (tilew == 16 || tilew == 64), MGX_ERR_ARG,
                        "tuning table: bad amr_tile_width");
        } else if (tag == "amr_tile_width32") {
            in >> tilew32;
            MGX_REQUIRE(!in.fail() && (tilew32 == 16 || tilew32 == 64), MGX_ERR_ARG,
                        "tuning table: bad amr_tile_width32");
        } else if (tag == "amr_tile_kz") {
            in >> tilekz;
            MGX_REQUIRE(!in.fail() && tilekz >= 2 && tilekz <= 64, MGX_ERR_ARG,
                        "tuning table: bad amr_tile_kz");
        } else if (tag == "mac_coarse_cells") {
            in >> mac_coarse;
            MGX_REQUIRE(!in.fail() && mac_coarse >= 0, MGX_ERR_ARG, "tuning table: bad mac_coarse_cells");
        } else if (tag == "jacobi_pairs") {
            in >> pairs;
            MGX_REQUIRE(!in.fail() && (pairs == 0 || pairs == 1), MGX_ERR_ARG,
                        "tuning table: bad jacobi_pairs");
        } else {
            throw Error(MGX_ERR_ARG, "tuning table: unknown line '" + tag + "'");
        }
    }
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    g_shapes = std::move(shapes);
    g_coarse_nodes = coarse;
    g_mac_coarse_cells = mac_coarse;
    g_jacobi_pairs = pairs;
    g_amr_tile_width = tilew;
    g_amr_tile_width32 = tilew32;
    g_amr_tile_kz = tilekz;
}

// ---- learned coarse-level predictor ---------------------------------------

static NodalGeom level_geom(const NodalGeom& g0, int lvl) {
    NodalGeom g = g0;
    for (int l = 0; l < lvl; ++l) g = coarsen(g);
    return g;
}

void NodalSolver::check_learned() const {
    MGX_REQUIRE(net != nullptr, MGX_ERR_STATE,
                "learned mode: no network attached to the nodal projector");
    MGX_REQUIRE(g0.per[0] && g0.per[1] && g0.per[2], MGX_ERR_UNSUPPORTED,
                "learned mode: only fully periodic domains are supported");
    MGX_REQUIRE(net_level >= 0 && net_level < nlev, MGX_ERR_ARG,
                "learned mode: network level is not a multigrid level of this problem");
    const NodalGeom gl = level_geom(g0, net_level);
    const int m = 1 << (net->levels - 1);
    for (int d = 0; d < 3; ++d)
        MGX_REQUIRE(gl.nn[d] % m == 0, MGX_ERR_UNSUPPORTED,
                    "learned mode: the network level's size must be a multiple of 2^(levels-1)");
}

// H32.L[0].b holds an FP32 right-hand side; restrict it to level lvl and pack
// the normalised features there.  Scalars land in scal[40..42].
void NodalSolver::learned_features(int lvl, float* in, hipStream_t s) {
    for (int l = 0; l < lvl; ++l)
        nodal_restrict_kernel<float><<<grid3(H32.L[l + 1].g.nn[0], H32.L[l + 1].g.nn[1], H32.L[l + 1].g.nn[2]),
                                       block3(), 0, s>>>(H32.L[l].kg, H32.L[l + 1].kg, H32.L[l].b,
                                                          H32.L[l + 1].b);
    MGLevel<float>& lv = H32.L[lvl];
    const long n = lv.g.nnode();
    const int np = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
    double* sc = scal.as<double>() + 40;
    const double h2 = std::pow(lv.g.dx[0] * lv.g.dx[1] * lv.g.dx[2], 2.0 / 3.0);
    if (logsig_lvl != lvl) {
        feature_partials_kernel<float><<<np, 256, 0, s>>>(lv.kg, lv.b, lv.sig, 1, part.as<double>());
        feature_scalars_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, n, 1, h2, sc);
        logsig_lvl = lvl;
    }
    feature_partials_kernel<float><<<np, 256, 0, s>>>(lv.kg, lv.b, lv.sig, 0, part.as<double>());
    feature_scalars_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, n, 0, h2, sc);
    pack_features_kernel<float><<<grid3(lv.g.nn[0], lv.g.nn[1], lv.g.nn[2]), block3(), 0, s>>>(
        lv.kg, lv.b, lv.sig, sc, in);
}

void NodalSolver::learned_correction(hipStream_t s) {
    const int lc = net_level;
    MGLevel<float>& lv = H32.L[lc];
    const long n = lv.g.nnode();
    net_in.alloc(8 * n * sizeof(float));
    net_out.alloc(n * sizeof(float));
    learned_features(lc, net_in.as<float>(), s);
    NetShape sh{};
    sh.nx = lv.g.nn[0];
    sh.ny = lv.g.nn[1];
    sh.nz = lv.g.nn[2];
    sh.batch = 1;
    sh.per[0] = sh.per[1] = sh.per[2] = 1;
    if (H32.graphs) {
        net->forward_graph(net_in.as<float>(), sh, net_out.as<float>(), s);
    } else {
        const hipStream_t saved = net->stream;
        net->stream = s;
        try {
            net->forward(net_in.as<float>(), sh, net_out.as<float>());
        } catch (...) {
            net->stream = saved;
            throw;
        }
        net->stream = saved;
    }
    unscale_kernel<float><<<grid1d(n), 256, 0, s>>>(net_out.as<float>(), n, scal.as<double>() + 40, lv.x);
    // upward half of a V-cycle from the predicted coarse correction, with the
    // smoother the V-cycles use (set here: this can run before the first cycle)
    H32.cheb = use_chebyshev(opts, true);
    for (int l = lc - 1; l >= 0; --l) {
        MGLevel<float>& f = H32.L[l];
        zero_kernel<float><<<grid1d(f.g.nnode()), 256, 0, s>>>(f.x, f.g.nnode());
        nodal_prolong_add_kernel<float><<<grid3(f.g.nn[0], f.g.nn[1], f.g.nn[2]), block3(), 0, s>>>(
            f.kg, H32.L[l + 1].kg, H32.L[l + 1].x, f.x);
        // not level 0: the first refinement step's V-cycles pre-smooth from
        // this guess there
        H32.smooth(l, l == 0 ? 0 : opts.nu2, false, opts, s);
    }
}

void NodalSolver::solve(double* b, double rtol, double atol, int mode, mgx_solve_info* info,
                        hipStream_t s) {
    const bool zero_start = phi_zero;  // one solve only, also when it fails
    phi_zero = false;
    MGX_REQUIRE(sigma_set, MGX_ERR_STATE, "nodal solve: sigma has not been set");
    MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED || mode == MGX_MODE_LEARNED,
                MGX_ERR_UNSUPPORTED, "nodal solve: unknown mode");
    if (eb && eb->ready) {
        MGX_REQUIRE(!mask_on, MGX_ERR_UNSUPPORTED,
                    "nodal solve: an overset mask with embedded boundaries is not supported");
        MGX_REQUIRE(mode == MGX_MODE_FP64, MGX_ERR_UNSUPPORTED,
                    "nodal solve: embedded boundaries run in FP64 mode only");
        eb->set_sigma(sigma0(), s);
        eb->solve(b, phi_a.as<double>(), phi_b.as<double>(), phi, rtol, atol, opts, scal, part,
                  h_pinned, info, s);
        return;
    }
    const bool learned = (mode == MGX_MODE_LEARNED);
    if (learned) check_learned();
    const NodalGeom& g = g0;
    const KGeom kg = kgeom(g);
    const long nn = g.nnode();
    bool mixed = (mode != MGX_MODE_FP64);
    // the mixed V-cycle runs on the FP16-stored finest level when there is a
    // level below it; the learned step works on the FP32 hierarchy
    MGX_REQUIRE(!(mask_on && learned), MGX_ERR_UNSUPPORTED,
                "nodal solve: the learned mode does not take an overset mask");
    MGX_REQUIRE(!mask_on || opts.nu2 >= 1, MGX_ERR_UNSUPPORTED,
                "nodal solve: an overset mask needs at least one post-smoothing sweep");
    const bool half = mixed && use_half && nlev >= 2 && !mask_on;
    H64.timer = &fine_timer;
    H32.timer = &fine_timer;
    if (mixed && (!half || learned)) prepare_h32(s);
    if (half) prepare_h16(s);
    if (!mixed) prepare_h64(*this, s);
    // overset mask on every level of the hierarchy in use; a problem with a
    // fixed node is not singular
    const unsigned char* mask0 = mask_on ? mask_lv[0].as<unsigned char>() : nullptr;
    auto attach = [&](auto& H) {
        H.masks.clear();
        H.bottom_fixed = mask_bottom_fixed;
        if (mask_on)
            for (int l = 0; l < nlev; ++l) H.masks.push_back(mask_lv[l].as<unsigned char>());
    };
    if (H32.built) attach(H32);
    if (H64.built) attach(H64);
    const bool sing = g0.singular() && !mask_fixed;
    if (mask0) mask_zero_kernel<<<grid1d(nn), 256, 0, s>>>(mask0, b, nn);

    const bool has_dir = g.dlo[0] || g.dlo[1] || g.dlo[2] || g.dhi[0] || g.dhi[1] || g.dhi[2];
    if (has_dir) {
        zero_dirichlet_kernel<double><<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, s>>>(kg, b);
        zero_dirichlet_kernel<double><<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, s>>>(kg, phi);
    }
    {
        // one pass for sum/min/max of b; ||b - mean||_inf follows from min/max
        const int np = (int)std::min<long>(1024, std::max<long>(1, (nn + 255) / 256));
        stats_kernel<<<np, 256, 0, s>>>(b, nn, part.as<double>(), np);
        final_stats_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, nn, sing ? 1 : 0,
                                             scal.as<double>() + 33, scal.as<unsigned long long>());
        if (sing)
            subtract_scalar_kernel<double><<<grid1d(nn), 256, 0, s>>>(b, nn, scal.as<double>() + 33);
    }
    MGX_HIP(hipMemsetAsync(ibuf.p, 0, sizeof(int), s));
    const double bnorm = read_norm(0, s);
    const Coef<double> c64 = make_coef<double>(g.dx);

    // The outer residual pass also writes the next V-cycle's first Jacobi
    // sweep from zero (omega r / diag) into the hierarchy's spare level-0
    // array (the live one may be the correction this pass reads).
    auto first_sweep = [&](auto& a, auto& H) {
        a.nmask = mask0;
        if (opts.nu1 > 0 && H.L.size() > 1 && !mask0) {
            a.x1out = H.L[0].x2;
            a.omega1 = sweep_omega(g, opts, use_chebyshev(opts, sizeof(*H.L[0].x) == 4), 0,
                                   opts.nu1);
        }
    };
    auto done_first_sweep = [&](auto& a, auto& H) {
        if (a.x1out) {
            std::swap(H.L[0].x, H.L[0].x2);
            H.x0_ready = true;
        }
    };
    // FP16 level: the stored residual is r / s16 and the stored first sweep
    // belongs to the scaled operator (L dx0^2 / sigma_max)
    double s16 = 1.0;
    auto half_scales = [&](SArgs<double, __half>& a, double sc) {
        s16 = sc > 0.0 ? sc : 1.0;
        a.rscale = 1.0 / s16;
        if (opts.nu1 > 0) {
            a.x1out = H16.x2;
            a.omega1 = nodal_omega(g, opts.omega);
            a.x1scale = H16.sigmax / (s16 * H16.dx0 * H16.dx0);
        }
    };
    auto half_done = [&](const SArgs<double, __half>& a) {
        if (a.x1out) {
            std::swap(H16.x, H16.x2);
            H16.ready = true;
        }
    };
    enum { P64 = 0, P32 = 1, P16 = 2 };
    // r = b - L phi into the V-cycle's level-0 right-hand side; sc: a scale
    // for the FP16 residual (an estimate of ||r||_inf)
    auto residual = [&](int prec, double sc) {
        zero_norm(1, s);
        if (prec == P16) {
            SArgs<double, __half> a{};
            a.g = kg; a.c = c64; a.x = phi; a.b = b; a.sig = sigma0();
            a.rout = H16.b.as<__half>(); a.norm = scal.as<unsigned long long>() + 1;
            half_scales(a, sc);
            launch_stencil<double, __half, KS_RESID>(g, a, s, false);
            half_done(a);
        } else if (prec == P32) {
            SArgs<double, float> a{};
            a.g = kg; a.c = c64; a.x = phi; a.b = b; a.sig = sigma0();
            a.rout = res32.as<float>(); a.norm = scal.as<unsigned long long>() + 1;
            first_sweep(a, H32);
            launch_stencil<double, float, KS_RESID>(g, a, s, false);
            done_first_sweep(a, H32);
        } else {
            SArgs<double, double> a{};
            a.g = kg; a.c = c64; a.x = phi; a.b = b; a.sig = sigma0();
            a.rout = res64.as<double>(); a.norm = scal.as<unsigned long long>() + 1;
            first_sweep(a, H64);
            launch_stencil<double, double, KS_RESID>(g, a, s, false);
            done_first_sweep(a, H64);
        }
        return read_norm(1, s);
    };
    // phi_out = phi + cor, r = b - L phi_out (FP64), returns ||r||_inf
    auto update = [&](int prec, double* pout, double sc) {
        zero_norm(1, s);
        if (prec == P16) {
            SArgs<double, __half> a{};
            a.g = kg; a.c = c64; a.x = phi; a.cor = H16.x; a.b = b; a.sig = sigma0();
            a.cor_scale = H16.dx0 * H16.dx0 * s16 / H16.sigmax;
            a.xout = pout; a.rout = H16.b.as<__half>(); a.norm = scal.as<unsigned long long>() + 1;
            half_scales(a, sc);
            launch_stencil<double, __half, KS_UPDRES>(g, a, s, false);
            half_done(a);
        } else if (prec == P32) {
            SArgs<double, float> a{};
            a.g = kg; a.c = c64; a.x = phi; a.cor = H32.L[0].x; a.b = b; a.sig = sigma0();
            a.xout = pout; a.rout = res32.as<float>(); a.norm = scal.as<unsigned long long>() + 1;
            first_sweep(a, H32);
            launch_stencil<double, float, KS_UPDRES>(g, a, s, false);
            done_first_sweep(a, H32);
        } else {
            SArgs<double, double> a{};
            a.g = kg; a.c = c64; a.x = phi; a.cor = H64.L[0].x; a.b = b; a.sig = sigma0();
            a.xout = pout; a.rout = res64.as<double>(); a.norm = scal.as<unsigned long long>() + 1;
            first_sweep(a, H64);
            launch_stencil<double, double, KS_UPDRES>(g, a, s, false);
            done_first_sweep(a, H64);
        }
        return read_norm(1, s);
    };
    auto other = [&]() { return (phi == phi_a.as<double>()) ? phi_b.as<double>() : phi_a.as<double>(); };

    const int vprec = mixed ? (half ? P16 : P32) : P64;
    const int prec0 = (learned || !half) ? (mixed ? P32 : P64) : P16;
    const bool from_zero = zero_start && prec0 != P16;
    // phi = 0: r0 = b (Dirichlet rows zeroed, mean removed), ||r0|| = ||b||;
    // the first V-cycle then runs its own first sweep
    const double r0 = from_zero ? bnorm : residual(prec0, bnorm);
    if (from_zero) {
        if (prec0 == P32) {
            cast_kernel<float, double><<<grid1d(nn), 256, 0, s>>>(b, res32.as<float>(), nn);
            H32.x0_ready = false;
        } else {
            MGX_HIP(hipMemcpyAsync(res64.p, b, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
            H64.x0_ready = false;
        }
    }
    int vcycles = 0;
    const double tol = std::max(atol, rtol * bnorm);
    double rn = r0;
    int it = 0;
    int fallback = 0;
    // a non-finite norm (NaN mapped to +inf in the reductions) ends the solve unconverged
    bool conv = std::isfinite(bnorm) && r0 <= tol;
    // learned correction of the current FP32 residual, judged in FP32 on the
    // correction equation: kept (as the refinement step's initial guess, its
    // V-cycles pre-smooth from it) only if ||r - L e|| < ||r||; the FP64
    // residual after that step judges the iterate as always
    auto learned_guess = [&](double rnow) {
        ++vcycles;
        learned_correction(s);
        MGLevel<float>& l0 = H32.L[0];
        SArgs<float, float> a{};
        a.g = l0.kg;
        a.c = l0.coef;
        a.x = l0.x;
        a.b = l0.b;
        a.sig = l0.sig;
        a.rout = l0.r;
        zero_norm(2, s);
        a.norm = scal.as<unsigned long long>() + 2;
        launch_stencil<float, float, KS_RESID>(l0.g, a, s, false);
        const double rl = read_norm(2, s);
        if (opts.verbose)
            fprintf(stderr, "mgx nodal MLMG: learned step |r - L e| %.6e (from %.6e)\n", rl, rnow);
        H32.x0_ready = false;
        H32.x0_guess = rl < rnow;
        return rl < rnow;
    };
    int lsteps = 0;  // refinement steps started from the network so far
    if (learned && !conv && opts.maxiter > 0 && !half) {
        ++lsteps;
        if (!learned_guess(rn)) fallback = 1;
    } else if (learned && !conv && opts.maxiter > 0) {
        // FP16-stored cycles: the step is an FP64 update of its own
        ++it;
        ++vcycles;
        learned_correction(s);
        double* pout = other();
        const double rl = update(P32, pout, rn);
        if (opts.verbose)
            fprintf(stderr, "mgx nodal MLMG: learned step |r| %.6e (from %.6e)\n", rl, rn);
        if (rl < rn) {
            phi = pout;
            rn = rl;
            if (half) rn = residual(P16, rn);
        } else {
            fallback = 1;
            rn = residual(vprec, rn);
        }
        conv = (rn <= tol);
    }
    int prec = vprec;
    int stall = 0;
    int ncyc = opts.cycles_per_update;  // FP32 cycles of the next refinement step
    while (!conv && it < opts.maxiter && std::isfinite(rn)) {
        ++it;
        double* pout = other();
        int ran = 1;
        if (learned && mixed && prec == P32 && it > 1 &&
            (opts.learned_steps < 0 || lsteps < opts.learned_steps)) {
            // a further step from the network's correction of its residual
            ++lsteps;
            learned_guess(rn);
        }
        if (prec == P16) vcycle16(s);
        else if (prec == P32) {
            // further FP32 cycles on the same FP64 residual, each from the
            // correction so far (the FP64 update and residual run once)
            H32.vcycle(opts, s, ibuf.as<int>());
            for (; ran < ncyc; ++ran) {
                H32.x0_guess = true;
                H32.vcycle(opts, s, ibuf.as<int>());
            }
        } else H64.vcycle(opts, s, ibuf.as<int>());
        vcycles += ran;
        const double prev = rn;
        rn = update(prec, pout, rn);
        phi = pout;
        if (prec == P32 && opts.cycles_per_update > 1 && rn > tol && rn < prev) {
            // the next step runs the cycles the measured per-cycle contraction
            // says are left, at most cycles_per_update
            const double rho = std::pow(rn / prev, 1.0 / ran);
            const double need = std::ceil(std::log(tol / rn) / std::log(rho));
            ncyc = (int)std::max(1.0, std::min((double)opts.cycles_per_update, need));
        }
        if (opts.verbose)
            fprintf(stderr, "mgx nodal MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                    bnorm > 0 ? rn / bnorm : rn);
        conv = (rn <= tol);
        if (learned && mixed && !conv && opts.stall_iters > 0) {
            // plain FP64 MLMG takes over when the reduced-precision cycle stalls
            stall = (rn > 0.9 * prev) ? stall + 1 : 0;
            if (stall >= opts.stall_iters) {
                fallback = 1;
                mixed = false;
                prec = P64;
                prepare_h64(*this, s);
                attach(H64);
                rn = residual(P64, rn);
            }
        }
    }
    if (sing) remove_mean(*this, phi, nn, s);
    int bot = 0;
    MGX_HIP(hipMemcpyAsync(&bot, ibuf.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    if (info) {
        info->iters = it;
        info->vcycles = vcycles;
        info->converged = conv ? 1 : 0;
        info->nlevels = nlev;
        info->fallback_used = fallback;
        info->bnorm = bnorm;
        info->resnorm0 = r0;
        info->resnorm = rn;
        info->rel_resid = bnorm > 0 ? rn / bnorm : rn;
    }
}

// Training pair for the learned predictor from the canonical rhs b and the
// converged phi: in = features of b at level lvl, target = phi injected there.
void NodalSolver::training_pair(const double* b, int lvl, float* in, float* target, hipStream_t s) {
    MGX_REQUIRE(sigma_set, MGX_ERR_STATE, "training pair: sigma has not been set");
    MGX_REQUIRE(g0.per[0] && g0.per[1] && g0.per[2], MGX_ERR_UNSUPPORTED,
                "training pair: only fully periodic domains are supported");
    MGX_REQUIRE(lvl >= 0 && lvl < nlev, MGX_ERR_ARG, "training pair: level out of range");
    prepare_h32(s);
    const long nn = g0.nnode();
    cast_kernel<float, double><<<grid1d(nn), 256, 0, s>>>(b, res32.as<float>(), nn);
    learned_features(lvl, in, s);
    const MGLevel<float>& lv = H32.L[lvl];
    inject_target_kernel<<<grid3(lv.g.nn[0], lv.g.nn[1], lv.g.nn[2]), block3(), 0, s>>>(
        H32.L[0].kg, lv.kg, 1 << lvl, phi, scal.as<double>() + 40, target);
    MGX_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------

void BoxLayout::build(const int* b6, int nb, const NodalGeom& g, bool partial) {
    MGX_REQUIRE(nb > 0 && b6 != nullptr, MGX_ERR_ARG, "box list is empty");
    boxes.assign(b6, b6 + 6 * nb);
    nboxes = nb;
    long vol = 0;
    for (int b = 0; b < nb; ++b) {
        for (int d = 0; d < 3; ++d) {
            const int lo = b6[6 * b + d], hi = b6[6 * b + 3 + d];
            MGX_REQUIRE(lo >= 0 && hi < g.n[d] && lo <= hi, MGX_ERR_ARG,
                        "box " + std::to_string(b) + " lies outside the domain");
        }
        vol += (long)(b6[6 * b + 3] - b6[6 * b] + 1) * (b6[6 * b + 4] - b6[6 * b + 1] + 1) *
               (b6[6 * b + 5] - b6[6 * b + 2] + 1);
    }
    MGX_REQUIRE(partial || vol == (long)g.n[0] * g.n[1] * g.n[2], MGX_ERR_UNSUPPORTED,
                "single-level solve needs boxes that tile the whole domain");
    for (int a = 0; a < nb; ++a)
        for (int b = a + 1; b < nb; ++b) {
            bool ov = true;
            for (int d = 0; d < 3; ++d)
                if (b6[6 * a + 3 + d] < b6[6 * b + d] || b6[6 * b + 3 + d] < b6[6 * a + d]) ov = false;
            MGX_REQUIRE(!ov, MGX_ERR_ARG, "boxes overlap");
        }
    std::vector<BoxInfo> info(nb);
    long pc[4] = {0, 0, 0, 0}, pn[4] = {0, 0, 0, 0}, pf[3][4] = {};
    std::vector<Tile> tiles[REG_NREG];
    for (int b = 0; b < nb; ++b) {
        BoxInfo& bi = info[b];
        for (int d = 0; d < 3; ++d) {
            bi.lo[d] = b6[6 * b + d];
            bi.len[d] = b6[6 * b + 3 + d] - b6[6 * b + d] + 1;
            const bool glo = !g.per[d] && bi.lo[d] == 0;
            const bool ghi = !g.per[d] && bi.lo[d] + bi.len[d] == g.n[d];
            bi.rlo[d] = bi.lo[d] - (glo ? 1 : 0);
            bi.rlen[d] = bi.len[d] + (glo ? 1 : 0) + (ghi ? 1 : 0);
        }
        for (int q = 0; q < 4; ++q) {
            bi.pc[q] = pc[q];
            bi.pn[q] = pn[q];
            for (int d = 0; d < 3; ++d) bi.pf[d][q] = pf[d][q];
        }
        const long a0 = bi.len[0], a1 = bi.len[1], a2 = bi.len[2];
        pc[0] += a0 * a1 * a2; pc[1] += a0 * a1 + a1 * a2 + a0 * a2; pc[2] += a0 + a1 + a2; pc[3] += 1;
        const long n0 = a0 + 1, n1 = a1 + 1, n2 = a2 + 1;
        pn[0] += n0 * n1 * n2; pn[1] += n0 * n1 + n1 * n2 + n0 * n2; pn[2] += n0 + n1 + n2; pn[3] += 1;
        long vols[REG_NREG] = {a0 * a1 * a2, (long)bi.rlen[0] * bi.rlen[1] * bi.rlen[2], n0 * n1 * n2};
        for (int d = 0; d < 3; ++d) {
            const long f0 = a0 + (d == 0), f1 = a1 + (d == 1), f2 = a2 + (d == 2);
            pf[d][0] += f0 * f1 * f2; pf[d][1] += f0 * f1 + f1 * f2 + f0 * f2;
            pf[d][2] += f0 + f1 + f2; pf[d][3] += 1;
            vols[REG_FACE0 + d] = f0 * f1 * f2;
        }
        for (int r = 0; r < REG_NREG; ++r)
            for (long st = 0; st < vols[r]; st += 256) tiles[r].push_back(Tile{b, (int)st});
    }
    dboxes.alloc(nb * sizeof(BoxInfo));
    MGX_HIP(hipMemcpy(dboxes.p, info.data(), nb * sizeof(BoxInfo), hipMemcpyHostToDevice));
    for (int r = 0; r < REG_NREG; ++r) {
        ntiles[r] = (int)tiles[r].size();
        dtiles[r].alloc(tiles[r].size() * sizeof(Tile));
        MGX_HIP(hipMemcpy(dtiles[r].p, tiles[r].data(), tiles[r].size() * sizeof(Tile),
                          hipMemcpyHostToDevice));
    }
}

}  // namespace mgx

// mac_mg.hip -- host side of the MAC projection's cell-centred multigrid:
// red-black Gauss-Seidel V-cycle (cell-average restriction, piecewise-constant
// prolongation, face-averaged coarse beta, one-workgroup FP64 BiCGStab
// bottom) inside the same outer MLMG iteration as the nodal solve (FP64
// residual, FP64 or FP32 V-cycle correction).  Same steps as the oracle
// (oracle/orc_mac.c).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mac_mg.h"
#include "mgx_ctx.h"

namespace mgx {

template <class T>
MacHierarchy<T>::~MacHierarchy() {
    drop_graphs();
    if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

template <class T>
void MacHierarchy<T>::drop_graphs() {
    for (CycleGraph& c : cache_) (void)hipGraphExecDestroy(c.exec);
    cache_.clear();
}

template <class T>
void MacHierarchy<T>::build(const MacGeom& g0, int nlev) {
    drop_graphs();  // they hold the old levels' addresses
    L.clear();
    L.resize(nlev);
    for (int l = 0; l < nlev; ++l) {
        MacLevel<T>& lv = L[l];
        lv.g = (l == 0) ? g0 : L[l - 1].g.coarse();
        lv.kg = lv.g.k();
        lv.c = make_mcoef<T>(lv.g.dx);
        for (int d = 0; d < 3; ++d) {
            lv.beta_buf[d].alloc(lv.g.nface(d) * sizeof(T));
            lv.beta[d] = lv.beta_buf[d].template as<T>();
        }
        const long n = lv.g.ncell();
        if (lv.g.has_a) {
            lv.a_buf.alloc(n * sizeof(T));
            lv.a = lv.a_buf.template as<T>();
        } else {
            lv.a = nullptr;
        }
        if (l > 0) lv.b_buf.alloc(n * sizeof(T));
        lv.x_buf.alloc(n * sizeof(T));
        if (l < nlev - 1) lv.r_buf.alloc(n * sizeof(T));
        lv.b = lv.b_buf.template as<T>();
        lv.x = lv.x_buf.template as<T>();
        lv.r = lv.r_buf.template as<T>();
    }
    bottom_work.alloc(7 * L.back().g.ncell() * sizeof(double));
    built = true;
}

template <class T>
void MacHierarchy<T>::set_beta(const double* const b0[3], hipStream_t s,
                               const std::vector<const double*>* rscale) {
    for (int d = 0; d < 3; ++d) {
        const long nf = L[0].g.nface(d);
        cast_kernel<T, double><<<grid1d(nf), 256, 0, s>>>(b0[d], L[0].beta[d], nf);
    }
    for (size_t l = 1; l < L.size(); ++l)
        for (int d = 0; d < 3; ++d) {
            const MacGeom& gc = L[l].g;
            mac_beta_coarsen_kernel<T, T>
                <<<grid3(gc.n[0] + (d == 0), gc.n[1] + (d == 1), gc.n[2] + (d == 2)), block3(), 0, s>>>(
                    L[l - 1].kg, L[l].kg, d, L[l - 1].beta[d], L[l].beta[d]);
        }
    // Robin planes last: every level was coarsened from unscaled values
    if (rscale)
        for (size_t l = 0; l < L.size(); ++l)
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    const double* sc = (*rscale)[l * 6 + 2 * d + side];
                    if (!sc) continue;
                    const MacGeom& gl = L[l].g;
                    const long np = (long)gl.n[d == 0 ? 1 : 0] * gl.n[d == 2 ? 1 : 2];
                    robin_beta_kernel<T><<<grid1d(np), 256, 0, s>>>(L[l].kg, d, side, L[l].beta[d],
                                                                    sc);
                }
}

template <class T>
void MacHierarchy<T>::set_acoef(const double* a0, hipStream_t s) {
    if (!L[0].a) return;
    const long n0 = L[0].g.ncell();
    cast_kernel<T, double><<<grid1d(n0), 256, 0, s>>>(a0, L[0].a, n0);
    for (size_t l = 1; l < L.size(); ++l) {
        const MacGeom& gc = L[l].g;
        mac_restrict_kernel<T><<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            L[l - 1].kg, L[l].kg, L[l - 1].a, L[l].a);
    }
}

template <class T>
static void gsrb(MacLevel<T>& lv, hipStream_t s) {
    for (int color = 0; color < 2; ++color)
        mac_gsrb_kernel<T><<<grid3((lv.g.n[0] + 1) / 2, lv.g.n[1], lv.g.n[2]), block3(), 0, s>>>(
            lv.kg, lv.c, lv.beta[0], lv.beta[1], lv.beta[2], lv.a, lv.b, lv.x, color);
}

// the options a V-cycle's launches depend on
static bool same_cycle_opts(const MGOpts& a, const MGOpts& b) {
    return a.nu1 == b.nu1 && a.nu2 == b.nu2 && a.bottom_maxiter == b.bottom_maxiter &&
           a.bottom_rtol == b.bottom_rtol && a.bottom_solver == b.bottom_solver;
}

template <class T>
void MacHierarchy<T>::vcycle(const MGOpts& o, hipStream_t s, int* bottom_iters) {
    if (!graphs) {
        vcycle_body(o, s, bottom_iters);
        return;
    }
    const bool guess0 = x0_guess;
    const long coarse = mac_coarse_cells();
    const CycleGraph* hit = nullptr;
    for (const CycleGraph& c : cache_)
        if (c.x0 == L[0].x && c.b0 == L[0].b && c.guess0 == guess0 && c.bottom_iters == bottom_iters &&
            c.coarse == coarse && same_cycle_opts(c.o, o))
            hit = &c;
    if (!hit) {
        if (cache_.size() >= 16) drop_graphs();
        bool ok = true;
        if (!cap_stream_ && hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking) != hipSuccess) {
            cap_stream_ = nullptr;
            ok = false;
        }
        if (ok && hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeRelaxed) != hipSuccess) ok = false;
        if (ok) {
            bool body_ok = true;
            try {
                vcycle_body(o, cap_stream_, bottom_iters);
            } catch (...) {
                body_ok = false;
            }
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(cap_stream_, &g);
            hipGraphExec_t exec = nullptr;
            if (e == hipSuccess && body_ok) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e == hipSuccess && body_ok) {
                cache_.push_back(CycleGraph{L[0].x, L[0].b, guess0, bottom_iters, coarse, o, exec});
                hit = &cache_.back();
            } else {
                ok = false;
            }
        }
        if (!ok) {
            // capture not available here: plain launches from now on
            (void)hipGetLastError();
            graphs = false;
            x0_guess = guess0;
            vcycle_body(o, s, bottom_iters);
            return;
        }
    }
    MGX_HIP(hipGraphLaunch(hit->exec, s));
    x0_guess = false;
}

static_assert(sizeof(MacCoarseArgs<double>) <= 3072, "coarse-cycle kernel arguments");

// the first level of the one-workgroup coarse cycle (nl - 1: none)
template <class T>
static int mac_coarse_first(const std::vector<MacLevel<T>>& L) {
    const int nl = (int)L.size();
    const long maxc = mac_coarse_cells();
    int lc = nl - 1;
    while (lc - 1 >= 1 && L[lc - 1].g.ncell() <= maxc && nl - 1 - (lc - 1) <= kMacCoarseMax) --lc;
    return lc;
}

template <class T>
static MacCoarseArgs<T> mac_coarse_args(std::vector<MacLevel<T>>& L, int lc, const MGOpts& o, int up) {
    MacCoarseArgs<T> a{};
    const int nl = (int)L.size();
    a.nlev = nl - 1 - lc;
    a.nu1 = o.nu1;
    a.nu2 = o.nu2;
    a.up = up;
    for (int q = 0; q <= a.nlev; ++q) {
        MacLevel<T>& v = L[lc + q];
        a.L[q].g = v.kg;
        a.L[q].c = v.c;
        for (int d = 0; d < 3; ++d) a.L[q].beta[d] = v.beta[d];
        a.L[q].a = v.a;
        a.L[q].b = v.b;
        a.L[q].x = v.x;
        a.L[q].r = v.r;
    }
    return a;
}

template <class T>
void MacHierarchy<T>::vcycle_body(const MGOpts& o, hipStream_t s, int* bottom_iters) {
    const int nl = (int)L.size();
    // levels lc..nl-2 in one workgroup (lc == nl - 1: none)
    const int lc = mac_coarse_first(L);
    for (int l = 0; l < lc; ++l) {
        MacLevel<T>& lv = L[l];
        const long n = lv.g.ncell();
        if (l > 0 || !x0_guess) zero_kernel<T><<<grid1d(n), 256, 0, s>>>(lv.x, n);
        for (int q = 0; q < o.nu1; ++q) gsrb(lv, s);
        mac_resid_kernel<T, T, MK_RESID><<<grid3(lv.g.n[0], lv.g.n[1], lv.g.n[2]), block3(), 0, s>>>(
            lv.kg, lv.c, lv.beta[0], lv.beta[1], lv.beta[2], lv.a, lv.b, lv.x, nullptr, nullptr,
            lv.r, nullptr);
        MacLevel<T>& cv = L[l + 1];
        mac_restrict_kernel<T><<<grid3(cv.g.n[0], cv.g.n[1], cv.g.n[2]), block3(), 0, s>>>(
            lv.kg, cv.kg, lv.r, cv.b);
    }
    if (lc < nl - 1) mac_coarse_cycle_kernel<T><<<1, 1024, 0, s>>>(mac_coarse_args(L, lc, o, 0));
    {
        MacLevel<T>& lb = L[nl - 1];
        const long nb = lb.g.ncell();
        const int threads = nb >= 1024 ? 1024 : (nb >= 256 ? 256 : 64);
        mac_bottom_kernel<T><<<1, threads, 0, s>>>(lb.kg, make_mcoef<double>(lb.g.dx), lb.beta[0],
                                                   lb.beta[1], lb.beta[2], lb.a, lb.b, lb.x,
                                                   bottom_work.as<double>(), o.bottom_solver,
                                                   o.bottom_maxiter,
                                                   o.bottom_rtol, lb.g.singular() ? 1 : 0,
                                                   bottom_iters);
    }
    x0_guess = false;
    if (lc < nl - 1) mac_coarse_cycle_kernel<T><<<1, 1024, 0, s>>>(mac_coarse_args(L, lc, o, 1));
    for (int l = lc - 1; l >= 0; --l) {
        MacLevel<T>& lv = L[l];
        mac_prolong_add_kernel<T><<<grid3(lv.g.n[0], lv.g.n[1], lv.g.n[2]), block3(), 0, s>>>(
            lv.kg, L[l + 1].kg, L[l + 1].x, lv.x);
        for (int q = 0; q < o.nu2; ++q) gsrb(lv, s);
    }
}

template class MacHierarchy<double>;
template class MacHierarchy<float>;

// ---------------------------------------------------------------------------

MacSolver::~MacSolver() {
    if (h_pinned) (void)hipHostFree(h_pinned);
}

static int mac_nlevels(const MacGeom& g, int max_coarsening_level) {
    NodalGeom ng{};
    for (int d = 0; d < 3; ++d) ng.n[d] = g.n[d];
    return nodal_nlevels(ng, max_coarsening_level);
}

void MacSolver::setup(const MacGeom& g, const MGOpts& o) {
    // face coefficients do not depend on the boundary types or the multigrid
    // options: a re-setup for new ones keeps them
    const bool keep_beta = beta_set && g.n[0] == g0.n[0] && g.n[1] == g0.n[1] && g.n[2] == g0.n[2] &&
                           g.per[0] == g0.per[0] && g.per[1] == g0.per[1] && g.per[2] == g0.per[2];
    g0 = g;
    opts = o;
    nlev = mac_nlevels(g, o.max_coarsening_level);
    for (int d = 0; d < 3; ++d) beta0[d].alloc(g.nface(d) * sizeof(double));
    const long n = g.ncell();
    if (g.has_a) acoef0.alloc(n * sizeof(double));
    h32_a = h64_a = false;
    phi_a.alloc(n * sizeof(double));
    phi_b.alloc(n * sizeof(double));
    MGX_HIP(hipMemset(phi_a.p, 0, n * sizeof(double)));
    phi = phi_a.as<double>();
    scal.alloc(64 * sizeof(double));
    part.alloc(4096 * sizeof(double));
    ibuf.alloc(16 * sizeof(int));
    if (!h_pinned) MGX_HIP(hipHostMalloc((void**)&h_pinned, 64 * sizeof(unsigned long long), 0));
    H64.built = H32.built = false;
    beta_set = keep_beta;
    acoef_set = acoef_set && keep_beta && g.has_a;
    h32_beta = h64_beta = false;
    rob_levels = 0;  // the level count may have changed
    beta0s_ok = false;
}

void MacSolver::set_robin(int d, int side, std::vector<double> a, std::vector<double> b) {
    rob_a[d][side] = std::move(a);
    rob_b[d][side] = std::move(b);
    robin_on = true;
    rob_levels = 0;
    beta0s_ok = false;
    h32_beta = h64_beta = false;
}

void MacSolver::clear_robin() {
    for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side) {
            rob_a[d][side].clear();
            rob_b[d][side].clear();
            rob_scale[d][side].clear();
        }
    robin_on = false;
    rob_levels = 0;
    rob_ptr.clear();
    beta0s_ok = false;
    h32_beta = h64_beta = false;
}

void MacSolver::robin_prepare(hipStream_t s) {
    if (!robin_on) return;
    if (rob_levels != nlev) {
        rob_ptr.assign((size_t)nlev * 6, nullptr);
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                rob_scale[d][side].clear();
                if (rob_a[d][side].empty()) continue;
                rob_scale[d][side].resize(nlev);
                const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
                int m0 = g0.n[a0], m1 = g0.n[a1];
                std::vector<double> a = rob_a[d][side], b = rob_b[d][side];
                double h = g0.dx[d];
                for (int l = 0; l < nlev; ++l) {
                    if (l > 0) {  // face averages over 2 x 2 fine faces
                        const int c0 = m0 / 2, c1 = m1 / 2;
                        std::vector<double> ac((size_t)c0 * c1), bc((size_t)c0 * c1);
                        for (int q = 0; q < c1; ++q)
                            for (int p = 0; p < c0; ++p) {
                                double sa = 0, sb = 0;
                                for (int e = 0; e < 4; ++e) {
                                    const long f = (2 * p + (e & 1)) + (long)m0 * (2 * q + (e >> 1));
                                    sa += a[f];
                                    sb += b[f];
                                }
                                ac[p + (long)c0 * q] = 0.25 * sa;
                                bc[p + (long)c0 * q] = 0.25 * sb;
                            }
                        a.swap(ac);
                        b.swap(bc);
                        m0 = c0;
                        m1 = c1;
                        h *= 2;
                    }
                    std::vector<double> sc(a.size());
                    for (size_t t = 0; t < a.size(); ++t) sc[t] = a[t] * h / (a[t] * h + 2.0 * b[t]);
                    DevBuf& buf = rob_scale[d][side][l];
                    buf.alloc(sc.size() * sizeof(double));
                    MGX_HIP(hipMemcpyAsync(buf.p, sc.data(), sc.size() * sizeof(double),
                                           hipMemcpyHostToDevice, s));
                    MGX_HIP(hipStreamSynchronize(s));
                    rob_ptr[(size_t)l * 6 + 2 * d + side] = buf.as<double>();
                }
            }
        rob_levels = nlev;
        beta0s_ok = false;
    }
    if (!beta0s_ok) {
        for (int d = 0; d < 3; ++d) {
            const long nf = g0.nface(d);
            beta0s[d].alloc(nf * sizeof(double));
            MGX_HIP(hipMemcpyAsync(beta0s[d].p, beta(d), nf * sizeof(double), hipMemcpyDeviceToDevice, s));
            for (int side = 0; side < 2; ++side) {
                const double* sc = rob_ptr[2 * d + side];
                if (!sc) continue;
                const long np = (long)g0.n[d == 0 ? 1 : 0] * g0.n[d == 2 ? 1 : 2];
                robin_beta_kernel<double><<<grid1d(np), 256, 0, s>>>(g0.k(), d, side,
                                                                     beta0s[d].as<double>(), sc);
            }
        }
        beta0s_ok = true;
    }
}

const double* MacSolver::fine_beta(int d, hipStream_t s) {
    if (!robin_on) return beta(d);
    robin_prepare(s);
    return beta0s[d].as<double>();
}

double MacSolver::read_norm(int slot, hipStream_t s) {
    MGX_HIP(hipMemcpyAsync(h_pinned + slot, scal.as<unsigned long long>() + slot,
                           sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    MGX_HIP(hipStreamSynchronize(s));
    double d;
    std::memcpy(&d, h_pinned + slot, sizeof(d));
    return d;
}

void MacSolver::solve(double* b, double rtol, double atol, int mode, mgx_solve_info* info,
                      hipStream_t s) {
    MGX_REQUIRE(beta_set, MGX_ERR_STATE, "MAC solve: beta has not been set");
    MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_UNSUPPORTED,
                "MAC solve: mode not available for the MAC projection");
    const MacGeom& g = g0;
    const MGeomK kg = g.k();
    const long n = g.ncell();
    const bool mixed = (mode == MGX_MODE_MIXED);
    robin_prepare(s);
    // unscaled coefficients feed the level coarsening; the operator uses the
    // Robin-scaled level-0 ones
    const double* braw0[3] = {beta(0), beta(1), beta(2)};
    const double* b0[3] = {fine_beta(0, s), fine_beta(1, s), fine_beta(2, s)};
    const std::vector<const double*>* rs = robin_on ? &rob_ptr : nullptr;
    MGX_REQUIRE(!g.has_a || acoef_set, MGX_ERR_STATE, "Helmholtz solve: a coefficient not set");
    const double* A0 = acoef();
    if (mixed) {
        if (!H32.built || (int)H32.L.size() != nlev) { H32.build(g, nlev); h32_beta = h32_a = false; }
        if (!h32_beta) { H32.set_beta(braw0, s, rs); h32_beta = true; }
        if (A0 && !h32_a) { H32.set_acoef(A0, s); h32_a = true; }
        res32.alloc(n * sizeof(float));
        H32.L[0].b = res32.as<float>();
    } else {
        if (!H64.built || (int)H64.L.size() != nlev) { H64.build(g, nlev); h64_beta = h64_a = false; }
        if (!h64_beta) { H64.set_beta(braw0, s, rs); h64_beta = true; }
        if (A0 && !h64_a) { H64.set_acoef(A0, s); h64_a = true; }
        res64.alloc(n * sizeof(double));
        H64.L[0].b = res64.as<double>();
    }
    {
        const int np = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
        stats_kernel<<<np, 256, 0, s>>>(b, n, part.as<double>(), np);
        final_stats_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, n, g.singular() ? 1 : 0,
                                             scal.as<double>() + 33, scal.as<unsigned long long>());
        if (g.singular())
            subtract_scalar_kernel<double><<<grid1d(n), 256, 0, s>>>(b, n, scal.as<double>() + 33);
    }
    MGX_HIP(hipMemsetAsync(ibuf.p, 0, sizeof(int), s));
    const double bnorm = read_norm(0, s);
    const MCoef<double> c64 = make_mcoef<double>(g.dx);
    const dim3 gr = grid3(g.n[0], g.n[1], g.n[2]);
    unsigned long long* nrm = scal.as<unsigned long long>() + 1;

    zero_u64_kernel<<<1, 64, 0, s>>>(nrm, 1);
    if (mixed)
        mac_resid_kernel<double, float, MK_RESID><<<gr, block3(), 0, s>>>(
            kg, c64, b0[0], b0[1], b0[2], A0, b, phi, nullptr, nullptr, res32.as<float>(), nrm);
    else
        mac_resid_kernel<double, double, MK_RESID><<<gr, block3(), 0, s>>>(
            kg, c64, b0[0], b0[1], b0[2], A0, b, phi, nullptr, nullptr, res64.as<double>(), nrm);
    const double r0 = read_norm(1, s);
    const double tol = std::max(atol, rtol * bnorm);
    double rn = r0;
    int it = 0;
    int vcycles = 0;
    bool conv = std::isfinite(bnorm) && r0 <= tol;
    int ncyc = opts.cycles_per_update;  // FP32 cycles of the next refinement step
    while (!conv && it < opts.maxiter && std::isfinite(rn)) {
        ++it;
        zero_u64_kernel<<<1, 64, 0, s>>>(nrm, 1);
        double* pout = (phi == phi_a.as<double>()) ? phi_b.as<double>() : phi_a.as<double>();
        int ran = 1;
        const double prev = rn;
        if (mixed) {
            // further FP32 cycles on the same FP64 residual, each from the
            // correction so far
            H32.vcycle(opts, s, ibuf.as<int>());
            for (; ran < ncyc; ++ran) {
                H32.x0_guess = true;
                H32.vcycle(opts, s, ibuf.as<int>());
            }
            mac_resid_kernel<double, float, MK_UPDRES><<<gr, block3(), 0, s>>>(
                kg, c64, b0[0], b0[1], b0[2], A0, b, phi, H32.L[0].x, pout, res32.as<float>(), nrm);
        } else {
            H64.vcycle(opts, s, ibuf.as<int>());
            mac_resid_kernel<double, double, MK_UPDRES><<<gr, block3(), 0, s>>>(
                kg, c64, b0[0], b0[1], b0[2], A0, b, phi, H64.L[0].x, pout, res64.as<double>(), nrm);
        }
        phi = pout;
        rn = read_norm(1, s);
        vcycles += ran;
        if (opts.verbose)
            fprintf(stderr, "mgx MAC MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                    bnorm > 0 ? rn / bnorm : rn);
        conv = (rn <= tol);
        if (mixed && opts.cycles_per_update > 1 && !conv && rn < prev) {
            // the cycles the measured per-cycle contraction says are left
            const double rho = std::pow(rn / prev, 1.0 / ran);
            const double need = std::ceil(std::log(tol / rn) / std::log(rho));
            ncyc = (int)std::max(1.0, std::min((double)opts.cycles_per_update, need));
        }
    }
    if (g.singular()) {
        const int np = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
        stats_kernel<<<np, 256, 0, s>>>(phi, n, part.as<double>(), np);
        final_stats_kernel<<<1, 256, 0, s>>>(part.as<double>(), np, n, 1, scal.as<double>() + 34,
                                             scal.as<unsigned long long>() + 2);
        subtract_scalar_kernel<double><<<grid1d(n), 256, 0, s>>>(phi, n, scal.as<double>() + 34);
    }
    MGX_HIP(hipStreamSynchronize(s));
    if (info) {
        info->iters = it;
        info->vcycles = vcycles;
        info->converged = conv ? 1 : 0;
        info->nlevels = nlev;
        info->fallback_used = 0;
        info->bnorm = bnorm;
        info->resnorm0 = r0;
        info->resnorm = rn;
        info->rel_resid = bnorm > 0 ? rn / bnorm : rn;
    }
}

}  // namespace mgx

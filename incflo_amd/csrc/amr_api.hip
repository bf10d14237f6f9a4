// amr_api.hip -- two-level composite nodal projection (AMR, refinement ratio 2)
// and its C ABI.
//
// The reference builds one NodalProjector over all AMR levels
// (src/projection/incflo_apply_nodal_projection.cpp:124-215), projects the
// velocity of every level, returns phi / G phi per level and averages gp
// down afterwards (:258-266).  Here the composite problem is the conforming
// Q1 discretisation described in amr_kernels.h, solved by composite V-cycles:
//   fine pre-smoothing on the fine unknowns (dF held at 0), the composite
//   residual restricted to the whole coarse grid, one V-cycle of the coarse
//   multigrid hierarchy, trilinear prolongation onto the fine unknowns and dF,
//   fine post-smoothing
// inside an FP64 outer iteration on the composite residual (MLMG's stopping
// rule ||r||_inf <= max(atol, rtol ||b||_inf)); MGX_MODE_MIXED runs the
// V-cycle in FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "amr_kernels.h"
#include "mgx_ctx.h"
#include "nodal_mg.h"

namespace mgx {

// per node of g: whether any / all of its in-domain adjacent cells are set
static void node_any_all(const NodalGeom& g, const std::vector<unsigned char>& cells,
                         std::vector<unsigned char>& any, std::vector<unsigned char>& all) {
    any.assign(g.nnode(), 0);
    all.assign(g.nnode(), 0);
    for (int k = 0; k < g.nn[2]; ++k)
        for (int j = 0; j < g.nn[1]; ++j)
            for (int i = 0; i < g.nn[0]; ++i) {
                bool a = false, b = true;
                for (int q = 0; q < 8; ++q) {
                    int c[3] = {i - 1 + (q & 1), j - 1 + ((q >> 1) & 1), k - 1 + ((q >> 2) & 1)};
                    bool inside = true;
                    for (int d = 0; d < 3; ++d) {
                        if (g.per[d]) c[d] = (c[d] + g.n[d]) % g.n[d];
                        else if (c[d] < 0 || c[d] >= g.n[d]) inside = false;
                    }
                    if (!inside) continue;
                    const bool v = cells[c[0] + (long)g.n[0] * (c[1] + (long)g.n[1] * c[2])] != 0;
                    a = a || v;
                    b = b && v;
                }
                const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
                any[id] = a;
                all[id] = b;
            }
}

static NodalGeom refine_geom(const NodalGeom& c) {
    NodalGeom f = c;
    for (int d = 0; d < 3; ++d) {
        f.n[d] = 2 * c.n[d];
        f.nn[d] = c.per[d] ? f.n[d] : f.n[d] + 1;
        f.dx[d] = 0.5 * c.dx[d];
        f.sg[d] = f.n[d] + 2;
    }
    return f;
}

// block of the point kernels launched over a TileSet's blocks
static dim3 tile_block() { return dim3(TileSet::kTX, TileSet::kTY, 1); }

static bool is_dir_host(const NodalGeom& g, int i, int j, int k) {
    return (g.dlo[0] && i == 0) || (g.dhi[0] && i == g.nn[0] - 1) || (g.dlo[1] && j == 0) ||
           (g.dhi[1] && j == g.nn[1] - 1) || (g.dlo[2] && k == 0) || (g.dhi[2] && k == g.nn[2] - 1);
}

class AmrNodal {
   public:
    NodalGeom gc, gf;
    MGOpts opts;
    int nlev = 1;
    NodalHierarchy<double> H64;
    NodalHierarchy<float> H32;
    DevBuf bC;                          // level-0 rhs of the coarse correction (FP64-sized)
    DevBuf sigc, sigc_out, sigc_out32, sigf, sigf32;
    DevBuf ftype, cdof, ccov;
    DevBuf cnear;                       // coarse nodes with a node of F's closure among their 27 fine nodes
    DevBuf xc, bc, rtmp, xf, bf, yf;
    DevBuf ef_a, ef_b, rfT, yT, rcT;    // V-cycle work (FP64-sized, reused for FP32)
    DevBuf scal, part, ibuf;
    TileSet ftiles;                     // the fine level's nodes in or on F: every fine launch
    unsigned long long* h_pinned = nullptr;
    double ndof = 0.0;
    bool sigma_set = false;
    bool h64_sigma = false, h32_sigma = false;

    ~AmrNodal() {
        if (h_pinned) (void)hipHostFree(h_pinned);
    }

    void setup(const NodalGeom& g0, const std::vector<int>& fboxes, const MGOpts& o) {
        // the boxes and the periodicity are fixed per handle, so a set-up for
        // new boundary types or options keeps the coefficient arrays
        const bool keep = sigma_set;
        gc = g0;
        gf = refine_geom(g0);
        opts = o;
        nlev = nodal_nlevels(gc, o.max_coarsening_level);
        // host masks: fine cells of F, covered coarse cells, node types
        std::vector<unsigned char> fm((size_t)gf.n[0] * gf.n[1] * gf.n[2], 0);
        for (size_t b = 0; b + 5 < fboxes.size(); b += 6)
            for (int k = fboxes[b + 2]; k <= fboxes[b + 5]; ++k)
                for (int j = fboxes[b + 1]; j <= fboxes[b + 4]; ++j)
                    for (int i = fboxes[b + 0]; i <= fboxes[b + 3]; ++i)
                        fm[i + (size_t)gf.n[0] * (j + (size_t)gf.n[1] * k)] = 1;
        std::vector<unsigned char> cov((size_t)gc.n[0] * gc.n[1] * gc.n[2], 0);
        for (int k = 0; k < gc.n[2]; ++k)
            for (int j = 0; j < gc.n[1]; ++j)
                for (int i = 0; i < gc.n[0]; ++i)
                    cov[i + (size_t)gc.n[0] * (j + (size_t)gc.n[1] * k)] =
                        fm[2 * i + (size_t)gf.n[0] * (2 * j + (size_t)gf.n[1] * 2 * k)];
        std::vector<unsigned char> anyf, allf, anyc, allc;
        node_any_all(gf, fm, anyf, allf);
        node_any_all(gc, cov, anyc, allc);
        std::vector<unsigned char> ft(gf.nnode()), cd(gc.nnode());
        long nf = 0, nc = 0;
        for (long q = 0; q < gf.nnode(); ++q) {
            ft[q] = anyf[q] ? (allf[q] ? AMR_DOF : AMR_BDRY) : AMR_OUT;
            nf += ft[q] == AMR_DOF;
        }
        for (int k = 0; k < gc.nn[2]; ++k)
            for (int j = 0; j < gc.nn[1]; ++j)
                for (int i = 0; i < gc.nn[0]; ++i) {
                    const long id = i + (long)gc.nn[0] * (j + (long)gc.nn[1] * k);
                    cd[id] = allc[id] ? 0 : 1;
                    nc += cd[id] && !is_dir_host(gc, i, j, k);
                }
        ndof = (double)(nf + nc);
        {
            std::vector<unsigned char> act(ft.size());
            for (size_t q = 0; q < ft.size(); ++q) act[q] = ft[q] != AMR_OUT;
            ftiles.build(gf, act);
        }
        {
            // the coarse nodes whose full-weighting stencil reaches a fine node
            // of F's closure: the only ones the restrictions of fine data touch
            std::vector<unsigned char> nr(gc.nnode(), 0);
            for (int K = 0; K < gc.nn[2]; ++K)
                for (int J = 0; J < gc.nn[1]; ++J)
                    for (int I = 0; I < gc.nn[0]; ++I) {
                        bool any = false;
                        for (int oz = -1; oz <= 1 && !any; ++oz)
                            for (int oy = -1; oy <= 1 && !any; ++oy)
                                for (int ox = -1; ox <= 1 && !any; ++ox) {
                                    int f[3] = {2 * I + ox, 2 * J + oy, 2 * K + oz};
                                    bool in = true;
                                    for (int d = 0; d < 3; ++d) {
                                        if (gf.per[d]) f[d] = (f[d] + gf.nn[d]) % gf.nn[d];
                                        else if (f[d] < 0 || f[d] >= gf.nn[d]) in = false;
                                    }
                                    if (in)
                                        any = ft[f[0] + (size_t)gf.nn[0] * (f[1] + (size_t)gf.nn[1] * f[2])] !=
                                              AMR_OUT;
                                }
                        nr[I + (size_t)gc.nn[0] * (J + (size_t)gc.nn[1] * K)] = any;
                    }
            cnear.alloc(nr.size());
            MGX_HIP(hipMemcpy(cnear.p, nr.data(), nr.size(), hipMemcpyHostToDevice));
        }
        ftype.alloc(ft.size());
        cdof.alloc(cd.size());
        ccov.alloc(cov.size());
        MGX_HIP(hipMemcpy(ftype.p, ft.data(), ft.size(), hipMemcpyHostToDevice));
        MGX_HIP(hipMemcpy(cdof.p, cd.data(), cd.size(), hipMemcpyHostToDevice));
        MGX_HIP(hipMemcpy(ccov.p, cov.data(), cov.size(), hipMemcpyHostToDevice));
        const long nnc = gc.nnode(), nnf = gf.nnode();
        sigc.alloc(gc.nsig() * sizeof(double));
        sigc_out.alloc(gc.nsig() * sizeof(double));
        sigc_out32.alloc(gc.nsig() * sizeof(float));
        sigf.alloc(gf.nsig() * sizeof(double));
        sigf32.alloc(gf.nsig() * sizeof(float));
        // ghost cells outside non-periodic faces stay 0 (no element there)
        if (!keep) {
            MGX_HIP(hipMemset(sigc.p, 0, gc.nsig() * sizeof(double)));
            MGX_HIP(hipMemset(sigc_out.p, 0, gc.nsig() * sizeof(double)));
            MGX_HIP(hipMemset(sigf.p, 0, gf.nsig() * sizeof(double)));
        }
        for (DevBuf* b : {&xc, &bc, &rtmp, &rcT}) b->alloc(nnc * sizeof(double));
        // the fine launches visit only the tiles that hold nodes of F's
        // closure; what lies elsewhere in these arrays is never read but is
        // multiplied by the zero sigma outside F, so it starts finite (0)
        for (DevBuf* b : {&xf, &bf, &yf, &ef_a, &ef_b, &rfT, &yT}) {
            b->alloc(nnf * sizeof(double));
            MGX_HIP(hipMemset(b->p, 0, nnf * sizeof(double)));
        }
        bC.alloc(nnc * sizeof(double));
        scal.alloc(64 * sizeof(double));
        part.alloc(4096 * sizeof(double));
        ibuf.alloc(16 * sizeof(int));
        if (!h_pinned) MGX_HIP(hipHostMalloc((void**)&h_pinned, 64 * sizeof(unsigned long long), 0));
        H64.built = false;
        H32.built = false;
        sigma_set = keep;
        h64_sigma = h32_sigma = false;
    }

    // sigc / sigf (ghosted FP64) are filled by the caller
    void sigma_changed(hipStream_t s) {
        const KGeom kc = kgeom(gc);
        amr_sigma_out_kernel<<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            kc, ccov.as<unsigned char>(), sigc.as<double>(), sigc_out.as<double>());
        cast_kernel<float, double><<<grid1d(gf.nsig()), 256, 0, s>>>(sigf.as<double>(),
                                                                    sigf32.as<float>(), gf.nsig());
        cast_kernel<float, double><<<grid1d(gc.nsig()), 256, 0, s>>>(sigc_out.as<double>(),
                                                                    sigc_out32.as<float>(), gc.nsig());
        sigma_set = true;
        h64_sigma = h32_sigma = false;
    }

    void zero_norm(int slot, hipStream_t s) {
        zero_u64_kernel<<<1, 64, 0, s>>>(scal.as<unsigned long long>() + slot, 1);
    }
    double read_norm(int slot, hipStream_t s) {
        MGX_HIP(hipMemcpyAsync(h_pinned + slot, scal.as<unsigned long long>() + slot,
                               sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        MGX_HIP(hipStreamSynchronize(s));
        double d;
        std::memcpy(&d, &h_pinned[slot], sizeof(d));
        return d;
    }

    // v -= f * (weighted mean over the unknowns of both levels) for singular problems
    void remove_rhs_mean(hipStream_t s) {
        const int np = 512;
        amr_sum_kernel<<<np, 256, 0, s>>>(bc.as<double>(), cdof.as<unsigned char>(), 1, gc.nnode(),
                                          part.as<double>());
        amr_sum_kernel<<<np, 256, 0, s>>>(bf.as<double>(), ftype.as<unsigned char>(), AMR_DOF,
                                          gf.nnode(), part.as<double>() + np);
        amr_mean_kernel<<<1, 64, 0, s>>>(part.as<double>(), np, 8.0, part.as<double>() + np, np, 1.0,
                                         ndof, scal.as<double>() + 40);
        amr_shift_kernel<<<grid1d(gc.nnode()), 256, 0, s>>>(bc.as<double>(), cdof.as<unsigned char>(),
                                                            1, 0, gc.nnode(), scal.as<double>() + 40,
                                                            0.125);
        amr_shift_kernel<<<grid1d(gf.nnode()), 256, 0, s>>>(bf.as<double>(), ftype.as<unsigned char>(),
                                                            AMR_DOF, 0, gf.nnode(),
                                                            scal.as<double>() + 40, 1.0);
    }

    // composite residual of (xc, xf) into rfT / rcT (type TO); returns ||r||_inf
    template <class TO>
    double residual(hipStream_t s) {
        const KGeom kc = kgeom(gc), kf = kgeom(gf);
        const unsigned char* ft = ftype.as<unsigned char>();
        int nbk = 0;
        const int4* tl = ftiles.blocks(nbk);
        if (nbk)
            amr_fill_kernel<double><<<nbk, tile_block(), 0, s>>>(kf, kc, ft, xc.as<double>(),
                                                             xf.as<double>(), tl);
        SArgs<double, double> a{};
        a.g = kf;
        a.c = make_coef<double>(gf.dx);
        a.x = xf.as<double>();
        a.sig = sigf.as<double>();
        a.rout = yf.as<double>();
        launch_stencil<double, double, KS_APPLY>(gf, a, s, false, &ftiles);
        zero_norm(1, s);
        if (nbk)
            amr_fine_resid_kernel<TO><<<nbk, tile_block(), 0, s>>>(kf, ft, bf.as<double>(),
                                                               yf.as<double>(), rfT.as<TO>(),
                                                               scal.as<unsigned long long>() + 1, tl);
        SArgs<double, double> c{};
        c.g = kc;
        c.c = make_coef<double>(gc.dx);
        c.x = xc.as<double>();
        c.b = bc.as<double>();
        c.sig = sigc_out.as<double>();
        c.rout = rtmp.as<double>();
        launch_stencil<double, double, KS_RESID>(gc, c, s, false);
        amr_coarse_combine_kernel<TO><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kf, kc, ft, cdof.as<unsigned char>(), rtmp.as<double>(), yf.as<double>(), -1.0,
            rcT.as<TO>(), scal.as<unsigned long long>() + 1, cnear.as<unsigned char>());
        return read_norm(1, s);
    }

    // one composite V-cycle on the residual in rfT / rcT; the correction is
    // added to (xc, xf)
    template <class T>
    void vcycle(NodalHierarchy<T>& H, const T* sfT, const T* scoT, hipStream_t s) {
        const KGeom kc = kgeom(gc), kf = kgeom(gf);
        const unsigned char* ft = ftype.as<unsigned char>();
        T* e = ef_a.as<T>();
        T* e2 = ef_b.as<T>();
        SArgs<T, T> a{};
        a.g = kf;
        a.c = make_coef<T>(gf.dx);
        a.b = rfT.as<T>();
        a.sig = sfT;
        a.omega = (T)nodal_omega(gf, opts.omega);
        a.nmask = ft;
        int nbk = 0;
        const int4* tl = ftiles.blocks(nbk);
        if (opts.nu1 > 0) {
            a.xout = e;
            launch_stencil<T, T, KS_JFIRST>(gf, a, s, false, &ftiles);
            for (int q = 1; q < opts.nu1; ++q) {
                a.x = e;
                a.xout = e2;
                launch_stencil<T, T, KS_JACOBI>(gf, a, s, false, &ftiles);
                std::swap(e, e2);
            }
        } else {
            zero_kernel<T><<<grid1d(gf.nnode()), 256, 0, s>>>(e, gf.nnode());
        }
        SArgs<T, T> ap{};
        ap.g = kf;
        ap.c = a.c;
        ap.x = e;
        ap.sig = sfT;
        ap.rout = yT.as<T>();
        launch_stencil<T, T, KS_APPLY>(gf, ap, s, false, &ftiles);
        amr_coarse_rhs_kernel<T><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kf, kc, ft, cdof.as<unsigned char>(), rfT.as<T>(), yT.as<T>(), rcT.as<T>(), H.L[0].b,
            cnear.as<unsigned char>());
        H.vcycle(opts, s, ibuf.as<int>());
        if (nbk) amr_prolong_kernel<T><<<nbk, tile_block(), 0, s>>>(kf, kc, ft, H.L[0].x, e, tl);
        for (int q = 0; q < opts.nu2; ++q) {
            a.x = e;
            a.xout = e2;
            launch_stencil<T, T, KS_JACOBI>(gf, a, s, false, &ftiles);
            std::swap(e, e2);
        }
        {
            // the coarse unknowns next to F: one Jacobi sweep with the
            // composite operator (e on dF follows the coarse correction)
            SArgs<T, T> af{};
            af.g = kf;
            af.c = a.c;
            af.x = e;
            af.sig = sfT;
            af.rout = yT.as<T>();
            launch_stencil<T, T, KS_APPLY>(gf, af, s, false, &ftiles);
            SArgs<T, T> ac{};
            ac.g = kc;
            ac.c = make_coef<T>(gc.dx);
            ac.x = H.L[0].x;
            ac.sig = scoT;
            ac.rout = rtmp.as<T>();
            launch_stencil<T, T, KS_APPLY>(gc, ac, s, false);
            amr_coarse_jacobi_kernel<T><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
                kf, kc, ft, cdof.as<unsigned char>(), rcT.as<T>(), rtmp.as<T>(), yT.as<T>(), H.L[0].sig,
                make_coef<double>(gc.dx).diag4, nodal_omega(gc, opts.omega), H.L[0].x,
                cnear.as<unsigned char>());
        }
        if (nbk) amr_add_kernel<T><<<nbk, tile_block(), 0, s>>>(kf, ft, AMR_DOF, e, xf.as<double>(), tl);
        amr_add_kernel<T><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kc, cdof.as<unsigned char>(), 1, H.L[0].x, xc.as<double>());
    }

    template <class T>
    void prepare(NodalHierarchy<T>& H, bool& have_sigma, hipStream_t s) {
        if (!H.built || (int)H.L.size() != nlev) {
            H.build(gc, nlev);
            have_sigma = false;
        }
        if (!have_sigma) {
            H.set_sigma(sigc.as<double>(), s);
            have_sigma = true;
        }
        H.L[0].b = bC.as<T>();
    }

    // composite solve of the right-hand side in (bc, bf); result in (xc, xf)
    // with dF filled, coarse nodes under F injected and (singular) the mean
    // of the coarse phi removed
    void solve(double rtol, double atol, int mode, mgx_solve_info* info, hipStream_t s) {
        MGX_REQUIRE(sigma_set, MGX_ERR_STATE, "AMR nodal projector: sigma not set");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_UNSUPPORTED,
                    "AMR nodal projector: FP64 and mixed modes only");
        const bool mixed = mode == MGX_MODE_MIXED;
        const KGeom kc = kgeom(gc), kf = kgeom(gf);
        const unsigned char* ft = ftype.as<unsigned char>();
        ftiles.wide = amr_tile_width() == 64;
        if (mixed) prepare(H32, h32_sigma, s);
        else prepare(H64, h64_sigma, s);
        MGX_HIP(hipMemsetAsync(ibuf.p, 0, sizeof(int), s));
        if (gc.singular()) remove_rhs_mean(s);
        // ||b||_inf over the unknowns of both levels
        zero_norm(0, s);
        int nbk = 0;
        const int4* tl = ftiles.blocks(nbk);
        if (nbk)
            amr_fine_resid_kernel<double><<<nbk, tile_block(), 0, s>>>(
                kf, ft, bf.as<double>(), nullptr, yf.as<double>(), scal.as<unsigned long long>(), tl);
        amr_coarse_combine_kernel<double><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kf, kc, ft, cdof.as<unsigned char>(), bc.as<double>(), yf.as<double>(), 0.0,
            rtmp.as<double>(), scal.as<unsigned long long>(), cnear.as<unsigned char>());
        const double bnorm = read_norm(0, s);
        MGX_HIP(hipMemsetAsync(xc.p, 0, gc.nnode() * sizeof(double), s));
        MGX_HIP(hipMemsetAsync(xf.p, 0, gf.nnode() * sizeof(double), s));
        const double tol = std::max(atol, rtol * bnorm);
        const double r0 = mixed ? residual<float>(s) : residual<double>(s);
        double rn = r0;
        int it = 0;
        bool conv = std::isfinite(bnorm) && rn <= tol;
        while (!conv && it < opts.maxiter && std::isfinite(rn)) {
            ++it;
            if (mixed) {
                vcycle<float>(H32, sigf32.as<float>(), sigc_out32.as<float>(), s);
                rn = residual<float>(s);
            } else {
                vcycle<double>(H64, sigf.as<double>(), sigc_out.as<double>(), s);
                rn = residual<double>(s);
            }
            if (opts.verbose)
                fprintf(stderr, "mgx AMR nodal MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                        bnorm > 0 ? rn / bnorm : rn);
            conv = rn <= tol;
        }
        // dF from the coarse level, coarse nodes under F from the fine level
        if (nbk)
            amr_fill_kernel<double><<<nbk, tile_block(), 0, s>>>(kf, kc, ft, xc.as<double>(),
                                                             xf.as<double>(), tl);
        amr_inject_kernel<<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kf, kc, cdof.as<unsigned char>(), xf.as<double>(), xc.as<double>());
        if (gc.singular()) {
            const int np = 512;
            amr_sum_kernel<<<np, 256, 0, s>>>(xc.as<double>(), nullptr, 0, gc.nnode(), part.as<double>());
            amr_mean_kernel<<<1, 64, 0, s>>>(part.as<double>(), np, 1.0, nullptr, 0, 0.0,
                                             (double)gc.nnode(), scal.as<double>() + 41);
            amr_shift_kernel<<<grid1d(gc.nnode()), 256, 0, s>>>(xc.as<double>(), nullptr, 0, 0,
                                                                gc.nnode(), scal.as<double>() + 41, 1.0);
            amr_shift_kernel<<<grid1d(gf.nnode()), 256, 0, s>>>(xf.as<double>(), ft, AMR_OUT, 1,
                                                                gf.nnode(), scal.as<double>() + 41, 1.0);
        }
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

struct mgx_amr_s {
    mgx_ctx_s* ctx = nullptr;
    int n_cell[3] = {0, 0, 0};
    double prob_lo[3] = {0, 0, 0}, prob_hi[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    std::vector<int> cboxes, fboxes;
    int bclo[3] = {0, 0, 0}, bchi[3] = {0, 0, 0};
    bool bc_set = false;
    MGOpts opts;
    bool dirty = true;
    AmrNodal solver;
    BoxLayout lay_c, lay_f;
    DevBuf ugc, ugf;
    hipStream_t stream() const { return ctx->stream; }

    void ensure_setup() {
        if (!dirty) return;
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(periodic[d] || bc_set, MGX_ERR_STATE,
                        "AMR nodal projector: setDomainBC is required for non-periodic domains");
        double dx[3];
        for (int d = 0; d < 3; ++d) dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
        const NodalGeom g = make_nodal_geom(n_cell, periodic, bclo, bchi, dx);
        solver.setup(g, fboxes, opts);
        lay_c.build(cboxes.data(), (int)cboxes.size() / 6, g);
        if (!fboxes.empty()) lay_f.build(fboxes.data(), (int)fboxes.size() / 6, solver.gf, true);
        ugc.alloc(3 * g.nsig() * sizeof(double));
        ugf.alloc(3 * solver.gf.nsig() * sizeof(double));
        MGX_HIP(hipMemset(ugc.p, 0, 3 * g.nsig() * sizeof(double)));
        MGX_HIP(hipMemset(ugf.p, 0, 3 * solver.gf.nsig() * sizeof(double)));
        dirty = false;
    }
};

extern "C" {

MGX_API int mgx_amr_nodal_create(mgx_ctx ctx, const mgx_level_desc* levels, int nlevels,
                                 mgx_amr_nodal* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    auto h = std::make_unique<mgx_amr_s>();
    int status = guarded(ctx, [&] {
        MGX_REQUIRE(levels != nullptr && (nlevels == 1 || nlevels == 2), MGX_ERR_UNSUPPORTED,
                    "AMR nodal projector: one coarse level and at most one fine level (ratio 2)");
        const mgx_level_desc& c = levels[0];
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(c.n_cell[d] >= 2 && c.n_cell[d] % 2 == 0, MGX_ERR_ARG,
                        "AMR nodal projector: coarse extents must be even and >= 2");
            h->n_cell[d] = c.n_cell[d];
            h->prob_lo[d] = c.prob_lo[d];
            h->prob_hi[d] = c.prob_hi[d];
            h->periodic[d] = c.is_periodic[d] ? 1 : 0;
        }
        MGX_REQUIRE(c.nboxes > 0 && c.boxes, MGX_ERR_ARG, "AMR nodal projector: empty coarse box list");
        h->cboxes.assign(c.boxes, c.boxes + 6 * c.nboxes);
        if (nlevels == 2) {
            const mgx_level_desc& f = levels[1];
            for (int d = 0; d < 3; ++d) {
                MGX_REQUIRE(f.n_cell[d] == 2 * c.n_cell[d], MGX_ERR_UNSUPPORTED,
                            "AMR nodal projector: refinement ratio 2 only");
                MGX_REQUIRE(f.prob_lo[d] == c.prob_lo[d] && f.prob_hi[d] == c.prob_hi[d] &&
                                (f.is_periodic[d] != 0) == (c.is_periodic[d] != 0),
                            MGX_ERR_ARG, "AMR nodal projector: levels must share the domain");
            }
            MGX_REQUIRE(f.nboxes >= 0 && (f.nboxes == 0 || f.boxes), MGX_ERR_ARG,
                        "AMR nodal projector: bad fine box list");
            for (int b = 0; b < f.nboxes; ++b)
                for (int d = 0; d < 3; ++d) {
                    const int lo = f.boxes[6 * b + d], hi = f.boxes[6 * b + 3 + d];
                    MGX_REQUIRE(lo >= 0 && hi < f.n_cell[d] && lo <= hi, MGX_ERR_ARG,
                                "AMR nodal projector: fine box outside the domain");
                    MGX_REQUIRE(lo % 2 == 0 && hi % 2 == 1, MGX_ERR_ARG,
                                "AMR nodal projector: fine boxes must align with coarse cells");
                    MGX_REQUIRE(c.is_periodic[d] || (lo > 0 && hi < f.n_cell[d] - 1),
                                MGX_ERR_UNSUPPORTED,
                                "AMR nodal projector: fine boxes must stay off non-periodic domain faces");
                }
            h->fboxes.assign(f.boxes, f.boxes + 6 * f.nboxes);
        }
        h->ctx = ctx;
    });
    if (status == MGX_OK) *out = h.release();
    return status;
}

MGX_API int mgx_amr_nodal_destroy(mgx_amr_nodal h) {
    if (!h) return MGX_ERR_ARG;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_amr_nodal_set_domain_bc(mgx_amr_nodal h, const int lo[3], const int hi[3]) {
    if (!h || !lo || !hi) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        for (int d = 0; d < 3; ++d) {
            if (h->periodic[d]) continue;
            for (int v : {lo[d], hi[d]}) {
                MGX_REQUIRE(v == MGX_BC_DIRICHLET || v == MGX_BC_NEUMANN || v == MGX_BC_INFLOW,
                            MGX_ERR_UNSUPPORTED, "AMR nodal projector: unsupported boundary type");
            }
            h->bclo[d] = lo[d];
            h->bchi[d] = hi[d];
        }
        h->bc_set = true;
        h->dirty = true;
    });
}

MGX_API int mgx_amr_nodal_set_options(mgx_amr_nodal h, const mgx_mlmg_opts* o) {
    if (!h || !o) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGOpts m;
        to_opts(o, m);
        h->opts = m;
        h->dirty = true;
    });
}

MGX_API int mgx_amr_nodal_set_sigma(mgx_amr_nodal h, const double* const* sigma, int ng) {
    if (!h || !sigma) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(sigma[0] != nullptr && (h->fboxes.empty() || sigma[1] != nullptr), MGX_ERR_ARG,
                    "AMR nodal projector: sigma needed on every level");
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "sigma ghost width out of range");
        h->ensure_setup();
        AmrNodal& S = h->solver;
        gather_sigma_kernel<<<h->lay_c.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            kgeom(S.gc), h->lay_c.dev_boxes(), h->lay_c.tiles(REG_CELL), sigma[0], ng,
            S.sigc.as<double>());
        MGX_HIP(hipMemsetAsync(S.sigf.p, 0, S.gf.nsig() * sizeof(double), h->stream()));
        if (!h->fboxes.empty())
            gather_sigma_kernel<<<h->lay_f.ntiles[REG_CELL], 256, 0, h->stream()>>>(
                kgeom(S.gf), h->lay_f.dev_boxes(), h->lay_f.tiles(REG_CELL), sigma[1], ng,
                S.sigf.as<double>());
        S.sigma_changed(h->stream());
    });
}

MGX_API int mgx_amr_nodal_project(mgx_amr_nodal h, double* const* vel, int vel_ng,
                                  double* const* phi, int phi_ng, double* const* gphi, int gphi_ng,
                                  double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h || !vel) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        const bool two = !h->fboxes.empty();
        MGX_REQUIRE(vel[0] != nullptr && (!two || vel[1] != nullptr), MGX_ERR_ARG,
                    "AMR nodal projector: velocity needed on every level");
        const bool any_nonper = !h->periodic[0] || !h->periodic[1] || !h->periodic[2];
        MGX_REQUIRE(vel_ng >= (any_nonper ? 1 : 0) && vel_ng <= 8, MGX_ERR_ARG,
                    "velocity needs one ghost cell at non-periodic domain faces");
        MGX_REQUIRE(phi_ng >= 0 && gphi_ng >= 0, MGX_ERR_ARG, "negative ghost width");
        h->ensure_setup();
        AmrNodal& S = h->solver;
        MGX_REQUIRE(S.sigma_set, MGX_ERR_STATE, "AMR nodal projector: sigma not set");
        hipStream_t s = h->stream();
        const KGeom kc = kgeom(S.gc), kf = kgeom(S.gf);
        // coarse divergence of the velocity outside F
        GatherVelArgs a{};
        a.g = kc;
        a.boxes = h->lay_c.dev_boxes();
        a.tiles = h->lay_c.tiles(REG_CELL_GROWN);
        a.vel = vel[0];
        a.vel_ng = vel_ng;
        a.scale = 1.0;
        a.const_rho = 1.0;
        a.ug = h->ugc.as<double>();
        a.ghost_rule = 0;
        gather_vel_kernel<<<h->lay_c.ntiles[REG_CELL_GROWN], 256, 0, s>>>(a);
        amr_zero_covered_kernel<<<grid3(S.gc.n[0], S.gc.n[1], S.gc.n[2]), block3(), 0, s>>>(
            kc, S.ccov.as<unsigned char>(), h->ugc.as<double>());
        nodal_rhs_kernel<<<grid3(S.gc.nn[0], S.gc.nn[1], S.gc.nn[2]), block3(), 0, s>>>(
            kc, h->ugc.as<double>(), S.rtmp.as<double>(), 0.25 / S.gc.dx[0], 0.25 / S.gc.dx[1],
            0.25 / S.gc.dx[2]);
        // fine divergence of the velocity inside F (the gather writes the
        // same cells every call; the rest of ugf stays 0 from the set-up)
        if (two) {
            GatherVelArgs f = a;
            f.g = kf;
            f.boxes = h->lay_f.dev_boxes();
            f.tiles = h->lay_f.tiles(REG_CELL_GROWN);
            f.vel = vel[1];
            f.ug = h->ugf.as<double>();
            gather_vel_kernel<<<h->lay_f.ntiles[REG_CELL_GROWN], 256, 0, s>>>(f);
        }
        int nbk = 0;
        const int4* tl = S.ftiles.blocks(nbk);
        const unsigned char* ft = S.ftype.as<unsigned char>();
        if (nbk) {
            nodal_rhs_kernel<<<nbk, tile_block(), 0, s>>>(kf, h->ugf.as<double>(), S.yf.as<double>(),
                                                      0.25 / S.gf.dx[0], 0.25 / S.gf.dx[1],
                                                      0.25 / S.gf.dx[2], tl);
            amr_fine_resid_kernel<double><<<nbk, tile_block(), 0, s>>>(
                kf, ft, S.yf.as<double>(), nullptr, S.bf.as<double>(), nullptr, tl);
        }
        amr_coarse_combine_kernel<double><<<grid3(S.gc.nn[0], S.gc.nn[1], S.gc.nn[2]), block3(), 0, s>>>(
            kf, kc, ft, S.cdof.as<unsigned char>(), S.rtmp.as<double>(), S.yf.as<double>(), 1.0,
            S.bc.as<double>(), nullptr, S.cnear.as<unsigned char>());
        S.solve(rtol, atol, mode, info, s);
        // u -= sigma G phi on both levels, G phi out; covered coarse cells averaged down
        UpdateArgs uc{};
        uc.g = kc;
        uc.boxes = h->lay_c.dev_boxes();
        uc.tiles = h->lay_c.tiles(REG_CELL);
        uc.phi = S.xc.as<double>();
        uc.ug = h->ugc.as<double>();
        uc.sig = S.sigc.as<double>();
        for (int d = 0; d < 3; ++d) uc.f[d] = 0.25 / S.gc.dx[d];
        uc.vel = vel[0];
        uc.vel_ng = vel_ng;
        uc.gp = gphi ? gphi[0] : nullptr;
        uc.gp_ng = gphi_ng;
        update_cells_kernel<<<h->lay_c.ntiles[REG_CELL], 256, 0, s>>>(uc);
        if (two) {
            UpdateArgs uf = uc;
            uf.g = kf;
            uf.boxes = h->lay_f.dev_boxes();
            uf.tiles = h->lay_f.tiles(REG_CELL);
            uf.phi = S.xf.as<double>();
            uf.ug = h->ugf.as<double>();
            uf.sig = S.sigf.as<double>();
            for (int d = 0; d < 3; ++d) uf.f[d] = 0.25 / S.gf.dx[d];
            uf.vel = vel[1];
            uf.gp = gphi ? gphi[1] : nullptr;
            update_cells_kernel<<<h->lay_f.ntiles[REG_CELL], 256, 0, s>>>(uf);
            AvgDownArgs ad{};
            ad.gc = kc;
            ad.gf = kf;
            ad.boxes = h->lay_c.dev_boxes();
            ad.tiles = h->lay_c.tiles(REG_CELL);
            ad.cov = S.ccov.as<unsigned char>();
            ad.phif = S.xf.as<double>();
            ad.ugf = h->ugf.as<double>();
            ad.sigf = S.sigf.as<double>();
            for (int d = 0; d < 3; ++d) ad.ff[d] = 0.25 / S.gf.dx[d];
            ad.vel = vel[0];
            ad.vel_ng = vel_ng;
            ad.gp = gphi ? gphi[0] : nullptr;
            ad.gp_ng = gphi_ng;
            amr_avgdown_kernel<<<h->lay_c.ntiles[REG_CELL], 256, 0, s>>>(ad);
        }
        if (phi) {
            if (phi[0])
                scatter_nodes_kernel<<<h->lay_c.ntiles[REG_NODE], 256, 0, s>>>(
                    kc, h->lay_c.dev_boxes(), h->lay_c.tiles(REG_NODE), S.xc.as<double>(), phi[0],
                    phi_ng, 0);
            if (two && phi[1])
                scatter_nodes_kernel<<<h->lay_f.ntiles[REG_NODE], 256, 0, s>>>(
                    kf, h->lay_f.dev_boxes(), h->lay_f.tiles(REG_NODE), S.xf.as<double>(), phi[1],
                    phi_ng, 0);
        }
        MGX_HIP(hipStreamSynchronize(s));
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "AMR nodal projector: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

// ---- AMR synchronisation ---------------------------------------------------

namespace {

// both levels' layouts (the fine boxes need not tile the domain), checked for
// ratio 2 and coarse-aligned fine boxes
void sync_layouts(const mgx_level_desc* c, const mgx_level_desc* f, BoxLayout& lc, BoxLayout& lf,
                  KGeom& kf) {
    MGX_REQUIRE(c && f && c->boxes && f->boxes && c->nboxes > 0 && f->nboxes > 0, MGX_ERR_ARG,
                "average down: bad level descriptions");
    double dxc[3], dxf[3];
    int per[3], zero[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        MGX_REQUIRE(f->n_cell[d] == 2 * c->n_cell[d], MGX_ERR_UNSUPPORTED,
                    "average down: refinement ratio 2 only");
        per[d] = c->is_periodic[d] ? 1 : 0;
        dxc[d] = (c->prob_hi[d] - c->prob_lo[d]) / c->n_cell[d];
        dxf[d] = (f->prob_hi[d] - f->prob_lo[d]) / f->n_cell[d];
    }
    for (int b = 0; b < f->nboxes; ++b)
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(f->boxes[6 * b + d] % 2 == 0 && f->boxes[6 * b + 3 + d] % 2 == 1,
                        MGX_ERR_ARG, "average down: fine boxes must align with coarse cells");
    const NodalGeom gc = make_nodal_geom(c->n_cell, per, zero, zero, dxc);
    const NodalGeom gf = make_nodal_geom(f->n_cell, per, zero, zero, dxf);
    lc.build(c->boxes, c->nboxes, gc);
    lf.build(f->boxes, f->nboxes, gf, true);
    kf = kgeom(gf);
}

}  // namespace

MGX_API int mgx_amr_average_down_faces(mgx_ctx ctx, const mgx_level_desc* crse,
                                       const mgx_level_desc* fine, const double* const* fine_faces,
                                       int fine_ng, double* const* crse_faces, int crse_ng,
                                       int ncomp) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(fine_faces && crse_faces && ncomp >= 1 && fine_ng >= 0 && crse_ng >= 0,
                    MGX_ERR_ARG, "average_down_faces: bad fields");
        BoxLayout lc, lf;
        KGeom kf;
        sync_layouts(crse, fine, lc, lf, kf);
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(fine_faces[d] && crse_faces[d], MGX_ERR_ARG, "average_down_faces: null face field");
            average_down_faces_kernel<<<lf.ntiles[REG_FACE0 + d], 256, 0, ctx->stream>>>(
                kf, lf.dev_boxes(), lf.tiles(REG_FACE0 + d), d, fine_faces[d], fine_ng,
                lc.dev_boxes(), lc.nboxes, crse_faces[d], crse_ng, ncomp);
        }
        MGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

MGX_API int mgx_amr_average_down(mgx_ctx ctx, const mgx_level_desc* crse,
                                 const mgx_level_desc* fine, const double* fine_cells, int fine_ng,
                                 double* crse_cells, int crse_ng, int ncomp) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(fine_cells && crse_cells && ncomp >= 1 && fine_ng >= 0 && crse_ng >= 0,
                    MGX_ERR_ARG, "average_down: bad fields");
        BoxLayout lc, lf;
        KGeom kf;
        sync_layouts(crse, fine, lc, lf, kf);
        average_down_cells_kernel<<<lf.ntiles[REG_CELL], 256, 0, ctx->stream>>>(
            kf, lf.dev_boxes(), lf.tiles(REG_CELL), fine_cells, fine_ng, lc.dev_boxes(), lc.nboxes,
            crse_cells, crse_ng, ncomp);
        MGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"

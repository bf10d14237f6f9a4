// amr_api.hip -- composite nodal projection over AMR levels (refinement ratio
// 2 between neighbours) and its C ABI.
//
// The reference builds one NodalProjector over all AMR levels
// (src/projection/incflo_apply_nodal_projection.cpp:124-215), projects the
// velocity of every level, returns phi / G phi per level and averages gp
// down afterwards (:258-266).  Here the composite problem is the conforming
// Q1 discretisation described in amr_kernels.h, solved by composite V-cycles
// from the finest level down: on each level above 0, pre-smoothing of its
// correction over its region (its boundary held at 0), the residual restricted
// to the level below, that level's correction (level 0: one V-cycle of its
// multigrid hierarchy over the whole domain), trilinear prolongation onto the
// level's nodes, post-smoothing and one composite Jacobi sweep of the level
// below's unknowns next to it; inside an FP64 outer iteration on the composite residual (MLMG's stopping
// rule ||r||_inf <= max(atol, rtol ||b||_inf)); MGX_MODE_MIXED runs the
// V-cycle in FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
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

// One level above level 0: dense canonical arrays over the whole domain at
// its resolution; only the tiles holding nodes of its region's closure are
// ever visited.
struct AmrFine {
    NodalGeom g;
    bool finer = false;                 // a still finer level lies above
    DevBuf ftype;                       // AMR_OUT / AMR_DOF / AMR_BDRY against the level below
    DevBuf dof;                         // composite unknowns: AMR_DOF nodes not surrounded by covered cells
    DevBuf ccov, rw;                    // (finer) cells under the finer level; restriction of its values (FP64-sized)
    DevBuf sig, sig_out, sig32, sig_out32;  // full / covered cells 0 (sig_out only when finer)
    DevBuf x, b, y, rT, ea, eb, yT, bC, rtmp;
    DevBuf wm;                          // its dF values for the restriction onto the level below (FP64, 0 elsewhere)
    TileSet tiles;
    const int4* blocks(int& n) { return tiles.blocks(n); }
    template <class T> const T* sigT() const {
        return std::is_same<T, float>::value ? (const T*)sig32.p : (const T*)sig.p;
    }
    template <class T> const T* sigoutT() const {
        if (!finer) return sigT<T>();
        return std::is_same<T, float>::value ? (const T*)sig_out32.p : (const T*)sig_out.p;
    }
    const double* sigout64() const { return finer ? sig_out.as<double>() : sig.as<double>(); }
};

// host masks of one level: every node's any / all of its adjacent cells set
static std::vector<unsigned char> cells_of_boxes(const NodalGeom& g, const std::vector<int>& boxes) {
    std::vector<unsigned char> fm((size_t)g.n[0] * g.n[1] * g.n[2], 0);
    for (size_t b = 0; b + 5 < boxes.size(); b += 6)
        for (int k = boxes[b + 2]; k <= boxes[b + 5]; ++k)
            for (int j = boxes[b + 1]; j <= boxes[b + 4]; ++j)
                for (int i = boxes[b + 0]; i <= boxes[b + 3]; ++i)
                    fm[i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k)] = 1;
    return fm;
}

class AmrNodal {
   public:
    NodalGeom gc;
    MGOpts opts;
    int nlev = 1;                       // multigrid levels of level 0
    NodalHierarchy<double> H64;
    NodalHierarchy<float> H32;
    DevBuf bC;                          // level-0 rhs of the coarse correction (FP64-sized)
    DevBuf sigc, sigc_out, sigc_out32;
    DevBuf cdof, ccov;
    DevBuf rw0;                         // restriction of level 1's values onto level 0 (FP64-sized)
    DevBuf xc, bc, rtmp, rcT;
    DevBuf scal, part, ibuf;
    std::vector<std::unique_ptr<AmrFine>> F;  // F[l - 1]: level l
    std::vector<void*> ecur;            // per level: the buffer holding its correction after a cycle
    unsigned long long* h_pinned = nullptr;
    double ndof = 0.0;
    bool sigma_set = false;
    bool h64_sigma = false, h32_sigma = false;
    int last_mixed = -1;                // precision of the last solve's V-cycle buffers

    ~AmrNodal() {
        if (h_pinned) (void)hipHostFree(h_pinned);
    }
    int L() const { return (int)F.size(); }
    AmrFine& lev(int l) { return *F[l - 1]; }

    // lboxes[l - 1]: level l's boxes (6 ints each, level-l cell indices)
    void setup(const NodalGeom& g0, const std::vector<std::vector<int>>& lboxes, const MGOpts& o) {
        // the boxes and the periodicity are fixed per handle, so a set-up for
        // new boundary types or options keeps the coefficient arrays
        const bool keep = sigma_set && (int)F.size() == (int)lboxes.size();
        gc = g0;
        opts = o;
        nlev = nodal_nlevels(gc, o.max_coarsening_level);
        const int L = (int)lboxes.size();
        if ((int)F.size() != L) {
            F.clear();
            for (int l = 1; l <= L; ++l) F.emplace_back(new AmrFine());
        }
        ecur.assign(L + 1, nullptr);
        std::vector<NodalGeom> G(L + 1);
        G[0] = gc;
        for (int l = 1; l <= L; ++l) G[l] = refine_geom(G[l - 1]);
        std::vector<std::vector<unsigned char>> fm(L + 1), cov(L + 1), ft(L + 1);
        fm[0].assign((size_t)gc.n[0] * gc.n[1] * gc.n[2], 1);
        for (int l = 1; l <= L; ++l) fm[l] = cells_of_boxes(G[l], lboxes[l - 1]);
        for (int l = 0; l <= L; ++l) {
            const NodalGeom& g = G[l];
            cov[l].assign((size_t)g.n[0] * g.n[1] * g.n[2], 0);
            if (l == L) continue;
            const NodalGeom& gf = G[l + 1];
            for (int k = 0; k < g.n[2]; ++k)
                for (int j = 0; j < g.n[1]; ++j)
                    for (int i = 0; i < g.n[0]; ++i) {
                        const size_t c = i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
                        cov[l][c] = fm[l + 1][2 * i + (size_t)gf.n[0] * (2 * j + (size_t)gf.n[1] * 2 * k)];
                        MGX_REQUIRE(!cov[l][c] || fm[l][c], MGX_ERR_ARG,
                                    "AMR nodal projector: a level reaches outside the level below it");
                    }
        }
        for (int l = 1; l <= L; ++l) {
            std::vector<unsigned char> anyf, allf;
            node_any_all(G[l], fm[l], anyf, allf);
            ft[l].resize(G[l].nnode());
            for (long q = 0; q < G[l].nnode(); ++q)
                ft[l][q] = anyf[q] ? (allf[q] ? AMR_DOF : AMR_BDRY) : AMR_OUT;
        }
        long count = 0;
        for (int l = 0; l <= L; ++l) {
            const NodalGeom& g = G[l];
            std::vector<unsigned char> anyc, allc;
            node_any_all(g, cov[l], anyc, allc);
            std::vector<unsigned char> dof(g.nnode());
            for (int k = 0; k < g.nn[2]; ++k)
                for (int j = 0; j < g.nn[1]; ++j)
                    for (int i = 0; i < g.nn[0]; ++i) {
                        const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
                        const bool in = l == 0 || ft[l][id] == AMR_DOF;
                        // proper nesting: a node touching cells under the
                        // finer level lies inside this level's region
                        MGX_REQUIRE(!(anyc[id] && !in), MGX_ERR_ARG,
                                    "AMR nodal projector: a level is not properly nested in the level below it");
                        dof[id] = in && !allc[id];
                        count += dof[id] && !is_dir_host(g, i, j, k);
                    }
            DevBuf& d = l == 0 ? cdof : lev(l).dof;
            d.alloc(dof.size());
            MGX_HIP(hipMemcpy(d.p, dof.data(), dof.size(), hipMemcpyHostToDevice));
            DevBuf& cv = l == 0 ? ccov : lev(l).ccov;
            if (l < L || l == 0) {
                cv.alloc(cov[l].size());
                MGX_HIP(hipMemcpy(cv.p, cov[l].data(), cov[l].size(), hipMemcpyHostToDevice));
            }
            if (l == L) continue;
            DevBuf& rb = l == 0 ? rw0 : lev(l).rw;
            rb.alloc(g.nnode() * sizeof(double));
        }
        ndof = (double)count;
        const long nnc = gc.nnode();
        sigc.alloc(gc.nsig() * sizeof(double));
        sigc_out.alloc(gc.nsig() * sizeof(double));
        sigc_out32.alloc(gc.nsig() * sizeof(float));
        // ghost cells outside non-periodic faces stay 0 (no element there)
        if (!keep) {
            MGX_HIP(hipMemset(sigc.p, 0, gc.nsig() * sizeof(double)));
            MGX_HIP(hipMemset(sigc_out.p, 0, gc.nsig() * sizeof(double)));
        }
        for (DevBuf* b : {&xc, &bc, &rtmp, &rcT}) b->alloc(nnc * sizeof(double));
        for (int l = 1; l <= L; ++l) {
            AmrFine& f = lev(l);
            f.g = G[l];
            f.finer = l < L;
            f.ftype.alloc(ft[l].size());
            MGX_HIP(hipMemcpy(f.ftype.p, ft[l].data(), ft[l].size(), hipMemcpyHostToDevice));
            {
                std::vector<unsigned char> act(ft[l].size());
                for (size_t q = 0; q < act.size(); ++q) act[q] = ft[l][q] != AMR_OUT;
                f.tiles = TileSet();
                f.tiles.build(f.g, act);
            }
            const long nn = f.g.nnode(), ns = f.g.nsig();
            f.sig.alloc(ns * sizeof(double));
            f.sig32.alloc(ns * sizeof(float));
            if (!keep) MGX_HIP(hipMemset(f.sig.p, 0, ns * sizeof(double)));
            if (f.finer) {
                f.sig_out.alloc(ns * sizeof(double));
                f.sig_out32.alloc(ns * sizeof(float));
                if (!keep) MGX_HIP(hipMemset(f.sig_out.p, 0, ns * sizeof(double)));
            }
            // the launches visit only the tiles that hold nodes of the level's
            // closure; what lies elsewhere in these arrays is never read but is
            // multiplied by the zero sigma outside, so it starts finite (0)
            std::vector<DevBuf*> bufs = {&f.x, &f.b, &f.y, &f.rT, &f.ea, &f.eb, &f.yT, &f.wm};
            if (f.finer) { bufs.push_back(&f.bC); bufs.push_back(&f.rtmp); }
            for (DevBuf* b : bufs) {
                b->alloc(nn * sizeof(double));
                MGX_HIP(hipMemset(b->p, 0, nn * sizeof(double)));
            }
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

    // sigc / level sig (ghosted FP64) are filled by the caller
    void sigma_changed(hipStream_t s) {
        const KGeom kc = kgeom(gc);
        amr_sigma_out_kernel<<<grid3(gc.n[0], gc.n[1], gc.n[2]), block3(), 0, s>>>(
            kc, ccov.as<unsigned char>(), sigc.as<double>(), sigc_out.as<double>());
        cast_kernel<float, double><<<grid1d(gc.nsig()), 256, 0, s>>>(sigc_out.as<double>(),
                                                                    sigc_out32.as<float>(), gc.nsig());
        for (int l = 1; l <= L(); ++l) {
            AmrFine& f = lev(l);
            cast_kernel<float, double><<<grid1d(f.g.nsig()), 256, 0, s>>>(f.sig.as<double>(),
                                                                         f.sig32.as<float>(), f.g.nsig());
            if (!f.finer) continue;
            amr_sigma_out_kernel<<<grid3(f.g.n[0], f.g.n[1], f.g.n[2]), block3(), 0, s>>>(
                kgeom(f.g), f.ccov.as<unsigned char>(), f.sig.as<double>(), f.sig_out.as<double>());
            cast_kernel<float, double><<<grid1d(f.g.nsig()), 256, 0, s>>>(
                f.sig_out.as<double>(), f.sig_out32.as<float>(), f.g.nsig());
        }
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

    // the unknown mask of level l >= 1 (the finest level: its AMR_DOF nodes)
    const unsigned char* umask(int l, unsigned char& want) {
        AmrFine& f = lev(l);
        want = f.finer ? 1 : AMR_DOF;
        return f.finer ? f.dof.as<unsigned char>() : f.ftype.as<unsigned char>();
    }

    // b_l -= mu / 8^(L-l) on the unknowns, mu the mean of the right-hand side
    // over all unknowns in the symmetric scaling (singular problems)
    void remove_rhs_mean(hipStream_t s) {
        const int np = 512;
        amr_sum_kernel<<<np, 256, 0, s>>>(bc.as<double>(), cdof.as<unsigned char>(), 1, gc.nnode(),
                                          part.as<double>());
        for (int l = 1; l <= L(); ++l) {
            unsigned char want;
            const unsigned char* m = umask(l, want);
            amr_sum_kernel<<<np, 256, 0, s>>>(lev(l).b.as<double>(), m, want, lev(l).g.nnode(),
                                              part.as<double>() + (long)l * np);
        }
        amr_mean_levels_kernel<<<1, 64, 0, s>>>(part.as<double>(), np, L() + 1, ndof,
                                                scal.as<double>() + 40);
        amr_shift_kernel<<<grid1d(gc.nnode()), 256, 0, s>>>(bc.as<double>(), cdof.as<unsigned char>(),
                                                            1, 0, gc.nnode(), scal.as<double>() + 40,
                                                            std::pow(0.125, L()));
        for (int l = 1; l <= L(); ++l) {
            unsigned char want;
            const unsigned char* m = umask(l, want);
            amr_shift_kernel<<<grid1d(lev(l).g.nnode()), 256, 0, s>>>(
                lev(l).b.as<double>(), m, want, 0, lev(l).g.nnode(), scal.as<double>() + 40,
                std::pow(0.125, L() - l));
        }
    }

    // level l >= 1's closure: dF_l from level l - 1 (the levels filled upward)
    void fill_levels(hipStream_t s) {
        for (int l = 1; l <= L(); ++l) {
            AmrFine& f = lev(l);
            int nbk = 0;
            const int4* tl = f.blocks(nbk);
            const KGeom kcl = kgeom(l == 1 ? gc : lev(l - 1).g);
            const double* xcl = l == 1 ? xc.as<double>() : lev(l - 1).x.as<double>();
            if (nbk)
                amr_fill_kernel<double><<<nbk, tile_block(), 0, s>>>(
                    kgeom(f.g), kcl, f.ftype.as<unsigned char>(), xcl, f.x.as<double>(), tl);
        }
    }

    // composite residual of the levels' x into rT (levels >= 1) / rcT (level
    // 0), type TO; returns ||r||_inf
    template <class TO>
    double residual(hipStream_t s) {
        fill_levels(s);
        zero_norm(1, s);
        const double* yB = nullptr;  // level l+1's contribution to restrict, with sign sB
        double sB = 0.0;
        for (int l = L(); l >= 1; --l) {
            AmrFine& f = lev(l);
            const KGeom kf = kgeom(f.g);
            int nbk = 0;
            const int4* tl = f.blocks(nbk);
            if (!f.finer) {
                SArgs<double, double> a{};
                a.g = kf;
                a.c = make_coef<double>(f.g.dx);
                a.x = f.x.as<double>();
                a.sig = f.sig.as<double>();
                a.rout = f.y.as<double>();
                launch_stencil<double, double, KS_APPLY>(f.g, a, s, false, &f.tiles);
                if (nbk)
                    amr_fine_resid_kernel<TO><<<nbk, tile_block(), 0, s>>>(
                        kf, f.ftype.as<unsigned char>(), f.b.as<double>(), f.y.as<double>(), f.rT.as<TO>(),
                        scal.as<unsigned long long>() + 1, tl, f.wm.as<double>());
                yB = f.wm.as<double>();  // L x on dF, 0 elsewhere
                sB = -1.0;
            } else {
                AmrFine& u = lev(l + 1);
                SArgs<double, double> a{};
                a.g = kf;
                a.c = make_coef<double>(f.g.dx);
                a.x = f.x.as<double>();
                a.b = f.b.as<double>();
                a.sig = f.sig_out.as<double>();
                a.rout = f.rtmp.as<double>();
                launch_stencil<double, double, KS_RESID>(f.g, a, s, false, &f.tiles);
                restrict_full<double>(u.g, f.g, yB, f.rw.as<double>(), s);
                if (nbk)
                    amr_coarse_combine_kernel<TO><<<nbk, tile_block(), 0, s>>>(
                        kf, f.dof.as<unsigned char>(), f.rtmp.as<double>(), f.rw.as<double>(), sB,
                        f.rT.as<TO>(), scal.as<unsigned long long>() + 1, tl, f.ftype.as<unsigned char>(),
                        f.wm.as<double>());
                // level l's L x on dF_l is -(b - L x) there (b is 0 off the unknowns)
                yB = f.wm.as<double>();
                sB = 1.0;
            }
        }
        const KGeom kc = kgeom(gc);
        SArgs<double, double> c{};
        c.g = kc;
        c.c = make_coef<double>(gc.dx);
        c.x = xc.as<double>();
        c.b = bc.as<double>();
        c.sig = sigc_out.as<double>();
        c.rout = rtmp.as<double>();
        launch_stencil<double, double, KS_RESID>(gc, c, s, false);
        if (L() > 0) restrict_full<double>(lev(1).g, gc, yB, rw0.as<double>(), s);
        amr_coarse_combine_kernel<TO><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kc, cdof.as<unsigned char>(), rtmp.as<double>(), L() > 0 ? rw0.as<double>() : nullptr, sB,
            rcT.as<TO>(), scal.as<unsigned long long>() + 1);
        return read_norm(1, s);
    }

    // the correction of level l >= 1 from the right-hand side rhs: pre-smooth,
    // the level below's correction (recursively; level 0: one V-cycle of its
    // multigrid hierarchy), prolongation, post-smooth, and one composite
    // Jacobi sweep of the level below's unknowns next to this level
    template <class T>
    void cycle_level(NodalHierarchy<T>& H, int l, const T* rhs, hipStream_t s) {
        AmrFine& f = lev(l);
        const KGeom kf = kgeom(f.g);
        const unsigned char* ft = f.ftype.as<unsigned char>();
        T* e = f.ea.as<T>();
        T* e2 = f.eb.as<T>();
        SArgs<T, T> a{};
        a.g = kf;
        a.c = make_coef<T>(f.g.dx);
        a.b = rhs;
        a.sig = f.sigT<T>();
        // weighted Jacobi (the Chebyshev factors of the level-0 hierarchy
        // measured one composite cycle more at the two-level 128^3 / 256^3 case)
        a.omega = (T)nodal_omega(f.g, opts.omega);
        a.nmask = ft;
        int nbk = 0;
        const int4* tl = f.blocks(nbk);
        if (opts.nu1 > 0) {
            a.xout = e;
            launch_stencil<T, T, KS_JFIRST>(f.g, a, s, false, &f.tiles);
            for (int q = 1; q < opts.nu1; ++q) {
                a.x = e;
                a.xout = e2;
                launch_stencil<T, T, KS_JACOBI>(f.g, a, s, false, &f.tiles);
                std::swap(e, e2);
            }
        } else {
            zero_kernel<T><<<grid1d(f.g.nnode()), 256, 0, s>>>(e, f.g.nnode());
        }
        // rhs - L e on the level's closure (0 from rhs on dF, and from L e
        // away from the closure), restricted to the level below
        SArgs<T, T> ap{};
        ap.g = kf;
        ap.c = a.c;
        ap.x = e;
        ap.b = rhs;
        ap.sig = f.sigT<T>();
        ap.rout = f.yT.as<T>();
        launch_stencil<T, T, KS_RESID>(f.g, ap, s, false, &f.tiles);
        T* ec;
        KGeom kcl;
        T* w = l == 1 ? rw0.as<T>() : lev(l - 1).rw.as<T>();
        const NodalGeom& gcl = l == 1 ? gc : lev(l - 1).g;
        restrict_full<T>(f.g, gcl, f.yT.as<T>(), w, s);
        if (l == 1) {
            kcl = kgeom(gc);
            amr_coarse_rhs_kernel<T><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
                kcl, cdof.as<unsigned char>(), rcT.as<T>(), w, H.L[0].b);
            H.vcycle(opts, s, ibuf.as<int>());
            ec = H.L[0].x;
        } else {
            AmrFine& c = lev(l - 1);
            kcl = kgeom(c.g);
            int nbc = 0;
            const int4* tlc = c.blocks(nbc);
            if (nbc)
                amr_coarse_rhs_kernel<T><<<nbc, tile_block(), 0, s>>>(
                    kcl, c.dof.as<unsigned char>(), c.rT.as<T>(), w, c.bC.as<T>(), tlc);
            cycle_level<T>(H, l - 1, c.bC.as<T>(), s);
            ec = (T*)ecur[l - 1];
        }
        if (nbk) amr_prolong_kernel<T><<<nbk, tile_block(), 0, s>>>(kf, kcl, ft, ec, e, tl);
        for (int q = 0; q < opts.nu2; ++q) {
            a.x = e;
            a.xout = e2;
            launch_stencil<T, T, KS_JACOBI>(f.g, a, s, false, &f.tiles);
            std::swap(e, e2);
        }
        {
            // the unknowns of the level below next to this level: one Jacobi
            // sweep with the composite operator (e on dF follows the
            // correction below)
            ap.x = e;
            ap.b = nullptr;
            launch_stencil<T, T, KS_APPLY>(f.g, ap, s, false, &f.tiles);
            if (nbk)
                amr_mask_bdry_kernel<T><<<nbk, tile_block(), 0, s>>>(kf, ft, f.yT.as<T>(), f.yT.as<T>(), tl);
            restrict_full<T>(f.g, gcl, f.yT.as<T>(), w, s);
            if (l == 1) {
                SArgs<T, T> ac{};
                ac.g = kcl;
                ac.c = make_coef<T>(gc.dx);
                ac.x = ec;
                ac.sig = std::is_same<T, float>::value ? (const T*)sigc_out32.p : (const T*)sigc_out.p;
                ac.rout = rtmp.as<T>();
                launch_stencil<T, T, KS_APPLY>(gc, ac, s, false);
                amr_coarse_jacobi_kernel<T><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
                    kcl, cdof.as<unsigned char>(), rcT.as<T>(), rtmp.as<T>(), w, H.L[0].sig,
                    make_coef<double>(gc.dx).diag4, nodal_omega(gc, opts.omega), ec);
            } else {
                AmrFine& c = lev(l - 1);
                int nbc = 0;
                const int4* tlc = c.blocks(nbc);
                SArgs<T, T> ac{};
                ac.g = kcl;
                ac.c = make_coef<T>(c.g.dx);
                ac.x = ec;
                ac.sig = c.sigoutT<T>();
                ac.rout = c.yT.as<T>();
                launch_stencil<T, T, KS_APPLY>(c.g, ac, s, false, &c.tiles);
                if (nbc)
                    amr_coarse_jacobi_kernel<T><<<nbc, tile_block(), 0, s>>>(
                        kcl, c.dof.as<unsigned char>(), c.rT.as<T>(), c.yT.as<T>(), w, c.sigT<T>(),
                        make_coef<double>(c.g.dx).diag4, nodal_omega(c.g, opts.omega), ec, tlc);
            }
        }
        ecur[l] = e;
    }

    // one composite V-cycle on the residual in rT / rcT; the corrections are
    // added to the levels' x
    template <class T>
    void vcycle(NodalHierarchy<T>& H, hipStream_t s) {
        if (L() == 0) {
            H.L[0].b = rcT.as<T>();
            H.vcycle(opts, s, ibuf.as<int>());
            H.L[0].b = bC.as<T>();
        } else {
            cycle_level<T>(H, L(), lev(L()).rT.as<T>(), s);
        }
        for (int l = 1; l <= L(); ++l) {
            AmrFine& f = lev(l);
            int nbk = 0;
            const int4* tl = f.blocks(nbk);
            unsigned char want;
            const unsigned char* m = umask(l, want);
            if (nbk)
                amr_add_kernel<T><<<nbk, tile_block(), 0, s>>>(kgeom(f.g), m, want, (const T*)ecur[l],
                                                               f.x.as<double>(), tl);
        }
        amr_add_kernel<T><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kgeom(gc), cdof.as<unsigned char>(), 1, H.L[0].x, xc.as<double>());
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

    // composite solve of the right-hand side in bc / the levels' b; result
    // in xc / the levels' x with every closure filled, nodes under a finer
    // level injected from it and (singular) the mean of the level-0 phi
    // removed
    void solve(double rtol, double atol, int mode, mgx_solve_info* info, hipStream_t s) {
        MGX_REQUIRE(sigma_set, MGX_ERR_STATE, "AMR nodal projector: sigma not set");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_UNSUPPORTED,
                    "AMR nodal projector: FP64 and mixed modes only");
        const bool mixed = mode == MGX_MODE_MIXED;
        const KGeom kc = kgeom(gc);
        if (mixed) prepare(H32, h32_sigma, s);
        else prepare(H64, h64_sigma, s);
        if (last_mixed != (int)mixed) {
            // the V-cycle's buffers change element type: parts of the levels
            // that no tile holds must read 0 again (their coefficient is 0, but
            // reinterpreted bytes of the other type need not be finite)
            for (auto& f : F)
                for (DevBuf* b : {&f->rT, &f->ea, &f->eb, &f->yT, &f->bC})
                    if (b->p) MGX_HIP(hipMemsetAsync(b->p, 0, f->g.nnode() * sizeof(double), s));
            last_mixed = (int)mixed;
        }
        MGX_HIP(hipMemsetAsync(ibuf.p, 0, sizeof(int), s));
        if (gc.singular()) remove_rhs_mean(s);
        // ||b||_inf over the unknowns of every level
        zero_norm(0, s);
        for (int l = L(); l >= 1; --l) {
            AmrFine& f = lev(l);
            int nbk = 0;
            const int4* tl = f.blocks(nbk);
            if (!nbk) continue;
            if (!f.finer)
                amr_fine_resid_kernel<double><<<nbk, tile_block(), 0, s>>>(
                    kgeom(f.g), f.ftype.as<unsigned char>(), f.b.as<double>(), nullptr, f.y.as<double>(),
                    scal.as<unsigned long long>(), tl);
            else
                amr_coarse_combine_kernel<double><<<nbk, tile_block(), 0, s>>>(
                    kgeom(f.g), f.dof.as<unsigned char>(), f.b.as<double>(), nullptr, 0.0, f.rtmp.as<double>(),
                    scal.as<unsigned long long>(), tl);
        }
        amr_coarse_combine_kernel<double><<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
            kc, cdof.as<unsigned char>(), bc.as<double>(), nullptr, 0.0, rtmp.as<double>(),
            scal.as<unsigned long long>());
        const double bnorm = read_norm(0, s);
        MGX_HIP(hipMemsetAsync(xc.p, 0, gc.nnode() * sizeof(double), s));
        for (auto& f : F) MGX_HIP(hipMemsetAsync(f->x.p, 0, f->g.nnode() * sizeof(double), s));
        const double tol = std::max(atol, rtol * bnorm);
        const double r0 = mixed ? residual<float>(s) : residual<double>(s);
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
                fprintf(stderr, "mgx AMR nodal MLMG: iter %d |r| %.6e rel %.3e\n", it, rn,
                        bnorm > 0 ? rn / bnorm : rn);
            conv = rn <= tol;
        }
        // closures from the level below, nodes under a finer level from it
        fill_levels(s);
        for (int l = L() - 1; l >= 0; --l) {
            AmrFine& u = lev(l + 1);
            if (l == 0) {
                amr_inject_kernel<<<grid3(gc.nn[0], gc.nn[1], gc.nn[2]), block3(), 0, s>>>(
                    kgeom(u.g), kc, cdof.as<unsigned char>(), u.x.as<double>(), xc.as<double>());
            } else {
                AmrFine& f = lev(l);
                int nbk = 0;
                const int4* tl = f.blocks(nbk);
                if (nbk)
                    amr_inject_kernel<<<nbk, tile_block(), 0, s>>>(
                        kgeom(u.g), kgeom(f.g), f.dof.as<unsigned char>(), u.x.as<double>(), f.x.as<double>(),
                        f.ftype.as<unsigned char>(), tl);
            }
        }
        if (gc.singular()) {
            const int np = 512;
            amr_sum_kernel<<<np, 256, 0, s>>>(xc.as<double>(), nullptr, 0, gc.nnode(), part.as<double>());
            amr_mean_kernel<<<1, 64, 0, s>>>(part.as<double>(), np, 1.0, nullptr, 0, 0.0,
                                             (double)gc.nnode(), scal.as<double>() + 41);
            amr_shift_kernel<<<grid1d(gc.nnode()), 256, 0, s>>>(xc.as<double>(), nullptr, 0, 0,
                                                                gc.nnode(), scal.as<double>() + 41, 1.0);
            for (auto& f : F)
                amr_shift_kernel<<<grid1d(f->g.nnode()), 256, 0, s>>>(
                    f->x.as<double>(), f->ftype.as<unsigned char>(), AMR_OUT, 1, f->g.nnode(),
                    scal.as<double>() + 41, 1.0);
        }
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

struct mgx_amr_s {
    mgx_ctx_s* ctx = nullptr;
    int n_cell[3] = {0, 0, 0};
    double prob_lo[3] = {0, 0, 0}, prob_hi[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    std::vector<int> cboxes;
    std::vector<std::vector<int>> lboxes;   // levels 1.. (6 ints per box)
    int bclo[3] = {0, 0, 0}, bchi[3] = {0, 0, 0};
    bool bc_set = false;
    MGOpts opts;
    bool dirty = true;
    AmrNodal solver;
    std::vector<std::unique_ptr<BoxLayout>> lay;   // per level (0: coarse)
    std::vector<std::unique_ptr<DevBuf>> ug, gcan; // per level: canonical ghosted u* / u_new, G phi
    hipStream_t stream() const { return ctx->stream; }
    int nfine() const { return (int)lboxes.size(); }

    void ensure_setup() {
        if (!dirty) return;
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(periodic[d] || bc_set, MGX_ERR_STATE,
                        "AMR nodal projector: setDomainBC is required for non-periodic domains");
        // fine levels stay off inflow faces (their ghost velocity there is not
        // part of the composite divergence)
        for (int l = 1; l <= nfine(); ++l) {
            const std::vector<int>& bx = lboxes[l - 1];
            const int nf[3] = {n_cell[0] << l, n_cell[1] << l, n_cell[2] << l};
            for (size_t b = 0; b + 5 < bx.size(); b += 6)
                for (int d = 0; d < 3; ++d) {
                    if (periodic[d]) continue;
                    MGX_REQUIRE(!(bx[b + d] == 0 && bclo[d] == MGX_BC_INFLOW) &&
                                    !(bx[b + 3 + d] == nf[d] - 1 && bchi[d] == MGX_BC_INFLOW),
                                MGX_ERR_UNSUPPORTED, "AMR nodal projector: fine boxes must stay off inflow faces");
                }
        }
        double dx[3];
        for (int d = 0; d < 3; ++d) dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
        const NodalGeom g = make_nodal_geom(n_cell, periodic, bclo, bchi, dx);
        solver.setup(g, lboxes, opts);
        const int L = nfine();
        lay.clear();
        ug.clear();
        gcan.clear();
        for (int l = 0; l <= L; ++l) {
            const NodalGeom& gl = l == 0 ? g : solver.lev(l).g;
            lay.emplace_back(new BoxLayout());
            if (l == 0) lay[0]->build(cboxes.data(), (int)cboxes.size() / 6, gl);
            else lay[l]->build(lboxes[l - 1].data(), (int)lboxes[l - 1].size() / 6, gl, true);
            ug.emplace_back(new DevBuf());
            ug[l]->alloc(3 * gl.nsig() * sizeof(double));
            MGX_HIP(hipMemset(ug[l]->p, 0, 3 * gl.nsig() * sizeof(double)));
            gcan.emplace_back(new DevBuf());
            if (l > 0) gcan[l]->alloc(3 * gl.nsig() * sizeof(double));
        }
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
        MGX_REQUIRE(levels != nullptr && nlevels >= 1 && nlevels <= MGX_AMR_MAX_LEVELS, MGX_ERR_UNSUPPORTED,
                    "AMR nodal projector: 1 to MGX_AMR_MAX_LEVELS levels (ratio 2)");
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
        for (int l = 1; l < nlevels; ++l) {
            const mgx_level_desc& p = levels[l - 1];
            const mgx_level_desc& f = levels[l];
            for (int d = 0; d < 3; ++d) {
                MGX_REQUIRE(f.n_cell[d] == 2 * p.n_cell[d], MGX_ERR_UNSUPPORTED,
                            "AMR nodal projector: refinement ratio 2 only");
                MGX_REQUIRE(f.prob_lo[d] == c.prob_lo[d] && f.prob_hi[d] == c.prob_hi[d] &&
                                (f.is_periodic[d] != 0) == (c.is_periodic[d] != 0),
                            MGX_ERR_ARG, "AMR nodal projector: levels must share the domain");
            }
            MGX_REQUIRE(f.nboxes >= 0 && (f.nboxes == 0 || f.boxes), MGX_ERR_ARG,
                        "AMR nodal projector: bad fine box list");
            MGX_REQUIRE(f.nboxes > 0 || l == nlevels - 1, MGX_ERR_ARG,
                        "AMR nodal projector: only the finest level may be empty");
            for (int b = 0; b < f.nboxes; ++b)
                for (int d = 0; d < 3; ++d) {
                    const int lo = f.boxes[6 * b + d], hi = f.boxes[6 * b + 3 + d];
                    MGX_REQUIRE(lo >= 0 && hi < f.n_cell[d] && lo <= hi, MGX_ERR_ARG,
                                "AMR nodal projector: fine box outside the domain");
                    MGX_REQUIRE(lo % 2 == 0 && hi % 2 == 1, MGX_ERR_ARG,
                                "AMR nodal projector: fine boxes must align with coarse cells");
                }
            if (f.nboxes > 0) h->lboxes.emplace_back(f.boxes, f.boxes + 6 * f.nboxes);
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
        for (int l = 0; l <= h->nfine(); ++l)
            MGX_REQUIRE(sigma[l] != nullptr, MGX_ERR_ARG, "AMR nodal projector: sigma needed on every level");
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "sigma ghost width out of range");
        h->ensure_setup();
        AmrNodal& S = h->solver;
        gather_sigma_kernel<<<h->lay[0]->ntiles[REG_CELL], 256, 0, h->stream()>>>(
            kgeom(S.gc), h->lay[0]->dev_boxes(), h->lay[0]->tiles(REG_CELL), sigma[0], ng,
            S.sigc.as<double>());
        for (int l = 1; l <= h->nfine(); ++l) {
            AmrFine& f = S.lev(l);
            MGX_HIP(hipMemsetAsync(f.sig.p, 0, f.g.nsig() * sizeof(double), h->stream()));
            gather_sigma_kernel<<<h->lay[l]->ntiles[REG_CELL], 256, 0, h->stream()>>>(
                kgeom(f.g), h->lay[l]->dev_boxes(), h->lay[l]->tiles(REG_CELL), sigma[l], ng,
                f.sig.as<double>());
        }
        S.sigma_changed(h->stream());
    });
}

MGX_API int mgx_amr_nodal_project(mgx_amr_nodal h, double* const* vel, int vel_ng,
                                  double* const* phi, int phi_ng, double* const* gphi, int gphi_ng,
                                  double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h || !vel) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        const int L = h->nfine();
        for (int l = 0; l <= L; ++l)
            MGX_REQUIRE(vel[l] != nullptr, MGX_ERR_ARG, "AMR nodal projector: velocity needed on every level");
        const bool any_nonper = !h->periodic[0] || !h->periodic[1] || !h->periodic[2];
        MGX_REQUIRE(vel_ng >= (any_nonper ? 1 : 0) && vel_ng <= 8, MGX_ERR_ARG,
                    "velocity needs one ghost cell at non-periodic domain faces");
        MGX_REQUIRE(phi_ng >= 0 && gphi_ng >= 0, MGX_ERR_ARG, "negative ghost width");
        h->ensure_setup();
        AmrNodal& S = h->solver;
        MGX_REQUIRE(S.sigma_set, MGX_ERR_STATE, "AMR nodal projector: sigma not set");
        hipStream_t s = h->stream();
        auto geom = [&](int l) -> const NodalGeom& { return l == 0 ? S.gc : S.lev(l).g; };
        // u* per level (the gather writes the same cells every call; the rest
        // of a fine level's array stays 0 from the set-up); cells under a
        // finer level carry no velocity of their own
        for (int l = 0; l <= L; ++l) {
            const KGeom kl = kgeom(geom(l));
            GatherVelArgs a{};
            a.g = kl;
            a.boxes = h->lay[l]->dev_boxes();
            a.tiles = h->lay[l]->tiles(REG_CELL_GROWN);
            a.vel = vel[l];
            a.vel_ng = vel_ng;
            a.scale = 1.0;
            a.const_rho = 1.0;
            a.ug = h->ug[l]->as<double>();
            a.ghost_rule = 0;
            gather_vel_kernel<<<h->lay[l]->ntiles[REG_CELL_GROWN], 256, 0, s>>>(a);
            if (l < L) {
                const NodalGeom& g = geom(l);
                amr_zero_covered_kernel<<<grid3(g.n[0], g.n[1], g.n[2]), block3(), 0, s>>>(
                    kl, l == 0 ? S.ccov.as<unsigned char>() : S.lev(l).ccov.as<unsigned char>(),
                    h->ug[l]->as<double>());
            }
        }
        // composite divergence: each level's nodal divergence, dF_{l+1}'s
        // restricted onto level l's unknowns
        for (int l = L; l >= 1; --l) {
            AmrFine& f = S.lev(l);
            const KGeom kf = kgeom(f.g);
            int nbk = 0;
            const int4* tl = f.blocks(nbk);
            if (!nbk) continue;
            nodal_rhs_kernel<<<nbk, tile_block(), 0, s>>>(kf, h->ug[l]->as<double>(), f.y.as<double>(),
                                                          0.25 / f.g.dx[0], 0.25 / f.g.dx[1],
                                                          0.25 / f.g.dx[2], tl);
            if (!f.finer) {
                amr_fine_resid_kernel<double><<<nbk, tile_block(), 0, s>>>(
                    kf, f.ftype.as<unsigned char>(), f.y.as<double>(), nullptr, f.b.as<double>(), nullptr, tl);
                amr_mask_bdry_kernel<double><<<nbk, tile_block(), 0, s>>>(kf, f.ftype.as<unsigned char>(),
                                                                          f.y.as<double>(), f.wm.as<double>(), tl);
            } else {
                AmrFine& u = S.lev(l + 1);
                restrict_full<double>(u.g, f.g, u.wm.as<double>(), f.rw.as<double>(), s);
                amr_coarse_combine_kernel<double><<<nbk, tile_block(), 0, s>>>(
                    kf, f.dof.as<unsigned char>(), f.y.as<double>(), f.rw.as<double>(), 1.0, f.b.as<double>(),
                    nullptr, tl, f.ftype.as<unsigned char>(), f.wm.as<double>());
            }
        }
        {
            const KGeom kc = kgeom(S.gc);
            nodal_rhs_kernel<<<grid3(S.gc.nn[0], S.gc.nn[1], S.gc.nn[2]), block3(), 0, s>>>(
                kc, h->ug[0]->as<double>(), S.rtmp.as<double>(), 0.25 / S.gc.dx[0], 0.25 / S.gc.dx[1],
                0.25 / S.gc.dx[2]);
            if (L > 0) restrict_full<double>(S.lev(1).g, S.gc, S.lev(1).wm.as<double>(), S.rw0.as<double>(), s);
            amr_coarse_combine_kernel<double><<<grid3(S.gc.nn[0], S.gc.nn[1], S.gc.nn[2]), block3(), 0, s>>>(
                kc, S.cdof.as<unsigned char>(), S.rtmp.as<double>(), L > 0 ? S.rw0.as<double>() : nullptr, 1.0,
                S.bc.as<double>(), nullptr);
        }
        S.solve(rtol, atol, mode, info, s);
        // u -= sigma G phi on every level, G phi out; from the finest level
        // down, cells under a finer level take the mean of their children
        for (int l = L; l >= 0; --l) {
            const NodalGeom& g = geom(l);
            UpdateArgs u{};
            u.g = kgeom(g);
            u.boxes = h->lay[l]->dev_boxes();
            u.tiles = h->lay[l]->tiles(REG_CELL);
            u.phi = l == 0 ? S.xc.as<double>() : S.lev(l).x.as<double>();
            u.ug = h->ug[l]->as<double>();
            u.sig = l == 0 ? S.sigc.as<double>() : S.lev(l).sig.as<double>();
            for (int d = 0; d < 3; ++d) u.f[d] = 0.25 / g.dx[d];
            u.vel = vel[l];
            u.vel_ng = vel_ng;
            u.gp = gphi ? gphi[l] : nullptr;
            u.gp_ng = gphi_ng;
            if (l > 0) {
                u.ucan = h->ug[l]->as<double>();
                u.gcan = gphi && gphi[l] ? h->gcan[l]->as<double>() : nullptr;
            }
            update_cells_kernel<<<h->lay[l]->ntiles[REG_CELL], 256, 0, s>>>(u);
            if (l < L) {
                AvgDownArgs ad{};
                ad.gc = kgeom(g);
                ad.gf = kgeom(geom(l + 1));
                ad.boxes = h->lay[l]->dev_boxes();
                ad.tiles = h->lay[l]->tiles(REG_CELL);
                ad.cov = l == 0 ? S.ccov.as<unsigned char>() : S.lev(l).ccov.as<unsigned char>();
                ad.uf = h->ug[l + 1]->as<double>();
                ad.gf_ = gphi && gphi[l + 1] ? h->gcan[l + 1]->as<double>() : nullptr;
                ad.vel = vel[l];
                ad.vel_ng = vel_ng;
                ad.gp = gphi ? gphi[l] : nullptr;
                ad.gp_ng = gphi_ng;
                if (l > 0) {
                    ad.uc = h->ug[l]->as<double>();
                    ad.gc_ = gphi && gphi[l] ? h->gcan[l]->as<double>() : nullptr;
                }
                amr_avgdown_kernel<<<h->lay[l]->ntiles[REG_CELL], 256, 0, s>>>(ad);
            }
        }
        if (phi)
            for (int l = 0; l <= L; ++l)
                if (phi[l])
                    scatter_nodes_kernel<<<h->lay[l]->ntiles[REG_NODE], 256, 0, s>>>(
                        kgeom(geom(l)), h->lay[l]->dev_boxes(), h->lay[l]->tiles(REG_NODE),
                        l == 0 ? S.xc.as<double>() : S.lev(l).x.as<double>(), phi[l], phi_ng, 0);
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

// both levels' layouts (neither needs to tile the domain), checked for ratio 2
// and coarse-aligned fine boxes
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
    // the coarse level may itself be a fine level (any pair of neighbouring
    // levels); a fine cell whose parent no coarse box holds is skipped
    lc.build(c->boxes, c->nboxes, gc, true);
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

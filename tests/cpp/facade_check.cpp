// facade_check.cpp -- drives the C++ facade (include/mgx_amrex.H) the way
// incflo drives AMReX-Hydro's projectors, on inputs written by
// tests/test_gpu_facade.py, and writes the results back for comparison.
//
//   facade_check <dir> <n> <max_grid_size>
// reads  <dir>/vel.bin (boxed cell, 3 comps, 1 ghost), <dir>/sigma.bin (boxed cell),
//        <dir>/umac{0,1,2}.bin (boxed faces, no ghost)
// writes <dir>/vel_out.bin, phi.bin (boxed nodal), gphi.bin (boxed cell, 3 comps),
//        umac_out{0,1,2}.bin, macphi.bin (boxed cell)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mgx_amrex.H"

static std::vector<double> read_file(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) mgx::Abort("cannot open " + p);
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f) / (long)sizeof(double);
    std::fseek(f, 0, SEEK_SET);
    std::vector<double> v(n);
    if (std::fread(v.data(), sizeof(double), n, f) != (size_t)n) mgx::Abort("short read " + p);
    std::fclose(f);
    return v;
}

static void write_file(const std::string& p, const std::vector<double>& v) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) mgx::Abort("cannot write " + p);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

struct Dev {
    mgx::Context& c;
    void* p = nullptr;
    size_t n = 0;
    Dev(mgx::Context& ctx, const std::vector<double>& h) : c(ctx), n(h.size()) {
        c.check(mgx_dev_malloc(c.get(), n * sizeof(double), &p), "malloc");
        c.check(mgx_memcpy_h2d(c.get(), p, h.data(), n * sizeof(double)), "h2d");
    }
    Dev(mgx::Context& ctx, size_t count) : c(ctx), n(count) {
        c.check(mgx_dev_malloc(c.get(), n * sizeof(double), &p), "malloc");
        // zero, like a fresh MultiFab set to 0 (phi doubles as the initial guess)
        c.check(mgx_memset_d(c.get(), p, 0, n * sizeof(double)), "memset");
    }
    ~Dev() { mgx_dev_free(c.get(), p); }
    std::vector<double> get() const {
        std::vector<double> h(n);
        c.check(mgx_memcpy_d2h(c.get(), h.data(), p, n * sizeof(double)), "d2h");
        return h;
    }
};

static std::vector<double> get_view(mgx::Context& c, const mgx::MultiFab* mf, size_t n) {
    std::vector<double> h(n);
    c.check(mgx_memcpy_d2h(c.get(), h.data(), mf->p, n * sizeof(double)), "d2h");
    return h;
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <dir> <n> <max_grid_size>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int n = std::atoi(argv[2]), mgs = std::atoi(argv[3]);
    mgx::BoxArray ba;
    for (int k = 0; k < n; k += mgs)
        for (int j = 0; j < n; j += mgs)
            for (int i = 0; i < n; i += mgs)
                ba.boxes.push_back({{i, j, k}, {std::min(i + mgs, n) - 1, std::min(j + mgs, n) - 1,
                                                std::min(k + mgs, n) - 1}});
    mgx::Geometry geom;
    geom.n_cell = {n, n, n};
    mgx::Context ctx(0);
    mgx::Context::make_default(ctx);

    // --- nodal projection, as ApplyNodalProjection does (variable sigma) ---
    Dev vel(ctx, read_file(dir + "/vel.bin"));
    Dev sig(ctx, read_file(dir + "/sigma.bin"));
    mgx::MultiFab velmf{(double*)vel.p, &ba, 3, 1};
    mgx::MultiFab sigmf{(double*)sig.p, &ba, 1, 0};
    mgx::LPInfo info;
    info.setMaxCoarseningLevel(100);
    {
        mgx::NodalProjector nodal_proj({&velmf}, {&sigmf}, {geom}, info);
        using BC = mgx::LinOpBCType;
        nodal_proj.setDomainBC({BC::Periodic, BC::Periodic, BC::Periodic},
                               {BC::Periodic, BC::Periodic, BC::Periodic});
        nodal_proj.project(1e-12, 0.0);
        std::printf("nodal: %d iterations, rel resid %.3e\n", nodal_proj.info().iters,
                    nodal_proj.info().rel_resid);
        long long nodes = 0, cells = 0;
        for (const mgx::Box& b : ba.boxes) {
            long long vn = 1, vc = 1;
            for (int d = 0; d < 3; ++d) {
                vn *= b.hi[d] - b.lo[d] + 2;
                vc *= b.hi[d] - b.lo[d] + 1;
            }
            nodes += vn;
            cells += vc;
        }
        write_file(dir + "/phi.bin", get_view(ctx, nodal_proj.getPhi()[0], nodes));
        write_file(dir + "/gphi.bin", get_view(ctx, nodal_proj.getGradPhi()[0], 3 * cells));
        write_file(dir + "/vel_out.bin", vel.get());
    }

    // --- MAC projection with constant beta (MacProjector + setUMAC + project) ---
    Dev u0(ctx, read_file(dir + "/umac0.bin")), u1(ctx, read_file(dir + "/umac1.bin")),
        u2(ctx, read_file(dir + "/umac2.bin"));
    mgx::MultiFab m0{(double*)u0.p, &ba, 1, 0}, m1{(double*)u1.p, &ba, 1, 0},
        m2{(double*)u2.p, &ba, 1, 0};
    long long cells = 0;
    for (const mgx::Box& b : ba.boxes)
        cells += (long long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
    Dev macphi(ctx, (size_t)cells);
    mgx::MultiFab phimf{(double*)macphi.p, &ba, 1, 0};
    {
        mgx::MacProjector macproj({geom});
        if (!macproj.needInitialization()) mgx::Abort("fresh projector must need initialisation");
        macproj.initProjector(ba, info, 0.25);
        using BC = mgx::LinOpBCType;
        macproj.setDomainBC({BC::Periodic, BC::Periodic, BC::Periodic},
                            {BC::Periodic, BC::Periodic, BC::Periodic});
        macproj.setUMAC({{&m0, &m1, &m2}});
        macproj.project({&phimf}, 1e-10, 0.0);
        std::printf("mac: %d iterations, rel resid %.3e\n", macproj.info().iters,
                    macproj.info().rel_resid);
    }
    write_file(dir + "/umac_out0.bin", u0.get());
    write_file(dir + "/umac_out1.bin", u1.get());
    write_file(dir + "/umac_out2.bin", u2.get());
    write_file(dir + "/macphi.bin", macphi.get());

    // --- two levels (refinement ratio 2), when the test wrote a fine box list:
    // NodalProjector / MacProjector over Vector<MultiFab*> of both levels ---
    if (FILE* probe = std::fopen((dir + "/fboxes.bin").c_str(), "rb")) {
        std::fclose(probe);
        const std::vector<double> fb = read_file(dir + "/fboxes.bin");
        mgx::BoxArray fba;
        for (size_t q = 0; q + 5 < fb.size(); q += 6)
            fba.boxes.push_back({{(int)fb[q], (int)fb[q + 1], (int)fb[q + 2]},
                                 {(int)fb[q + 3], (int)fb[q + 4], (int)fb[q + 5]}});
        mgx::Geometry fgeom;
        fgeom.n_cell = {2 * n, 2 * n, 2 * n};
        auto count = [](const mgx::BoxArray& b, int extra) {
            long long c = 0;
            for (const mgx::Box& x : b.boxes)
                c += (long long)(x.hi[0] - x.lo[0] + 1 + extra) * (x.hi[1] - x.lo[1] + 1 + extra) *
                     (x.hi[2] - x.lo[2] + 1 + extra);
            return c;
        };
        using BC = mgx::LinOpBCType;
        Dev vc(ctx, read_file(dir + "/amr_velc.bin")), vf(ctx, read_file(dir + "/amr_velf.bin"));
        Dev sc(ctx, read_file(dir + "/amr_sigc.bin")), sf(ctx, read_file(dir + "/amr_sigf.bin"));
        mgx::MultiFab vcm{(double*)vc.p, &ba, 3, 0}, vfm{(double*)vf.p, &fba, 3, 0};
        mgx::MultiFab scm{(double*)sc.p, &ba, 1, 0}, sfm{(double*)sf.p, &fba, 1, 0};
        {
            mgx::NodalProjector np({&vcm, &vfm}, {&scm, &sfm}, {geom, fgeom}, info);
            np.setDomainBC({BC::Periodic, BC::Periodic, BC::Periodic},
                           {BC::Periodic, BC::Periodic, BC::Periodic});
            np.project(1e-12, 0.0);
            std::printf("amr nodal: %d iterations, rel resid %.3e\n", np.info().iters,
                        np.info().rel_resid);
            write_file(dir + "/amr_phic.bin", get_view(ctx, np.getPhi()[0], count(ba, 1)));
            write_file(dir + "/amr_phif.bin", get_view(ctx, np.getPhi()[1], count(fba, 1)));
            write_file(dir + "/amr_gphic.bin", get_view(ctx, np.getGradPhi()[0], 3 * count(ba, 0)));
            write_file(dir + "/amr_gphif.bin", get_view(ctx, np.getGradPhi()[1], 3 * count(fba, 0)));
        }
        write_file(dir + "/amr_velc_out.bin", vc.get());
        write_file(dir + "/amr_velf_out.bin", vf.get());
        std::vector<Dev*> keep;
        mgx::MultiFab bm[2][3], um[2][3];
        for (int l = 0; l < 2; ++l)
            for (int d = 0; d < 3; ++d) {
                const char* lv = l ? "f" : "c";
                Dev* b = new Dev(ctx, read_file(dir + "/amr_beta" + lv + std::to_string(d) + ".bin"));
                Dev* u = new Dev(ctx, read_file(dir + "/amr_umac" + lv + std::to_string(d) + ".bin"));
                keep.push_back(b);
                keep.push_back(u);
                bm[l][d] = {(double*)b->p, l ? &fba : &ba, 1, 0};
                um[l][d] = {(double*)u->p, l ? &fba : &ba, 1, 0};
            }
        Dev mpc(ctx, (size_t)count(ba, 0)), mpf(ctx, (size_t)count(fba, 0));
        mgx::MultiFab mpcm{(double*)mpc.p, &ba, 1, 0}, mpfm{(double*)mpf.p, &fba, 1, 0};
        {
            mgx::MacProjector mp({geom, fgeom});
            mp.initProjector(info, {{&bm[0][0], &bm[0][1], &bm[0][2]},
                                    {&bm[1][0], &bm[1][1], &bm[1][2]}});
            mp.setDomainBC({BC::Periodic, BC::Periodic, BC::Periodic},
                           {BC::Periodic, BC::Periodic, BC::Periodic});
            mp.setUMAC({{&um[0][0], &um[0][1], &um[0][2]}, {&um[1][0], &um[1][1], &um[1][2]}});
            mp.project({&mpcm, &mpfm}, 1e-12, 0.0);
            std::printf("amr mac: %d iterations, rel resid %.3e\n", mp.info().iters,
                        mp.info().rel_resid);
        }
        for (int l = 0; l < 2; ++l)
            for (int d = 0; d < 3; ++d)
                write_file(dir + "/amr_umac" + (l ? "f" : "c") + std::to_string(d) + "_out.bin",
                           keep[2 * (3 * l + d) + 1]->get());
        write_file(dir + "/amr_macphic.bin", mpc.get());
        write_file(dir + "/amr_macphif.bin", mpf.get());
        for (Dev* q : keep) delete q;
    }
    std::printf("facade_check ok\n");
    return 0;
}

// step_api.hip -- C ABI of the device predictor step around the projections:
// the cell-centred viscous operator and Helmholtz solve (incflo's
// DiffusionScalarOp on the MAC hierarchy with a cell coefficient), forces,
// the velocity / density updates and periodic ghost filling.
#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "mac_mg.h"
#include "mgx_ctx.h"
#include "step_kernels.h"

using namespace mgx;

namespace {

NodalGeom level_geom(const mgx_level_desc* lv, double dx[3]) {
    int n[3], per[3], z[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        n[d] = lv->n_cell[d];
        per[d] = lv->is_periodic[d];
        dx[d] = (lv->prob_hi[d] - lv->prob_lo[d]) / n[d];
        MGX_REQUIRE(n[d] > 0 && dx[d] > 0.0, MGX_ERR_ARG, "bad level geometry");
    }
    return make_nodal_geom(n, per, z, z, dx);
}

KGeom cell_kgeom(const NodalGeom& g) {
    KGeom k{};
    for (int d = 0; d < 3; ++d) {
        k.n[d] = g.n[d];
        k.per[d] = g.per[d];
        k.nn[d] = g.n[d];
        k.sg[d] = g.n[d] + 2;
    }
    return k;
}

// Tiles over every box grown by ng cells (ghost filling), cached per ng.
struct GrownTiles {
    std::map<int, std::pair<DevBuf, int>> by_ng;
    const Tile* get(const std::vector<int>& b6, int ng, int& ntiles) {
        auto it = by_ng.find(ng);
        if (it == by_ng.end()) {
            std::vector<Tile> t;
            const int nb = (int)b6.size() / 6;
            for (int b = 0; b < nb; ++b) {
                long vol = 1;
                for (int d = 0; d < 3; ++d) vol *= (long)(b6[6 * b + 3 + d] - b6[6 * b + d] + 1 + 2 * ng);
                for (long st = 0; st < vol; st += 256) t.push_back(Tile{b, (int)st});
            }
            DevBuf buf;
            buf.alloc(t.size() * sizeof(Tile));
            MGX_HIP(hipMemcpy(buf.p, t.data(), t.size() * sizeof(Tile), hipMemcpyHostToDevice));
            it = by_ng.emplace(ng, std::make_pair(std::move(buf), (int)t.size())).first;
        }
        ntiles = it->second.second;
        return it->second.first.as<Tile>();
    }
};

// ---- fill_boundary without allocation ----------------------------------------
// A fixed device array stages the canonical copy of the field, so
// mgx_fill_boundary_noalloc allocates nothing.  Each GPU the library runs on
// holds its own copy of the array until the process ends; every call on one
// GPU shares it (calls that overlap on one GPU race).
constexpr long kFillScratchDoubles = 8L << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kFillScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
double* g_scratch_ptr[kMaxDevices] = {};

// this GPU's address of `scratch`, looked up once per device with the
// device current (the array's name is only a host-side stand-in)
double* scratch_ptr() {
    int dev = 0;
    MGX_HIP(hipGetDevice(&dev));
    MGX_REQUIRE(dev >= 0 && dev < kMaxDevices, MGX_ERR_UNSUPPORTED,
                "fill_boundary_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

// ---- the step handle: box layout of one level ------------------------------
struct mgx_step_s {
    mgx_ctx_s* ctx = nullptr;
    NodalGeom g{};
    double dx[3] = {1, 1, 1};
    BoxLayout layout;
    GrownTiles grown;
    DevBuf canon;
    hipStream_t stream() const { return ctx->stream; }
    int ntile() const { return layout.ntiles[REG_CELL]; }
};

// ---- the diffusion handle ---------------------------------------------------
struct mgx_diff_s {
    mgx_ctx_s* ctx = nullptr;
    NodalGeom g{};
    double dx[3] = {1, 1, 1};
    BoxLayout layout;
    MGOpts opts;
    // per component LinOpBCType per side (periodic directions: periodic)
    int bclo[3][3], bchi[3][3];
    DevBuf etaf[3];     // eta on faces
    bool eta_set = false;
    DevBuf ucan, acan, rhs, out, planes;
    std::map<std::array<int, 6>, std::unique_ptr<MacSolver>> solvers;
    hipStream_t stream() const { return ctx->stream; }

    DiffGeomK dg(int comp) const {
        DiffGeomK k{};
        for (int d = 0; d < 3; ++d) {
            k.n[d] = g.n[d];
            k.per[d] = g.per[d];
            k.dlo[d] = !g.per[d] && bclo[comp][d] == MGX_BC_DIRICHLET;
            k.dhi[d] = !g.per[d] && bchi[comp][d] == MGX_BC_DIRICHLET;
            k.ih2[d] = 1.0 / (dx[d] * dx[d]);
        }
        return k;
    }
    long plane(int d) const { return (long)g.n[d == 0 ? 1 : 0] * g.n[d == 2 ? 1 : 2]; }
    // boundary values of component comp from the field's first ghost layer,
    // for its Dirichlet sides; returns the six plane pointers (null: none)
    std::array<const double*, 6> load_planes(const double* field, int ng, int ncomp, int comp) {
        std::array<const double*, 6> P{};
        const DiffGeomK k = dg(comp);
        long off = 0;
        planes.alloc(2 * (plane(0) + plane(1) + plane(2)) * sizeof(double));
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                const long np = plane(d);
                if (side ? k.dhi[d] : k.dlo[d]) {
                    MGX_REQUIRE(ng >= 1, MGX_ERR_ARG,
                                "diffusion: Dirichlet sides read their values from one ghost cell");
                    double* dst = planes.as<double>() + off;
                    comp_face_ghost_kernel<<<grid1d(np), 256, 0, stream()>>>(
                        cell_kgeom(g), layout.dev_boxes(), layout.nboxes, field, ng, ncomp, comp, d,
                        side, dst);
                    P[2 * d + side] = dst;
                }
                off += np;
            }
        return P;
    }
    MacSolver& solver(int comp) {
        std::array<int, 6> key{};
        for (int d = 0; d < 3; ++d) {
            key[d] = g.per[d] ? MGX_BC_PERIODIC : bclo[comp][d];
            key[3 + d] = g.per[d] ? MGX_BC_PERIODIC : bchi[comp][d];
        }
        auto it = solvers.find(key);
        if (it == solvers.end()) {
            MacGeom mg{};
            for (int d = 0; d < 3; ++d) {
                mg.n[d] = g.n[d];
                mg.per[d] = g.per[d];
                mg.dlo[d] = !g.per[d] && key[d] == MGX_BC_DIRICHLET;
                mg.dhi[d] = !g.per[d] && key[3 + d] == MGX_BC_DIRICHLET;
                mg.dx[d] = dx[d];
            }
            mg.has_a = true;
            auto s = std::make_unique<MacSolver>();
            s->setup(mg, opts);
            it = solvers.emplace(key, std::move(s)).first;
        }
        return *it->second;
    }
    void cells_in(const double* boxed, int ng, int ncomp, int comp, double* canon) {
        comp_copy_kernel<<<layout.ntiles[REG_CELL], 256, 0, stream()>>>(
            cell_kgeom(g), layout.dev_boxes(), layout.tiles(REG_CELL), const_cast<double*>(boxed), ng,
            ncomp, comp, canon, 0);
    }
    void cells_out(double* boxed, int ng, int ncomp, int comp, double* canon) {
        comp_copy_kernel<<<layout.ntiles[REG_CELL], 256, 0, stream()>>>(
            cell_kgeom(g), layout.dev_boxes(), layout.tiles(REG_CELL), boxed, ng, ncomp, comp, canon, 1);
    }
};

extern "C" {

// ---- step handle ------------------------------------------------------------
MGX_API int mgx_step_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_step* out) {
    if (!ctx || !level || !out) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        auto h = std::make_unique<mgx_step_s>();
        h->ctx = ctx;
        h->g = level_geom(level, h->dx);
        h->layout.build(level->boxes, level->nboxes, h->g);
        h->canon.alloc(3 * (size_t)h->g.n[0] * h->g.n[1] * h->g.n[2] * sizeof(double));
        *out = h.release();
    });
}

MGX_API int mgx_step_destroy(mgx_step h) {
    if (!h) return MGX_ERR_ARG;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_fill_boundary(mgx_step h, double* field, int ng, int ncomp) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(field && ng >= 0 && ncomp >= 1, MGX_ERR_ARG, "fill_boundary: bad field");
        if (ng == 0) return;
        const long nc = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        h->canon.alloc((size_t)ncomp * nc * sizeof(double));
        const KGeom k = cell_kgeom(h->g);
        canon_all_kernel<<<h->ntile(), 256, 0, h->stream()>>>(
            k, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), field, ng, ncomp,
            h->canon.as<double>());
        int nt = 0;
        const Tile* t = h->grown.get(h->layout.boxes, ng, nt);
        fill_ghosts_kernel<<<nt, 256, 0, h->stream()>>>(k, h->layout.dev_boxes(), t, field, ng, ncomp,
                                                        h->canon.as<double>());
    });
}

MGX_API int mgx_fill_boundary_noalloc(mgx_step h, double* field, int ng, int ncomp) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(field && ng >= 0 && ncomp >= 1, MGX_ERR_ARG, "fill_boundary_noalloc: bad field");
        const long nc = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        MGX_REQUIRE((long)ncomp * nc <= kFillScratchDoubles, MGX_ERR_ARG,
                    "fill_boundary_noalloc: ncomp x cells = " + std::to_string((long)ncomp * nc) +
                        " exceeds the fixed scratch of " + std::to_string(kFillScratchDoubles) +
                        " doubles (use mgx_fill_boundary)");
        if (ng == 0) return;
        double* canon = scratch_ptr();
        const KGeom k = cell_kgeom(h->g);
        canon_all_kernel<<<h->ntile(), 256, 0, h->stream()>>>(
            k, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), field, ng, ncomp, canon);
        // grid (chunk, box) over the grown boxes: no tile table to allocate
        long vmax = 0;
        const std::vector<int>& b6 = h->layout.boxes;
        for (int b = 0; b < h->layout.nboxes; ++b) {
            long vol = 1;
            for (int d = 0; d < 3; ++d) vol *= (long)(b6[6 * b + 3 + d] - b6[6 * b + d] + 1 + 2 * ng);
            vmax = std::max(vmax, vol);
        }
        const dim3 grid((unsigned)((vmax + 255) / 256), (unsigned)h->layout.nboxes);
        fill_ghosts_boxgrid_kernel<<<grid, 256, 0, h->stream()>>>(k, h->layout.dev_boxes(), field, ng,
                                                                  ncomp, canon);
    });
}

MGX_API int mgx_compute_vel_forces(mgx_step h, const double* gp, int gp_ng, const double* rho,
                                   int rho_ng, const double* divtau, int divtau_ng,
                                   const double gravity[3], const double gp0[3],
                                   int include_pressure_gradient, double* forces, int f_ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(rho && forces && (gp || !include_pressure_gradient) && rho_ng >= 0 &&
                        f_ng >= 0 && gp_ng >= 0 && divtau_ng >= 0,
                    MGX_ERR_ARG, "compute_vel_forces: bad fields");
        ForceArgs a{};
        a.boxes = h->layout.dev_boxes();
        a.tiles = h->layout.tiles(REG_CELL);
        a.gp = gp; a.gp_ng = gp_ng;
        a.rho = rho; a.rho_ng = rho_ng;
        a.divtau = divtau; a.dt_ng = divtau_ng;
        a.out = forces; a.out_ng = f_ng;
        for (int d = 0; d < 3; ++d) {
            a.gravity[d] = gravity ? gravity[d] : 0.0;
            a.gp0[d] = gp0 ? gp0[d] : 0.0;
        }
        a.include_gp = include_pressure_gradient ? 1 : 0;
        vel_forces_kernel<<<h->ntile(), 256, 0, h->stream()>>>(a);
    });
}

MGX_API int mgx_update_velocity(mgx_step h, double* vel, int vel_ng, const double* vel_old,
                                int vo_ng, const double* conv, int conv_ng, const double* forces,
                                int f_ng, const double* divtau, int divtau_ng, double divtau_coef,
                                const double* rho_old, const double* rho_new,
                                const double* rho_nph, int rho_ng, double dt) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(vel && vel_old && conv && forces && dt > 0.0, MGX_ERR_ARG,
                    "update_velocity: bad fields");
        const bool mom = rho_old || rho_new || rho_nph;
        MGX_REQUIRE(!mom || (rho_old && rho_new && rho_nph), MGX_ERR_ARG,
                    "update_velocity: the momentum form needs rho_old, rho_new and rho_nph");
        UpdVelArgs a{};
        a.boxes = h->layout.dev_boxes();
        a.tiles = h->layout.tiles(REG_CELL);
        a.vel = vel; a.vel_ng = vel_ng;
        a.vel_old = vel_old; a.vo_ng = vo_ng;
        a.conv = conv; a.conv_ng = conv_ng;
        a.force = forces; a.f_ng = f_ng;
        a.divtau = divtau; a.dt_ng = divtau_ng;
        a.divtau_coef = divtau_coef;
        a.rho_old = rho_old; a.ro_ng = rho_ng;
        a.rho_new = rho_new; a.rn_ng = rho_ng;
        a.rho_nph = rho_nph; a.rh_ng = rho_ng;
        a.dt = dt;
        a.momentum = mom ? 1 : 0;
        update_velocity_kernel<<<h->ntile(), 256, 0, h->stream()>>>(a);
    });
}

MGX_API int mgx_update_density(mgx_step h, const double* rho_old, int ro_ng, const double* drdt,
                               int c_ng, double dt, double* rho_new, int rn_ng, double* rho_nph,
                               int rh_ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(rho_old && drdt && rho_new && dt > 0.0, MGX_ERR_ARG,
                    "update_density: bad fields");
        update_density_kernel<<<h->ntile(), 256, 0, h->stream()>>>(
            h->layout.dev_boxes(), h->layout.tiles(REG_CELL), rho_old, ro_ng, drdt, c_ng, dt,
            rho_new, rn_ng, rho_nph, rh_ng);
    });
}

// ---- diffusion handle ---------------------------------------------------------
MGX_API int mgx_diff_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_diff* out) {
    if (!ctx || !level || !out) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        auto h = std::make_unique<mgx_diff_s>();
        h->ctx = ctx;
        h->g = level_geom(level, h->dx);
        h->layout.build(level->boxes, level->nboxes, h->g);
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d)
                h->bclo[c][d] = h->bchi[c][d] = h->g.per[d] ? MGX_BC_PERIODIC : MGX_BC_NEUMANN;
        const long nc = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        for (int d = 0; d < 3; ++d)
            h->etaf[d].alloc((size_t)(h->g.n[0] + (d == 0)) * (h->g.n[1] + (d == 1)) *
                             (h->g.n[2] + (d == 2)) * sizeof(double));
        h->ucan.alloc(nc * sizeof(double));
        h->acan.alloc(nc * sizeof(double));
        h->rhs.alloc(nc * sizeof(double));
        h->out.alloc(nc * sizeof(double));
        // DiffusionScalarOp defaults (src/diffusion/DiffusionScalarOp.H:59-77)
        h->opts.maxiter = 100;
        h->opts.bottom_maxiter = 100;
        *out = h.release();
    });
}

MGX_API int mgx_diff_destroy(mgx_diff h) {
    if (!h) return MGX_ERR_ARG;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_diff_set_domain_bc(mgx_diff h, int comp, const int lo[3], const int hi[3]) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(lo && hi && comp >= -1 && comp < 3, MGX_ERR_ARG, "diffusion: bad BC call");
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? hi[d] : lo[d];
                MGX_REQUIRE((bc == MGX_BC_PERIODIC) == (h->g.per[d] != 0), MGX_ERR_ARG,
                            "diffusion: periodic BC must match the geometry's periodicity");
                MGX_REQUIRE(bc == MGX_BC_PERIODIC || bc == MGX_BC_DIRICHLET || bc == MGX_BC_NEUMANN,
                            MGX_ERR_UNSUPPORTED,
                            "diffusion: periodic, Dirichlet or Neumann sides (Robin not implemented)");
            }
        for (int c = 0; c < 3; ++c) {
            if (comp >= 0 && c != comp) continue;
            for (int d = 0; d < 3; ++d) {
                h->bclo[c][d] = lo[d];
                h->bchi[c][d] = hi[d];
            }
        }
    });
}

MGX_API int mgx_diff_set_options(mgx_diff h, const mgx_mlmg_opts* o) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGOpts m;
        to_opts(o, m);
        h->opts = m;
        h->solvers.clear();  // rebuilt with the new options
    });
}

MGX_API int mgx_diff_set_eta(mgx_diff h, const double* eta, int ng, double const_eta) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(eta || const_eta > 0.0, MGX_ERR_ARG,
                    "diffusion: need a cell eta field or a positive constant");
        const double* ec = nullptr;
        if (eta) {
            h->cells_in(eta, ng, 1, 0, h->acan.as<double>());
            ec = h->acan.as<double>();
        }
        const DiffGeomK k = h->dg(0);
        for (int d = 0; d < 3; ++d) {
            const long nf = (long)(h->g.n[0] + (d == 0)) * (h->g.n[1] + (d == 1)) * (h->g.n[2] + (d == 2));
            eta_faces_kernel<<<grid1d(nf), 256, 0, h->stream()>>>(k, d, ec, const_eta, 1.0,
                                                                 h->etaf[d].as<double>());
        }
        h->eta_set = true;
    });
}

MGX_API int mgx_diff_apply(mgx_diff h, const double* field, int ng, int ncomp, const double* rho,
                           int rho_ng, double* out, int out_ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(h->eta_set, MGX_ERR_STATE, "diffusion: eta not set");
        MGX_REQUIRE(field && out && ncomp >= 1 && ncomp <= 3 && ng >= 0 && out_ng >= 0, MGX_ERR_ARG,
                    "diffusion apply: bad fields");
        const long nc = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        const double* rc = nullptr;
        if (rho) {
            h->cells_in(rho, rho_ng, 1, 0, h->acan.as<double>());
            rc = h->acan.as<double>();
        }
        for (int c = 0; c < ncomp; ++c) {
            const auto P = h->load_planes(field, ng, ncomp, c);
            h->cells_in(field, ng, ncomp, c, h->ucan.as<double>());
            diff_apply_kernel<<<grid1d(nc), 256, 0, h->stream()>>>(
                h->dg(c), h->etaf[0].as<double>(), h->etaf[1].as<double>(), h->etaf[2].as<double>(),
                h->ucan.as<double>(), P[0], P[1], P[2], P[3], P[4], P[5], rc, h->out.as<double>());
            h->cells_out(out, out_ng, ncomp, c, h->out.as<double>());
        }
    });
}

MGX_API int mgx_diff_solve(mgx_diff h, double* field, int ng, int ncomp, const double* acoef,
                           int a_ng, double dt, double rtol, double atol, int mode,
                           mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    mgx_solve_info last{};
    bool conv_all = true;
    int status = guarded(h->ctx, [&] {
        MGX_REQUIRE(h->eta_set, MGX_ERR_STATE, "diffusion: eta not set");
        MGX_REQUIRE(field && ncomp >= 1 && ncomp <= 3 && ng >= 0 && dt > 0.0, MGX_ERR_ARG,
                    "diffusion solve: bad fields");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_ARG,
                    "diffusion solve modes: FP64 or mixed");
        const long nc = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        if (acoef) h->cells_in(acoef, a_ng, 1, 0, h->acan.as<double>());
        else const_fill_kernel<<<grid1d(nc), 256, 0, h->stream()>>>(h->acan.as<double>(), nc, 1.0);
        int iters = 0, vcycles = 0;
        for (int c = 0; c < ncomp; ++c) {
            MacSolver& S = h->solver(c);
            const DiffGeomK k = h->dg(c);
            // B = dt eta on faces, A = rho (MLABecLaplacian alpha = 1, beta = dt)
            for (int d = 0; d < 3; ++d) {
                const long nf = (long)(h->g.n[0] + (d == 0)) * (h->g.n[1] + (d == 1)) * (h->g.n[2] + (d == 2));
                scale_copy_kernel<<<grid1d(nf), 256, 0, h->stream()>>>(h->etaf[d].as<double>(), dt,
                                                                       S.beta(d), nf);
            }
            S.beta_changed();
            MGX_HIP(hipMemcpyAsync(S.acoef(), h->acan.p, nc * sizeof(double), hipMemcpyDeviceToDevice,
                                   h->stream()));
            S.acoef_changed();
            const auto P = h->load_planes(field, ng, ncomp, c);
            h->cells_in(field, ng, ncomp, c, S.phi);
            diff_rhs_kernel<<<grid1d(nc), 256, 0, h->stream()>>>(
                k, S.beta(0), S.beta(1), S.beta(2), S.phi, P[0], P[1], P[2], P[3], P[4], P[5],
                h->acan.as<double>(), h->rhs.as<double>());
            mgx_solve_info ci{};
            S.solve(h->rhs.as<double>(), rtol, atol, mode, &ci, h->stream());
            h->cells_out(field, ng, ncomp, c, S.phi);
            iters += ci.iters;
            vcycles += ci.vcycles;
            if (!ci.converged) conv_all = false;
            last = ci;
        }
        last.iters = iters;
        last.vcycles = vcycles;
        last.converged = conv_all ? 1 : 0;
    });
    if (info) *info = last;
    if (status == MGX_OK && !conv_all) {
        h->ctx->last_error = "diffusion: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

}  // extern "C"

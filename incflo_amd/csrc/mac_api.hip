// mac_api.hip -- C ABI of the MAC projection (Hydro::MacProjector surface).
#include <memory>

#include "eb_nodal.h"
#include "mac_mg.h"
#include "mgx_ctx.h"

using namespace mgx;

namespace {

__global__ void fill_kernel(double* a, long n, double v) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) a[id] = v;
}

}  // namespace

struct mgx_mac_s {
    mgx_ctx_s* ctx = nullptr;
    int n_cell[3] = {0, 0, 0};
    double prob_lo[3] = {0, 0, 0}, prob_hi[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    std::vector<int> boxes;
    int bclo[3] = {0, 0, 0}, bchi[3] = {0, 0, 0};
    bool bc_set = false;
    MGOpts opts;
    bool dirty = true;
    MacGeom g{};
    MacSolver solver;
    BoxLayout layout;
    DevBuf u[3], rhs, cells;
    bool have_phi = false;
    // embedded boundary: apertures, raw beta (without apertures), covered cells
    bool eb_on = false;
    double eb_cyl[3] = {0, 0, 0};
    DevBuf eb_ap[3], eb_braw[3], eb_cov;
    void build_eb() {
        if (!eb_on) return;
        std::vector<double> ap[3];
        std::vector<unsigned char> cov;
        eb_cylinder_apertures(g.n, prob_lo, g.dx, eb_cyl[0], eb_cyl[1], eb_cyl[2], ap, cov);
        for (int d = 0; d < 3; ++d) {
            eb_ap[d].alloc(ap[d].size() * sizeof(double));
            MGX_HIP(hipMemcpy(eb_ap[d].p, ap[d].data(), ap[d].size() * sizeof(double),
                              hipMemcpyHostToDevice));
            eb_braw[d].alloc(ap[d].size() * sizeof(double));
        }
        eb_cov.alloc(cov.size());
        MGX_HIP(hipMemcpy(eb_cov.p, cov.data(), cov.size(), hipMemcpyHostToDevice));
    }
    // after a beta setter: keep the raw beta, hand apertures * beta to the solver
    void eb_post_beta() {
        if (!eb_on) return;
        for (int d = 0; d < 3; ++d) {
            const long nf = g.nface(d);
            MGX_HIP(hipMemcpyAsync(eb_braw[d].p, solver.beta(d), nf * sizeof(double),
                                   hipMemcpyDeviceToDevice, stream()));
            eb_face_scale_kernel<<<grid1d(nf), 256, 0, stream()>>>(eb_ap[d].as<double>(),
                                                                    solver.beta(d), nf);
        }
    }
    void eb_zero_covered() {
        if (eb_on)
            eb_zero_covered_kernel<<<grid1d(g.ncell()), 256, 0, stream()>>>(
                eb_cov.as<unsigned char>(), solver.phi, g.ncell());
    }
    const double* raw_beta(int d) { return eb_on ? eb_braw[d].as<double>() : solver.beta(d); }

    KGeom kg() const {
        KGeom k{};
        for (int d = 0; d < 3; ++d) {
            k.n[d] = g.n[d];
            k.per[d] = g.per[d];
            k.nn[d] = g.n[d];
            k.sg[d] = g.n[d] + 2;
        }
        return k;
    }
    hipStream_t stream() const { return ctx->stream; }
    void ensure_setup() {
        if (!dirty) return;
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(periodic[d] || bc_set, MGX_ERR_STATE,
                        "MAC projector: setDomainBC is required for non-periodic domains");
        for (int d = 0; d < 3; ++d) {
            g.n[d] = n_cell[d];
            g.per[d] = periodic[d];
            g.dlo[d] = (!periodic[d] && bclo[d] == MGX_BC_DIRICHLET) ? 1 : 0;
            g.dhi[d] = (!periodic[d] && bchi[d] == MGX_BC_DIRICHLET) ? 1 : 0;
            g.dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
        }
        solver.setup(g, opts);
        int zero[3] = {0, 0, 0};
        double dx[3] = {g.dx[0], g.dx[1], g.dx[2]};
        NodalGeom ng = make_nodal_geom(n_cell, periodic, zero, zero, dx);
        layout.build(boxes.data(), (int)boxes.size() / 6, ng);
        for (int d = 0; d < 3; ++d) u[d].alloc(g.nface(d) * sizeof(double));
        rhs.alloc(g.ncell() * sizeof(double));
        cells.alloc(g.ncell() * sizeof(double));
        have_phi = false;
        dirty = false;
        build_eb();
    }
    void faces_in(int d, const double* boxed, int ng, double* canon) {
        face_copy_kernel<<<layout.ntiles[REG_FACE0 + d], 256, 0, stream()>>>(
            kg(), layout.dev_boxes(), layout.tiles(REG_FACE0 + d), d, const_cast<double*>(boxed),
            ng, canon, 0);
    }
    void faces_out(int d, double* boxed, int ng, double* canon) {
        face_copy_kernel<<<layout.ntiles[REG_FACE0 + d], 256, 0, stream()>>>(
            kg(), layout.dev_boxes(), layout.tiles(REG_FACE0 + d), d, boxed, ng, canon, 1);
    }
    void cells_in(const double* boxed, int ng, double* canon) {
        cell_copy_kernel<<<layout.ntiles[REG_CELL], 256, 0, stream()>>>(
            kg(), layout.dev_boxes(), layout.tiles(REG_CELL), const_cast<double*>(boxed), ng, canon,
            0);
    }
    void cells_out(double* boxed, int ng, double* canon) {
        cell_copy_kernel<<<layout.ntiles[REG_CELL], 256, 0, stream()>>>(
            kg(), layout.dev_boxes(), layout.tiles(REG_CELL), boxed, ng, canon, 1);
    }
    dim3 face_grid(int d) const {
        return grid3(g.n[0] + (d == 0), g.n[1] + (d == 1), g.n[2] + (d == 2));
    }
};

extern "C" {

MGX_API int mgx_mac_create(mgx_ctx ctx, const mgx_level_desc* lv, mgx_mac* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        MGX_REQUIRE(lv != nullptr && lv->boxes != nullptr && lv->nboxes > 0, MGX_ERR_ARG,
                    "level description needs a box list");
        auto h = std::make_unique<mgx_mac_s>();
        h->ctx = ctx;
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(lv->n_cell[d] >= 1, MGX_ERR_ARG, "n_cell must be positive");
            MGX_REQUIRE(lv->prob_hi[d] > lv->prob_lo[d], MGX_ERR_ARG, "prob_hi must exceed prob_lo");
            h->n_cell[d] = lv->n_cell[d];
            h->prob_lo[d] = lv->prob_lo[d];
            h->prob_hi[d] = lv->prob_hi[d];
            h->periodic[d] = lv->is_periodic[d] ? 1 : 0;
        }
        h->boxes.assign(lv->boxes, lv->boxes + 6 * lv->nboxes);
        // MAC defaults: maxiter / bottom_maxiter 200
        // (Docs/sphinx_documentation/source/InputsMultigrid.rst:23-30)
        h->opts.maxiter = 200;
        h->opts.bottom_maxiter = 200;
        {
            int zero[3] = {0, 0, 0};
            double dx[3];
            for (int d = 0; d < 3; ++d) dx[d] = (h->prob_hi[d] - h->prob_lo[d]) / h->n_cell[d];
            NodalGeom g = make_nodal_geom(h->n_cell, h->periodic, zero, zero, dx);
            BoxLayout probe;
            probe.build(h->boxes.data(), lv->nboxes, g);
        }
        *out = h.release();
    });
}

MGX_API int mgx_mac_destroy(mgx_mac h) {
    if (!h) return MGX_OK;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_mac_set_domain_bc(mgx_mac h, const int lo[3], const int hi[3]) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(lo && hi, MGX_ERR_ARG, "null bc arrays");
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? hi[d] : lo[d];
                MGX_REQUIRE(bc >= MGX_BC_PERIODIC && bc <= MGX_BC_ROBIN, MGX_ERR_ARG,
                            "MAC projector: undefined BC type");
                MGX_REQUIRE((bc == MGX_BC_PERIODIC) == (h->periodic[d] != 0), MGX_ERR_ARG,
                            "MAC projector: periodic BC must match the geometry's periodicity");
                MGX_REQUIRE(bc != MGX_BC_ROBIN, MGX_ERR_UNSUPPORTED,
                            "MAC projector: Robin (mixed) boundaries are not implemented");
            }
        for (int d = 0; d < 3; ++d) { h->bclo[d] = lo[d]; h->bchi[d] = hi[d]; }
        h->bc_set = true;
        h->dirty = true;
    });
}

MGX_API int mgx_mac_set_options(mgx_mac h, const mgx_mlmg_opts* o) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(o != nullptr && o->maxiter >= 1 && o->bottom_maxiter >= 1 && o->nu1 >= 0 &&
                        o->nu2 >= 0 && o->nu1 + o->nu2 >= 1,
                    MGX_ERR_ARG, "bad MLMG options");
        const bool relevel = o->max_coarsening_level != h->opts.max_coarsening_level;
        h->opts.verbose = o->verbose;
        h->opts.maxiter = o->maxiter;
        h->opts.bottom_maxiter = o->bottom_maxiter;
        h->opts.bottom_rtol = o->bottom_rtol;
        h->opts.max_coarsening_level = std::max(0, o->max_coarsening_level);
        h->opts.nu1 = o->nu1;
        h->opts.nu2 = o->nu2;
        if (relevel) h->dirty = true;
        else h->solver.opts = h->opts;
    });
}

MGX_API int mgx_mac_set_beta(mgx_mac h, const double* bx, const double* by, const double* bz,
                             int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(bx && by && bz && ng >= 0, MGX_ERR_ARG, "bad beta fields");
        h->ensure_setup();
        const double* b[3] = {bx, by, bz};
        for (int d = 0; d < 3; ++d) h->faces_in(d, b[d], ng, h->solver.beta(d));
        h->eb_post_beta();
        h->solver.beta_changed();
    });
}

MGX_API int mgx_mac_set_const_beta(mgx_mac h, double beta) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(beta > 0.0, MGX_ERR_ARG, "beta must be positive");
        h->ensure_setup();
        for (int d = 0; d < 3; ++d) {
            const long nf = h->g.nface(d);
            fill_kernel<<<grid1d(nf), 256, 0, h->stream()>>>(h->solver.beta(d), nf, beta);
        }
        h->eb_post_beta();
        h->solver.beta_changed();
    });
}

MGX_API int mgx_mac_set_beta_from_density(mgx_mac h, const double* rho, int rho_ng, double dt) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(rho && rho_ng >= 0 && dt > 0.0, MGX_ERR_ARG, "bad density / dt");
        h->ensure_setup();
        h->cells_in(rho, rho_ng, h->cells.as<double>());
        for (int d = 0; d < 3; ++d)
            mac_beta_from_rho_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                h->g.k(), d, h->cells.as<double>(), dt, h->solver.beta(d));
        h->eb_post_beta();
        h->solver.beta_changed();
    });
}

MGX_API int mgx_mac_project(mgx_mac h, double* u, double* v, double* w, int umac_ng, double* phi,
                            int phi_ng, double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        MGX_REQUIRE(u && v && w && umac_ng >= 0 && phi_ng >= 0, MGX_ERR_ARG, "bad MAC fields");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_ARG,
                    "MAC projection modes: FP64 or mixed");
        h->ensure_setup();
        MGX_REQUIRE(h->solver.beta_set, MGX_ERR_STATE, "MAC projector: beta not set");
        double* U[3] = {u, v, w};
        for (int d = 0; d < 3; ++d) h->faces_in(d, U[d], umac_ng, h->u[d].as<double>());
        const MGeomK k = h->g.k();
        if (h->eb_on)
            eb_mac_div_kernel<<<grid3(h->g.n[0], h->g.n[1], h->g.n[2]), block3(), 0, h->stream()>>>(
                k, h->u[0].as<double>(), h->u[1].as<double>(), h->u[2].as<double>(),
                h->eb_ap[0].as<double>(), h->eb_ap[1].as<double>(), h->eb_ap[2].as<double>(),
                1.0 / h->g.dx[0], 1.0 / h->g.dx[1], 1.0 / h->g.dx[2], h->rhs.as<double>());
        else
            mac_div_kernel<<<grid3(h->g.n[0], h->g.n[1], h->g.n[2]), block3(), 0, h->stream()>>>(
                k, h->u[0].as<double>(), h->u[1].as<double>(), h->u[2].as<double>(),
                1.0 / h->g.dx[0], 1.0 / h->g.dx[1], 1.0 / h->g.dx[2], h->rhs.as<double>());
        if (phi) h->cells_in(phi, phi_ng, h->solver.phi);
        else MGX_HIP(hipMemsetAsync(h->solver.phi, 0, h->g.ncell() * sizeof(double), h->stream()));
        h->solver.solve(h->rhs.as<double>(), rtol, atol, mode, info, h->stream());
        h->eb_zero_covered();
        for (int d = 0; d < 3; ++d) {
            if (h->eb_on)
                eb_mac_face_update_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                    k, d, h->raw_beta(d), h->eb_ap[d].as<double>(), h->solver.phi,
                    1.0 / h->g.dx[d], h->u[d].as<double>(), 0);
            else
                mac_face_update_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                    k, d, h->solver.beta(d), h->solver.phi, 1.0 / h->g.dx[d], h->u[d].as<double>(),
                    0);
            h->faces_out(d, U[d], umac_ng, h->u[d].as<double>());
        }
        if (phi) h->cells_out(phi, phi_ng, h->solver.phi);
        h->have_phi = true;
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "MAC projector: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

MGX_API int mgx_mac_get_fluxes(mgx_mac h, double* fx, double* fy, double* fz, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(h->have_phi, MGX_ERR_STATE, "getFluxes before project");
        MGX_REQUIRE(fx && fy && fz && ng >= 0, MGX_ERR_ARG, "bad flux fields");
        double* F[3] = {fx, fy, fz};
        const MGeomK k = h->g.k();
        for (int d = 0; d < 3; ++d) {
            if (h->eb_on)
                eb_mac_face_update_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                    k, d, h->raw_beta(d), h->eb_ap[d].as<double>(), h->solver.phi,
                    1.0 / h->g.dx[d], h->u[d].as<double>(), 1);
            else
                mac_face_update_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                    k, d, h->solver.beta(d), h->solver.phi, 1.0 / h->g.dx[d], h->u[d].as<double>(),
                    1);
            h->faces_out(d, F[d], ng, h->u[d].as<double>());
        }
    });
}

MGX_API int mgx_mac_solve(mgx_mac h, double* phi, int phi_ng, const double* rhs, int rhs_ng,
                          double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        MGX_REQUIRE(phi && rhs && phi_ng >= 0 && rhs_ng >= 0, MGX_ERR_ARG, "bad solve fields");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_ARG,
                    "MAC solve modes: FP64 or mixed");
        h->ensure_setup();
        h->cells_in(rhs, rhs_ng, h->rhs.as<double>());
        h->cells_in(phi, phi_ng, h->solver.phi);
        h->solver.solve(h->rhs.as<double>(), rtol, atol, mode, info, h->stream());
        h->eb_zero_covered();
        h->cells_out(phi, phi_ng, h->solver.phi);
        h->have_phi = true;
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "MAC MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

MGX_API int mgx_mac_set_eb_cylinder(mgx_mac h, double cx, double cy, double r) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(r > 0.0, MGX_ERR_ARG, "EB cylinder: radius must be positive");
        MGX_REQUIRE(!h->solver.beta_set, MGX_ERR_STATE,
                    "MAC projector: set the embedded boundary before the coefficients");
        bool dir = false;
        for (int d = 0; d < 3; ++d)
            if (!h->periodic[d] && (h->bclo[d] == MGX_BC_DIRICHLET || h->bchi[d] == MGX_BC_DIRICHLET))
                dir = true;
        MGX_REQUIRE(dir, MGX_ERR_UNSUPPORTED,
                    "MAC projector with an embedded boundary needs a Dirichlet (outflow) face");
        h->eb_on = true;
        h->eb_cyl[0] = cx;
        h->eb_cyl[1] = cy;
        h->eb_cyl[2] = r;
        if (!h->dirty) h->build_eb();
    });
}

}  // extern "C"

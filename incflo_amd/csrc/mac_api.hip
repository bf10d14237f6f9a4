// mac_api.hip -- C ABI of the MAC projection (Hydro::MacProjector surface).
#include <cmath>
#include <memory>
#include <vector>

#include "cc_kernels.h"
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
    DevBuf ug;  // cell-centred projection: ghosted canonical velocity
    bool have_phi = false;
    // embedded boundary: apertures and covered cells
    bool eb_on = false;
    double eb_cyl[3] = {0, 0, 0};
    DevBuf eb_ap[3], eb_cov;
    DevBuf ueb;  // EB inflow velocity (3 x cells)
    bool has_ueb = false;
    void build_eb() {
        if (!eb_on) return;
        std::vector<double> ap[3];
        std::vector<unsigned char> cov;
        eb_cylinder_apertures(g.n, prob_lo, g.dx, eb_cyl[0], eb_cyl[1], eb_cyl[2], ap, cov);
        for (int d = 0; d < 3; ++d) {
            eb_ap[d].alloc(ap[d].size() * sizeof(double));
            MGX_HIP(hipMemcpy(eb_ap[d].p, ap[d].data(), ap[d].size() * sizeof(double),
                              hipMemcpyHostToDevice));
        }
        eb_cov.alloc(cov.size());
        MGX_HIP(hipMemcpy(eb_cov.p, cov.data(), cov.size(), hipMemcpyHostToDevice));
    }
    void eb_zero_covered() {
        if (eb_on)
            eb_zero_covered_kernel<<<grid1d(g.ncell()), 256, 0, stream()>>>(
                eb_cov.as<unsigned char>(), solver.phi, g.ncell());
    }

    // Robin faces (setLevelBC robin_a / robin_b / robin_f): per side, the
    // beta scale a h / (a h + 2 b) and the constant c = f / (a h/2 + b)
    bool rob_set = false;    // coefficients received for the current BCs
    bool rob_has_f = false;  // some f != 0
    bool rob_fixes = false;  // some a > 0 (the problem is not singular)
    DevBuf rob_c[3][2];
    bool robin_side(int d, int side) const {
        return !periodic[d] && (side ? bchi[d] : bclo[d]) == MGX_BC_ROBIN;
    }
    bool any_robin() const {
        for (int d = 0; d < 3; ++d)
            if (robin_side(d, 0) || robin_side(d, 1)) return true;
        return false;
    }
    long plane(int d) const { return (long)g.n[d == 0 ? 1 : 0] * g.n[d == 2 ? 1 : 2]; }

    // The caller's face coefficients.  With apertures the solver runs on
    // a beta and braw keeps the caller's; otherwise the solver's arrays are
    // the caller's.  (The solver scales the Robin planes itself, per level.)
    DevBuf braw[3];
    bool raw_valid = false;  // braw holds the caller's current beta
    bool need_raw() const { return eb_on; }
    double* beta_dst(int d) {
        if (!need_raw()) return solver.beta(d);
        braw[d].alloc(g.nface(d) * sizeof(double));
        return braw[d].as<double>();
    }
    const double* raw_beta(int d) { return raw_valid ? braw[d].as<double>() : solver.beta(d); }
    // solver beta = caller beta x apertures
    void derive_beta() {
        if (!need_raw()) {
            if (raw_valid)  // back to the caller's values
                for (int d = 0; d < 3; ++d)
                    MGX_HIP(hipMemcpyAsync(solver.beta(d), braw[d].p, g.nface(d) * sizeof(double),
                                           hipMemcpyDeviceToDevice, stream()));
            raw_valid = false;
            return;
        }
        for (int d = 0; d < 3; ++d) {
            const long nf = g.nface(d);
            MGX_HIP(hipMemcpyAsync(solver.beta(d), braw[d].p, nf * sizeof(double),
                                   hipMemcpyDeviceToDevice, stream()));
            eb_face_scale_kernel<<<grid1d(nf), 256, 0, stream()>>>(eb_ap[d].as<double>(),
                                                                    solver.beta(d), nf);
        }
    }
    // after a beta setter wrote beta_dst()
    void post_beta() {
        raw_valid = need_raw();
        derive_beta();
        solver.beta_changed();
    }
    void robin_rhs(double* rhs) {
        if (!rob_has_f) return;
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side)
                if (robin_side(d, side))
                    robin_rhs_kernel<<<grid1d(plane(d)), 256, 0, stream()>>>(
                        g.k(), d, side, raw_beta(d), rob_c[d][side].as<double>(), 1.0 / g.dx[d],
                        rhs);
    }
    void robin_flux(int d, double* u) {
        if (!rob_has_f) return;
        for (int side = 0; side < 2; ++side)
            if (robin_side(d, side))
                robin_flux_kernel<<<grid1d(plane(d)), 256, 0, stream()>>>(
                    g.k(), d, side, raw_beta(d), rob_c[d][side].as<double>(), u);
    }

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
            g.dlo[d] = (!periodic[d] && (bclo[d] == MGX_BC_DIRICHLET || bclo[d] == MGX_BC_ROBIN)) ? 1 : 0;
            g.dhi[d] = (!periodic[d] && (bchi[d] == MGX_BC_DIRICHLET || bchi[d] == MGX_BC_ROBIN)) ? 1 : 0;
            g.rlo[d] = robin_side(d, 0) ? 1 : 0;
            g.rhi[d] = robin_side(d, 1) ? 1 : 0;
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
        g.robin_fixes = rob_set && rob_fixes;
        solver.set_robin_fixes(g.robin_fixes);
        have_phi = false;
        dirty = false;
        build_eb();
        // coefficients kept across the set-up: re-derive the solver's
        if (solver.beta_set) {
            if (!raw_valid && need_raw()) {
                for (int d = 0; d < 3; ++d) {
                    braw[d].alloc(g.nface(d) * sizeof(double));
                    MGX_HIP(hipMemcpyAsync(braw[d].p, solver.beta(d), g.nface(d) * sizeof(double),
                                           hipMemcpyDeviceToDevice, stream()));
                }
                raw_valid = true;
            }
            derive_beta();
            solver.beta_changed();
        } else {
            raw_valid = false;
        }
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
                MGX_REQUIRE(bc != MGX_BC_INOUT, MGX_ERR_UNSUPPORTED,
                            "MAC projector: direction-dependent (in/out) faces need the outflow rescaling of HydroUtils::enforceInOutSolvability, which this build does not implement");
                MGX_REQUIRE(bc >= MGX_BC_PERIODIC && bc <= MGX_BC_ROBIN, MGX_ERR_ARG,
                            "MAC projector: undefined BC type");
                MGX_REQUIRE((bc == MGX_BC_PERIODIC) == (h->periodic[d] != 0), MGX_ERR_ARG,
                            "MAC projector: periodic BC must match the geometry's periodicity");
                MGX_REQUIRE(bc != MGX_BC_ROBIN || !h->eb_on, MGX_ERR_UNSUPPORTED,
                            "MAC projector: Robin faces with an embedded boundary are not supported");
            }
        for (int d = 0; d < 3; ++d) { h->bclo[d] = lo[d]; h->bchi[d] = hi[d]; }
        h->bc_set = true;
        h->dirty = true;
        h->rob_set = h->rob_has_f = h->rob_fixes = false;  // setLevelBC follows
        h->solver.clear_robin();
    });
}

MGX_API int mgx_mac_set_options(mgx_mac h, const mgx_mlmg_opts* o) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGOpts m;
        to_opts(o, m);
        const bool relevel = m.max_coarsening_level != h->opts.max_coarsening_level;
        h->opts = m;
        if (relevel) h->dirty = true;
        else h->solver.opts = h->opts;
    });
}

MGX_API int mgx_mac_set_graphs(mgx_mac h, int on) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        h->solver.H64.graphs = on != 0;
        h->solver.H32.graphs = on != 0;
        if (!on) {
            h->solver.H64.drop_graphs();
            h->solver.H32.drop_graphs();
        }
    });
}

MGX_API int mgx_mac_set_beta(mgx_mac h, const double* bx, const double* by, const double* bz,
                             int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(bx && by && bz && ng >= 0, MGX_ERR_ARG, "bad beta fields");
        h->ensure_setup();
        const double* b[3] = {bx, by, bz};
        for (int d = 0; d < 3; ++d) h->faces_in(d, b[d], ng, h->beta_dst(d));
        h->post_beta();
    });
}

MGX_API int mgx_mac_set_const_beta(mgx_mac h, double beta) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(beta > 0.0, MGX_ERR_ARG, "beta must be positive");
        h->ensure_setup();
        for (int d = 0; d < 3; ++d) {
            const long nf = h->g.nface(d);
            fill_kernel<<<grid1d(nf), 256, 0, h->stream()>>>(h->beta_dst(d), nf, beta);
        }
        h->post_beta();
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
                h->g.k(), d, h->cells.as<double>(), dt, h->beta_dst(d));
        h->post_beta();
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
        MGX_REQUIRE(h->rob_set || !h->any_robin(), MGX_ERR_STATE,
                    "MAC projector: Robin faces need setLevelBC coefficients");
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
        if (h->eb_on && h->has_ueb)
            eb_mac_inflow_kernel<<<grid3(h->g.n[0], h->g.n[1], h->g.n[2]), block3(), 0, h->stream()>>>(
                k, h->ueb.as<double>(), h->eb_ap[0].as<double>(), h->eb_ap[1].as<double>(),
                h->eb_ap[2].as<double>(), 1.0 / h->g.dx[0], 1.0 / h->g.dx[1], 1.0 / h->g.dx[2],
                h->rhs.as<double>());
        h->robin_rhs(h->rhs.as<double>());
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
                    k, d, h->solver.fine_beta(d, h->stream()), h->solver.phi, 1.0 / h->g.dx[d],
                    h->u[d].as<double>(),
                    0);
            h->robin_flux(d, h->u[d].as<double>());
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
                    k, d, h->solver.fine_beta(d, h->stream()), h->solver.phi, 1.0 / h->g.dx[d],
                    h->u[d].as<double>(),
                    1);
            h->robin_flux(d, h->u[d].as<double>());
            h->faces_out(d, F[d], ng, h->u[d].as<double>());
        }
    });
}

MGX_API int mgx_apply_cc_projection(mgx_mac h, const double* density, int rho_ng,
                                    double const_rho, double* vel, int vel_ng,
                                    const double* vel_old, int vel_old_ng, double* gp, int gp_ng,
                                    double* p_cc, int p_ng, double scaling_factor, int incremental,
                                    const double* inflow_vel, double* umac, double* vmac,
                                    double* wmac, int umac_ng, double rtol, double atol, int mode,
                                    mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        MGX_REQUIRE(vel && gp && p_cc && vel_ng >= 1 && gp_ng >= 0 && p_ng >= 0, MGX_ERR_ARG,
                    "cc projection: velocity (>= 1 ghost), gp and p_cc are required");
        MGX_REQUIRE(density != nullptr || const_rho > 0.0, MGX_ERR_ARG,
                    "cc projection: need a density field or a positive constant density");
        MGX_REQUIRE(scaling_factor > 0.0, MGX_ERR_ARG, "cc projection: scaling factor must be positive");
        MGX_REQUIRE(!incremental || vel_old, MGX_ERR_ARG, "cc projection: incremental needs vel_old");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_ARG,
                    "cc projection modes: FP64 or mixed");
        MGX_REQUIRE(!h->eb_on, MGX_ERR_UNSUPPORTED,
                    "cc projection with embedded boundaries (the reference aborts there too, "
                    "src/projection/incflo_apply_cc_projection.cpp:377)");
        h->ensure_setup();
        MGX_REQUIRE(h->rob_set || !h->any_robin(), MGX_ERR_STATE,
                    "cc projection: Robin faces need setLevelBC coefficients");
        const KGeom kg = h->kg();
        const MacGeom& g = h->g;
        h->ug.alloc(3 * (size_t)kg.sg[0] * kg.sg[1] * kg.sg[2] * sizeof(double));
        // u' = u (+ s/rho gp) (- u_old), ghosts 0 / inflow (:209-250)
        GatherVelArgs a{};
        a.g = kg;
        a.boxes = h->layout.dev_boxes();
        a.tiles = h->layout.tiles(REG_CELL_GROWN);
        a.vel = vel;
        a.vel_ng = vel_ng;
        a.vel_old = incremental ? vel_old : nullptr;
        a.vold_ng = vel_old_ng;
        a.gp = incremental ? nullptr : gp;
        a.gp_ng = gp_ng;
        a.rho = density;
        a.rho_ng = rho_ng;
        a.const_rho = density ? 0.0 : const_rho;
        a.scale = scaling_factor;
        a.sigma_from_rho = 0;
        a.ug = h->ug.as<double>();
        a.ghost_rule = 1;
        for (int d = 0; d < 3; ++d) { a.bclo[d] = h->bclo[d]; a.bchi[d] = h->bchi[d]; }
        if (inflow_vel && !incremental) for (int q = 0; q < 18; ++q) a.inflow[q] = inflow_vel[q];
        else for (int d = 0; d < 3; ++d) { a.bclo[d] = a.bchi[d] = MGX_BC_NEUMANN; }
        gather_vel_kernel<<<h->layout.ntiles[REG_CELL_GROWN], 256, 0, h->stream()>>>(a);
        // beta = s / face-average(rho), or s / ro_0 (:253-320)
        if (density) {
            h->cells_in(density, rho_ng, h->cells.as<double>());
            for (int d = 0; d < 3; ++d)
                mac_beta_from_rho_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                    g.k(), d, h->cells.as<double>(), scaling_factor, h->beta_dst(d));
        } else {
            for (int d = 0; d < 3; ++d) {
                const long nf = g.nface(d);
                fill_kernel<<<grid1d(nf), 256, 0, h->stream()>>>(h->beta_dst(d), nf,
                                                                scaling_factor / const_rho);
            }
        }
        h->post_beta();
        // MOL face velocities (:329-343)
        for (int d = 0; d < 3; ++d) {
            const long nf = g.nface(d);
            mol_faces_kernel<<<grid1d(nf), 256, 0, h->stream()>>>(
                kg, d, (!g.per[d] && h->bclo[d] == MGX_BC_DIRICHLET) ? 1 : 0,
                (!g.per[d] && h->bchi[d] == MGX_BC_DIRICHLET) ? 1 : 0, h->ug.as<double>(),
                h->u[d].as<double>());
        }
        // MAC projection from phi = 0 (:345-348)
        const MGeomK k = g.k();
        mac_div_kernel<<<grid3(g.n[0], g.n[1], g.n[2]), block3(), 0, h->stream()>>>(
            k, h->u[0].as<double>(), h->u[1].as<double>(), h->u[2].as<double>(), 1.0 / g.dx[0],
            1.0 / g.dx[1], 1.0 / g.dx[2], h->rhs.as<double>());
        h->robin_rhs(h->rhs.as<double>());
        MGX_HIP(hipMemsetAsync(h->solver.phi, 0, g.ncell() * sizeof(double), h->stream()));
        h->solver.solve(h->rhs.as<double>(), rtol, atol, mode, info, h->stream());
        double* U[3] = {umac, vmac, wmac};
        for (int d = 0; d < 3; ++d) {
            const double* bd = h->solver.fine_beta(d, h->stream());
            if (U[d]) {  // the projected face velocities, when asked for
                mac_face_update_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                    k, d, bd, h->solver.phi, 1.0 / g.dx[d], h->u[d].as<double>(), 0);
                h->robin_flux(d, h->u[d].as<double>());
                h->faces_out(d, U[d], umac_ng, h->u[d].as<double>());
            }
            // getFluxes (:362-371): F = -beta grad phi
            mac_face_update_kernel<<<h->face_grid(d), block3(), 0, h->stream()>>>(
                k, d, bd, h->solver.phi, 1.0 / g.dx[d], h->u[d].as<double>(), 1);
            h->robin_flux(d, h->u[d].as<double>());
        }
        // u += gphi, p_cc, gp (:373-458)
        CCUpdateArgs c{};
        c.g = kg;
        c.boxes = h->layout.dev_boxes();
        c.tiles = h->layout.tiles(REG_CELL);
        c.ug = h->ug.as<double>();
        for (int d = 0; d < 3; ++d) c.F[d] = h->u[d].as<double>();
        c.phi = h->solver.phi;
        c.vel = vel;
        c.vel_ng = vel_ng;
        c.vel_old = vel_old;
        c.vold_ng = vel_old_ng;
        c.gp = gp;
        c.gp_ng = gp_ng;
        c.p = p_cc;
        c.p_ng = p_ng;
        c.rho = density;
        c.rho_ng = rho_ng;
        c.const_rho = density ? 0.0 : const_rho;
        c.scale = scaling_factor;
        c.incremental = incremental ? 1 : 0;
        cc_update_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(c);
        h->have_phi = true;
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "cc projection: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

MGX_API int mgx_mac_solve(mgx_mac h, double* phi, int phi_ng, const double* rhs, int rhs_ng,
                          double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        MGX_REQUIRE(phi && rhs && phi_ng >= 0 && rhs_ng >= 0, MGX_ERR_ARG, "bad solve fields");
        MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED, MGX_ERR_ARG,
                    "MAC solve modes: FP64 or mixed");
        h->ensure_setup();
        MGX_REQUIRE(h->rob_set || !h->any_robin(), MGX_ERR_STATE,
                    "MAC solve: Robin faces need setLevelBC coefficients");
        h->cells_in(rhs, rhs_ng, h->rhs.as<double>());
        h->robin_rhs(h->rhs.as<double>());
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

MGX_API int mgx_mac_set_level_bc(mgx_mac h, const double* robin_a, const double* robin_b,
                                 const double* robin_f, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(robin_a && robin_b && robin_f && ng >= 1, MGX_ERR_ARG,
                    "setLevelBC: robin_a / robin_b / robin_f need at least one ghost cell");
        h->ensure_setup();
        const KGeom kg = h->kg();
        bool fixes = false, has_f = false;
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                if (!h->robin_side(d, side)) continue;
                const long np = h->plane(d);
                DevBuf tmp;
                tmp.alloc(3 * np * sizeof(double));
                const double* src[3] = {robin_a, robin_b, robin_f};
                for (int q = 0; q < 3; ++q)
                    gather_face_ghost_kernel<<<grid1d(np), 256, 0, h->stream()>>>(
                        kg, h->layout.dev_boxes(), h->layout.nboxes, src[q], ng, d, side,
                        tmp.as<double>() + q * np);
                std::vector<double> v(3 * np), c(np);
                MGX_HIP(hipMemcpyAsync(v.data(), tmp.p, 3 * np * sizeof(double),
                                       hipMemcpyDeviceToHost, h->stream()));
                MGX_HIP(hipStreamSynchronize(h->stream()));
                const double hd = h->g.dx[d];
                for (long t = 0; t < np; ++t) {
                    const double a = v[t], b = v[np + t], f = v[2 * np + t];
                    MGX_REQUIRE(std::isfinite(a) && std::isfinite(b) && std::isfinite(f) && a >= 0.0 &&
                                    b >= 0.0 && a * hd / 2 + b > 0.0,
                                MGX_ERR_ARG,
                                "setLevelBC: Robin coefficients need a, b >= 0 and a h/2 + b > 0 "
                                "on every face of a Robin side");
                    c[t] = f / (a * hd / 2 + b);
                    if (a > 0.0) fixes = true;
                    if (f != 0.0) has_f = true;
                }
                h->rob_c[d][side].alloc(np * sizeof(double));
                MGX_HIP(hipMemcpyAsync(h->rob_c[d][side].p, c.data(), np * sizeof(double),
                                       hipMemcpyHostToDevice, h->stream()));
                MGX_HIP(hipStreamSynchronize(h->stream()));
                h->solver.set_robin(d, side, std::vector<double>(v.begin(), v.begin() + np),
                                    std::vector<double>(v.begin() + np, v.begin() + 2 * np));
            }
        h->rob_set = h->any_robin();
        h->rob_fixes = fixes;
        h->rob_has_f = has_f;
        h->g.robin_fixes = h->rob_set && fixes;
        h->solver.set_robin_fixes(h->g.robin_fixes);
    });
}

MGX_API int mgx_mac_set_eb_inflow_velocity(mgx_mac h, const double* eb_vel, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        if (!eb_vel) {
            h->has_ueb = false;
            return;
        }
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "EB inflow velocity: bad ghost width");
        MGX_REQUIRE(h->eb_on, MGX_ERR_STATE, "EB inflow velocity: no embedded boundary set");
        h->ensure_setup();
        const long nc = h->g.ncell();
        h->ueb.alloc(3 * nc * sizeof(double));
        gather_cells_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            h->kg(), h->layout.dev_boxes(), h->layout.tiles(REG_CELL), eb_vel, ng, 3,
            h->ueb.as<double>());
        h->has_ueb = true;
    });
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
        MGX_REQUIRE(!h->any_robin(), MGX_ERR_UNSUPPORTED,
                    "MAC projector: Robin faces with an embedded boundary are not supported");
        h->eb_on = true;
        h->eb_cyl[0] = cx;
        h->eb_cyl[1] = cy;
        h->eb_cyl[2] = r;
        if (!h->dirty) h->build_eb();
    });
}

}  // extern "C"

// mgx_api.hip -- the C ABI (include/mgx.h): context, device-memory plumbing,
// and the nodal projection entry points.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>

#include "eb_nodal.h"
#include "mgx_ctx.h"
#include "nodal_mg.h"

using namespace mgx;

namespace {

__global__ void fill_sigma_const_kernel(KGeom g, double v, double* sig) {
    int i, j, k;
    if (!grid_point(g.n[0], g.n[1], i, j, k)) return;
    write_cell_images(sig, g, i, j, k, v);
}

}  // namespace

static __global__ void __launch_bounds__(256) stream_copy_kernel(float4* __restrict__ dst,
                                                                 const float4* __restrict__ src,
                                                                 long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        dst[i] = src[i];
}

struct mgx_nodal_s {
    mgx_ctx_s* ctx = nullptr;
    int n_cell[3] = {0, 0, 0};
    double prob_lo[3] = {0, 0, 0}, prob_hi[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    std::vector<int> boxes;
    int bclo[3] = {0, 0, 0}, bchi[3] = {0, 0, 0};
    bool bc_set = false;
    MGOpts opts;
    bool dirty = true;
    NodalGeom g{};
    NodalSolver solver;
    BoxLayout layout;
    DevBuf ug, rhs;
    bool have_phi = false;
    // embedded boundary request: 0 none, 1 z-cylinder (eb_cyl = cx, cy, r), 2 moments
    // overset mask (canonical node flags) as the caller set it
    DevBuf omask;
    bool has_omask = false;
    int eb_kind = 0;
    double eb_cyl[3] = {0, 0, 0};
    std::vector<int> eb_type;
    std::vector<double> eb_mom;
    std::unique_ptr<EBNodal> eb;
    // EB inflow velocity (setEBInflowVelocity): body-surface weights per cut
    // cell (cylinder geometry) and the caller's velocity (3 x cells)
    DevBuf eb_bcut, ueb;
    bool has_ueb = false;

    void build_eb() {
        solver.eb = nullptr;
        eb.reset();
        if (eb_kind == 0) return;
        if (eb_kind == 1) eb_cylinder_moments(g, prob_lo, eb_cyl[0], eb_cyl[1], eb_cyl[2], eb_type, eb_mom);
        eb = std::make_unique<EBNodal>();
        eb->setup(g, solver.nlev, eb_type, eb_mom, stream());
        solver.eb = eb.get();
        if (eb_kind == 1) {
            std::vector<double> bw;
            eb_cylinder_surface_weights(g, prob_lo, eb_cyl[0], eb_cyl[1], eb_cyl[2], eb_type, bw);
            const std::vector<int>& cut = eb->L[0].cut_cells;
            std::vector<double> bc(24 * std::max<size_t>(cut.size(), 1), 0.0);
            for (size_t q = 0; q < cut.size(); ++q)
                std::copy(&bw[24L * cut[q]], &bw[24L * cut[q]] + 24, &bc[24 * q]);
            eb_bcut.alloc(bc.size() * sizeof(double));
            MGX_HIP(hipMemcpy(eb_bcut.p, bc.data(), bc.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }

    void ensure_setup() {
        if (!dirty) return;
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(periodic[d] || bc_set, MGX_ERR_STATE,
                        "nodal projector: setDomainBC is required for non-periodic domains");
        double dx[3];
        for (int d = 0; d < 3; ++d) dx[d] = (prob_hi[d] - prob_lo[d]) / n_cell[d];
        g = make_nodal_geom(n_cell, periodic, bclo, bchi, dx);
        solver.setup(g, opts);
        layout.build(boxes.data(), (int)boxes.size() / 6, g);
        ug.alloc(3 * g.nsig() * sizeof(double));
        MGX_HIP(hipMemset(ug.p, 0, 3 * g.nsig() * sizeof(double)));
        rhs.alloc(g.nnode() * sizeof(double));
        have_phi = false;
        dirty = false;
        build_eb();
        if (has_omask) solver.set_mask(omask.as<unsigned char>(), ctx->stream);
    }
    hipStream_t stream() const { return ctx->stream; }
    void compute_rhs() {
        const KGeom kg = kgeom(g);
        if (eb) {
            eb_rhs_kernel<<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, stream()>>>(
                kg, eb->cidx0(), eb->W0(), eb->mask0(), ug.as<double>(), rhs.as<double>(),
                0.25 / g.dx[0], 0.25 / g.dx[1], 0.25 / g.dx[2]);
            if (has_ueb && eb_kind == 1)
                eb_inflow_rhs_kernel<<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, stream()>>>(
                    kg, eb->cidx0(), eb_bcut.as<double>(), eb->mask0(), ueb.as<double>(),
                    rhs.as<double>());
            return;
        }
        nodal_rhs_kernel<<<grid3(g.nn[0], g.nn[1], g.nn[2]), block3(), 0, stream()>>>(
            kg, ug.as<double>(), rhs.as<double>(), 0.25 / g.dx[0], 0.25 / g.dx[1], 0.25 / g.dx[2]);
    }
    void zero_phi() {
        MGX_HIP(hipMemsetAsync(solver.phi, 0, g.nnode() * sizeof(double), stream()));
        solver.phi_zero = true;
    }
    UpdateArgs update_args() {
        UpdateArgs u{};
        u.g = kgeom(g);
        u.boxes = layout.dev_boxes();
        u.tiles = layout.tiles(REG_CELL);
        u.phi = solver.phi;
        u.sig = solver.sigma0();
        for (int d = 0; d < 3; ++d) u.f[d] = 0.25 / g.dx[d];
        if (eb) {
            u.eb_cidx = eb->cidx0();
            u.eb_W = eb->W0();
            u.eb_V = eb->V0();
        }
        return u;
    }
};

static void check_mode(int mode) {
    MGX_REQUIRE(mode == MGX_MODE_FP64 || mode == MGX_MODE_MIXED || mode == MGX_MODE_LEARNED,
                MGX_ERR_ARG, "unknown solve mode " + std::to_string(mode));
}


extern "C" {

MGX_API const char* mgx_version(void) { return "mgx 0.1 (gfx950)"; }

MGX_API void mgx_default_mlmg_opts(mgx_mlmg_opts* o) {
    if (!o) return;
    MGOpts m;
    o->verbose = 0;
    o->bottom_verbose = 0;
    o->maxiter = m.maxiter;
    o->bottom_maxiter = m.bottom_maxiter;
    o->bottom_rtol = m.bottom_rtol;
    o->bottom_atol = -1.0;
    o->max_coarsening_level = m.max_coarsening_level;
    o->nu1 = m.nu1;
    o->nu2 = m.nu2;
    o->omega = m.omega;
    o->stall_iters = m.stall_iters;
    o->cycles_per_update = m.cycles_per_update;
    o->bottom_solver = m.bottom_solver;
    o->smoother = m.smoother;
    o->learned_steps = m.learned_steps;
}

MGX_API int mgx_ctx_create(int device, void* stream, mgx_ctx* out) {
    if (!out) return MGX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MGX_ERR_HIP;
    auto* c = new mgx_ctx_s();
    c->device = device;
    c->stream = (hipStream_t)stream;
    if (hipSetDevice(device) != hipSuccess) { delete c; return MGX_ERR_HIP; }
    *out = c;
    return MGX_OK;
}

MGX_API int mgx_ctx_destroy(mgx_ctx ctx) {
    if (!ctx) return MGX_OK;
    ctx->make_current();
    if (ctx->events_made)
        for (auto& e : ctx->events) (void)hipEventDestroy(e);
    delete ctx;
    return MGX_OK;
}

MGX_API const char* mgx_ctx_last_error(mgx_ctx ctx) { return ctx ? ctx->last_error.c_str() : ""; }

MGX_API int mgx_ctx_synchronize(mgx_ctx ctx) {
    return guarded(ctx, [&] { MGX_HIP(hipStreamSynchronize(ctx->stream)); });
}

MGX_API int mgx_dev_malloc(mgx_ctx ctx, unsigned long long bytes, void** out) {
    return guarded(ctx, [&] {
        MGX_REQUIRE(out != nullptr, MGX_ERR_ARG, "null output pointer");
        MGX_HIP(hipMalloc(out, bytes ? bytes : 1));
    });
}

MGX_API int mgx_dev_free(mgx_ctx ctx, void* p) {
    return guarded(ctx, [&] { MGX_HIP(hipFree(p)); });
}

MGX_API int mgx_memcpy_h2d(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes) {
    return guarded(ctx, [&] {
        MGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        MGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

MGX_API int mgx_memcpy_d2h(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes) {
    return guarded(ctx, [&] {
        MGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        MGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

MGX_API int mgx_memcpy_d2d(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes) {
    return guarded(ctx, [&] {
        MGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}

MGX_API int mgx_stream_copy(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes) {
    return guarded(ctx, [&] {
        MGX_REQUIRE(dst && src && bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 &&
                        ((uintptr_t)src % 16) == 0,
                    MGX_ERR_ARG, "stream copy: 16-byte aligned pointers and sizes required");
        const long n = (long)(bytes / 16);
        if (n == 0) return;
        const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 8192);
        stream_copy_kernel<<<blocks, 256, 0, ctx->stream>>>((float4*)dst, (const float4*)src, n);
    });
}

MGX_API int mgx_memset_d(mgx_ctx ctx, void* dst, int value, unsigned long long bytes) {
    return guarded(ctx, [&] { MGX_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream)); });
}

MGX_API int mgx_ctx_record(mgx_ctx ctx, int slot) {
    return guarded(ctx, [&] {
        MGX_REQUIRE(slot >= 0 && slot < 16, MGX_ERR_ARG, "event slot out of range");
        if (!ctx->events_made) {
            for (auto& e : ctx->events) MGX_HIP(hipEventCreate(&e));
            ctx->events_made = true;
        }
        MGX_HIP(hipEventRecord(ctx->events[slot], ctx->stream));
    });
}

MGX_API int mgx_ctx_elapsed_ms(mgx_ctx ctx, int a, int b, double* ms) {
    return guarded(ctx, [&] {
        MGX_REQUIRE(ctx->events_made && a >= 0 && a < 16 && b >= 0 && b < 16 && ms, MGX_ERR_ARG,
                    "bad event slots");
        MGX_HIP(hipEventSynchronize(ctx->events[b]));
        float f = 0.f;
        MGX_HIP(hipEventElapsedTime(&f, ctx->events[a], ctx->events[b]));
        *ms = f;
    });
}

// ---- nodal ---------------------------------------------------------------

MGX_API int mgx_nodal_create(mgx_ctx ctx, const mgx_level_desc* lv, mgx_nodal* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        MGX_REQUIRE(lv != nullptr && lv->boxes != nullptr && lv->nboxes > 0, MGX_ERR_ARG,
                    "level description needs a box list");
        auto h = std::make_unique<mgx_nodal_s>();
        h->ctx = ctx;
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(lv->n_cell[d] >= 1, MGX_ERR_ARG, "n_cell must be positive");
            MGX_REQUIRE(lv->prob_hi[d] > lv->prob_lo[d], MGX_ERR_ARG, "prob_hi must exceed prob_lo");
            h->n_cell[d] = lv->n_cell[d];
            h->prob_lo[d] = lv->prob_lo[d];
            h->prob_hi[d] = lv->prob_hi[d];
            h->periodic[d] = lv->is_periodic[d] ? 1 : 0;
            h->bclo[d] = h->bchi[d] = MGX_BC_PERIODIC;
        }
        h->boxes.assign(lv->boxes, lv->boxes + 6 * lv->nboxes);
        // validate the box list now (errors at construction, like BoxArray checks)
        {
            double dx[3];
            for (int d = 0; d < 3; ++d) dx[d] = (h->prob_hi[d] - h->prob_lo[d]) / h->n_cell[d];
            NodalGeom g = make_nodal_geom(h->n_cell, h->periodic, h->bclo, h->bchi, dx);
            BoxLayout probe;
            probe.build(h->boxes.data(), lv->nboxes, g);
        }
        *out = h.release();
    });
}

MGX_API int mgx_nodal_destroy(mgx_nodal h) {
    if (!h) return MGX_OK;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_nodal_set_domain_bc(mgx_nodal h, const int lo[3], const int hi[3]) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(lo && hi, MGX_ERR_ARG, "null bc arrays");
        for (int d = 0; d < 3; ++d) {
            for (int side = 0; side < 2; ++side) {
                const int bc = side ? hi[d] : lo[d];
                MGX_REQUIRE(bc != MGX_BC_INOUT, MGX_ERR_UNSUPPORTED,
                            "nodal projector: direction-dependent (in/out) faces need the outflow rescaling of HydroUtils::enforceInOutSolvability, which this build does not implement");
                MGX_REQUIRE(bc >= MGX_BC_PERIODIC && bc <= MGX_BC_ROBIN, MGX_ERR_ARG,
                            "nodal projector: undefined BC type");
                MGX_REQUIRE((bc == MGX_BC_PERIODIC) == (h->periodic[d] != 0), MGX_ERR_ARG,
                            "nodal projector: periodic BC must match the geometry's periodicity");
                MGX_REQUIRE(bc != MGX_BC_ROBIN, MGX_ERR_UNSUPPORTED,
                            "nodal projector: Robin is not a nodal-projection BC");
            }
            h->bclo[d] = lo[d];
            h->bchi[d] = hi[d];
        }
        h->bc_set = true;
        h->dirty = true;
    });
}

MGX_API int mgx_nodal_set_options(mgx_nodal h, const mgx_mlmg_opts* o) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(o != nullptr, MGX_ERR_ARG, "null options");
        MGOpts m;
        to_opts(o, m);
        const bool relevel = m.max_coarsening_level != h->opts.max_coarsening_level;
        h->opts = m;
        if (relevel) h->dirty = true;
        else h->solver.opts = m;
    });
}

MGX_API int mgx_nodal_autotune(mgx_nodal h, int mode, char* report, int report_len) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        check_mode(mode);
        h->ensure_setup();
        const std::string r = h->solver.autotune(mode, h->stream());
        if (report && report_len > 0) {
            const size_t n = std::min<size_t>(r.size(), (size_t)report_len - 1);
            std::memcpy(report, r.data(), n);
            report[n] = 0;
        }
    });
}

MGX_API int mgx_tuning_get(char* buf, int len) {
    try {
        const std::string t = tuning_table();
        if (!buf || len <= (int)t.size()) return MGX_ERR_ARG;
        std::memcpy(buf, t.c_str(), t.size() + 1);
        return MGX_OK;
    } catch (...) {
        return MGX_ERR_ARG;
    }
}

MGX_API int mgx_tuning_set(const char* text) {
    if (!text) return MGX_ERR_ARG;
    try {
        set_tuning_table(text);
        return MGX_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (...) {
        return MGX_ERR_ARG;
    }
}

MGX_API int mgx_nodal_set_timing(mgx_nodal h, int on) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        h->solver.fine_timer.on = on != 0;
        h->solver.fine_timer.reset();
    });
}

MGX_API int mgx_nodal_set_graphs(mgx_nodal h, int on) {
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

MGX_API int mgx_nodal_set_overset_mask(mgx_nodal h, const int* mask, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        h->ensure_setup();
        if (!mask) {
            h->has_omask = false;
            h->solver.set_mask(nullptr, h->stream());
            return;
        }
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "overset mask ghost width out of range");
        MGX_REQUIRE(h->eb_kind == 0, MGX_ERR_UNSUPPORTED,
                    "overset mask with embedded boundaries is not supported");
        const long nn = h->g.nnode();
        h->omask.alloc(nn);
        MGX_HIP(hipMemsetAsync(h->omask.p, 1, nn, h->stream()));
        gather_node_flags_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kgeom(h->g), h->layout.dev_boxes(), h->layout.tiles(REG_NODE), mask, ng,
            h->omask.as<unsigned char>());
        h->has_omask = true;
        h->solver.set_mask(h->omask.as<unsigned char>(), h->stream());
    });
}

MGX_API int mgx_nodal_set_vcycle_bits(mgx_nodal h, int bits) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(bits == 16 || bits == 32, MGX_ERR_ARG, "V-cycle storage must be 16 or 32 bits");
        h->solver.use_half = bits == 16;
    });
}

MGX_API int mgx_nodal_get_timing(mgx_nodal h, int which, double* total_ms, long long* count) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(which == 0 && total_ms && count, MGX_ERR_ARG, "bad timing query");
        *total_ms = h->solver.fine_timer.total_ms;
        *count = h->solver.fine_timer.count;
    });
}

MGX_API int mgx_nodal_set_sigma(mgx_nodal h, const double* sigma, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(sigma != nullptr && ng >= 0, MGX_ERR_ARG, "bad sigma field");
        h->ensure_setup();
        gather_sigma_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            kgeom(h->g), h->layout.dev_boxes(), h->layout.tiles(REG_CELL), sigma, ng,
            h->solver.sigma0());
        h->solver.sigma_changed();
    });
}

MGX_API int mgx_nodal_set_const_sigma(mgx_nodal h, double v) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(v > 0.0, MGX_ERR_ARG, "sigma must be positive");
        h->ensure_setup();
        const long nc = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        fill_sigma_const_kernel<<<grid3(h->g.n[0], h->g.n[1], h->g.n[2]), block3(), 0, h->stream()>>>(kgeom(h->g), v,
                                                                     h->solver.sigma0());
        h->solver.sigma_changed();
    });
}

MGX_API int mgx_nodal_project(mgx_nodal h, double* vel, int vel_ng, double rtol, double atol,
                              int mode, mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        check_mode(mode);
        MGX_REQUIRE(vel != nullptr, MGX_ERR_ARG, "null velocity");
        h->ensure_setup();
        const bool any_nonper = !h->periodic[0] || !h->periodic[1] || !h->periodic[2];
        MGX_REQUIRE(vel_ng >= (any_nonper ? 1 : 0), MGX_ERR_ARG,
                    "velocity needs one ghost cell at non-periodic domain faces");
        MGX_REQUIRE(h->solver.sigma_set, MGX_ERR_STATE, "nodal projector: sigma not set");
        GatherVelArgs a{};
        a.g = kgeom(h->g);
        a.boxes = h->layout.dev_boxes();
        a.tiles = h->layout.tiles(REG_CELL_GROWN);
        a.vel = vel;
        a.vel_ng = vel_ng;
        a.scale = 1.0;
        a.const_rho = 1.0;
        a.ug = h->ug.as<double>();
        a.sig = h->solver.sigma0();
        a.ghost_rule = 0;
        gather_vel_kernel<<<h->layout.ntiles[REG_CELL_GROWN], 256, 0, h->stream()>>>(a);
        h->compute_rhs();
        h->zero_phi();
        h->solver.solve(h->rhs.as<double>(), rtol, atol, mode, info, h->stream());
        UpdateArgs u = h->update_args();
        u.ug = h->ug.as<double>();
        u.vel = vel;
        u.vel_ng = vel_ng;
        update_cells_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(u);
        h->have_phi = true;
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "nodal projector: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

MGX_API int mgx_nodal_get_phi(mgx_nodal h, double* phi, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(h->have_phi, MGX_ERR_STATE, "getPhi before project");
        MGX_REQUIRE(phi != nullptr && ng >= 0, MGX_ERR_ARG, "bad phi field");
        scatter_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kgeom(h->g), h->layout.dev_boxes(), h->layout.tiles(REG_NODE), h->solver.phi, phi, ng,
            0);
    });
}

MGX_API int mgx_nodal_get_gradphi(mgx_nodal h, double* gphi, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(h->have_phi, MGX_ERR_STATE, "getGradPhi before project");
        MGX_REQUIRE(gphi != nullptr && ng >= 0, MGX_ERR_ARG, "bad gradphi field");
        UpdateArgs u = h->update_args();
        u.gp = gphi;
        u.gp_ng = ng;
        update_cells_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(u);
    });
}

MGX_API int mgx_nodal_apply(mgx_nodal h, const double* phi, int phi_ng, double* out, int out_ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(phi && out && phi_ng >= 0 && out_ng >= 0, MGX_ERR_ARG, "bad apply fields");
        MGX_REQUIRE(h->solver.sigma_set, MGX_ERR_STATE, "nodal apply: sigma not set");
        h->ensure_setup();
        const KGeom kg = kgeom(h->g);
        gather_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kg, h->layout.dev_boxes(), h->layout.tiles(REG_NODE), phi, phi_ng, h->rhs.as<double>());
        if (h->eb) {
            h->eb->set_sigma(h->solver.sigma0(), h->stream());
            EBArgs e{};
            e.L = h->eb->lev(0);
            e.x = h->rhs.as<double>();
            e.xout = h->solver.phi;
            eb_mask_kernel<<<grid1d(h->g.nnode()), 256, 0, h->stream()>>>(
                h->eb->mask0(), h->rhs.as<double>(), h->g.nnode());
            eb_stencil_kernel<<<grid3(h->g.nn[0], h->g.nn[1], h->g.nn[2]), block3(), 0,
                                h->stream()>>>(e, EB_APPLY);
        } else {
            SArgs<double, double> a{};
            a.g = kg;
            a.c = make_coef<double>(h->g.dx);
            a.x = h->rhs.as<double>();
            a.sig = h->solver.sigma0();
            a.rout = h->solver.phi;
            launch_stencil<double, double, KS_APPLY>(h->g, a, h->stream(), false);
        }
        scatter_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kg, h->layout.dev_boxes(), h->layout.tiles(REG_NODE), h->solver.phi, out, out_ng, 0);
        h->have_phi = false;
    });
}

MGX_API int mgx_nodal_solve(mgx_nodal h, double* phi, int phi_ng, const double* rhs, int rhs_ng,
                            double rtol, double atol, int mode, mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        check_mode(mode);
        MGX_REQUIRE(phi && rhs && phi_ng >= 0 && rhs_ng >= 0, MGX_ERR_ARG, "bad solve fields");
        h->ensure_setup();
        const KGeom kg = kgeom(h->g);
        gather_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kg, h->layout.dev_boxes(), h->layout.tiles(REG_NODE), rhs, rhs_ng, h->rhs.as<double>());
        gather_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kg, h->layout.dev_boxes(), h->layout.tiles(REG_NODE), phi, phi_ng, h->solver.phi);
        h->solver.solve(h->rhs.as<double>(), rtol, atol, mode, info, h->stream());
        scatter_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kg, h->layout.dev_boxes(), h->layout.tiles(REG_NODE), h->solver.phi, phi, phi_ng, 0);
        h->have_phi = true;
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "nodal MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

MGX_API int mgx_apply_nodal_projection(mgx_nodal h, const double* density, int rho_ng,
                                       double const_rho, double* vel, int vel_ng,
                                       const double* vel_old, int vel_old_ng, double* gp,
                                       int gp_ng, double* p_nd, int p_ng, double scaling_factor,
                                       int incremental, const double* inflow_vel, double rtol,
                                       double atol, int mode, mgx_solve_info* info) {
    if (!h) return MGX_ERR_ARG;
    int status = guarded(h->ctx, [&] {
        check_mode(mode);
        MGX_REQUIRE(vel && gp && p_nd, MGX_ERR_ARG, "velocity, gp and p_nd are required");
        MGX_REQUIRE(density != nullptr || const_rho > 0.0, MGX_ERR_ARG,
                    "need a density field or a positive constant density");
        MGX_REQUIRE(scaling_factor > 0.0, MGX_ERR_ARG, "scaling factor must be positive");
        h->ensure_setup();
        const bool any_nonper = !h->periodic[0] || !h->periodic[1] || !h->periodic[2];
        MGX_REQUIRE(vel_ng >= (any_nonper ? 1 : 0), MGX_ERR_ARG,
                    "velocity needs one ghost cell at non-periodic domain faces");
        GatherVelArgs a{};
        a.g = kgeom(h->g);
        a.boxes = h->layout.dev_boxes();
        a.tiles = h->layout.tiles(REG_CELL_GROWN);
        a.vel = vel;
        a.vel_ng = vel_ng;
        a.vel_old = vel_old;
        a.vold_ng = vel_old_ng;
        a.gp = incremental ? nullptr : gp;
        a.gp_ng = gp_ng;
        a.rho = density;
        a.rho_ng = rho_ng;
        a.const_rho = const_rho;
        a.scale = scaling_factor;
        a.sigma_from_rho = 1;
        a.ug = h->ug.as<double>();
        a.sig = h->solver.sigma0();
        a.ghost_rule = 1;
        for (int d = 0; d < 3; ++d) { a.bclo[d] = h->bclo[d]; a.bchi[d] = h->bchi[d]; }
        if (inflow_vel) for (int q = 0; q < 18; ++q) a.inflow[q] = inflow_vel[q];
        else for (int d = 0; d < 3; ++d) { a.bclo[d] = a.bchi[d] = MGX_BC_NEUMANN; }
        gather_vel_kernel<<<h->layout.ntiles[REG_CELL_GROWN], 256, 0, h->stream()>>>(a);
        h->solver.sigma_changed();
        h->compute_rhs();
        h->zero_phi();
        h->solver.solve(h->rhs.as<double>(), rtol, atol, mode, info, h->stream());
        UpdateArgs u = h->update_args();
        u.ug = h->ug.as<double>();
        u.vel = vel;
        u.vel_ng = vel_ng;
        u.vel_old = vel_old;
        u.vold_ng = vel_old_ng;
        u.gp = gp;
        u.gp_ng = gp_ng;
        u.gp_add = incremental ? 1 : 0;
        update_cells_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(u);
        scatter_nodes_kernel<<<h->layout.ntiles[REG_NODE], 256, 0, h->stream()>>>(
            kgeom(h->g), h->layout.dev_boxes(), h->layout.tiles(REG_NODE), h->solver.phi, p_nd,
            p_ng, incremental ? 1 : 0);
        h->have_phi = true;
    });
    if (status == MGX_OK && info && !info->converged) {
        h->ctx->last_error = "nodal projection: MLMG failed to converge";
        status = MGX_ERR_NOT_CONVERGED;
    }
    return status;
}

MGX_API int mgx_nodal_set_unet(mgx_nodal h, mgx_unet u, int level) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(u == nullptr || u->ctx == h->ctx, MGX_ERR_ARG,
                    "the network must belong to the projector's context");
        MGX_REQUIRE(level >= 0, MGX_ERR_ARG, "network level must be >= 0");
        h->solver.net = u ? &u->net : nullptr;
        h->solver.net_level = level;
    });
}

MGX_API int mgx_nodal_training_pair(mgx_nodal h, int level, float* in, float* target) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(in && target, MGX_ERR_ARG, "null output tensors");
        MGX_REQUIRE(h->have_phi, MGX_ERR_STATE, "training pair: no converged solve yet");
        h->solver.training_pair(h->rhs.as<double>(), level, in, target, h->stream());
    });
}

MGX_API int mgx_nodal_set_eb_cylinder(mgx_nodal h, double cx, double cy, double r) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(r > 0.0, MGX_ERR_ARG, "EB cylinder: radius must be positive");
        h->eb_kind = 1;
        h->eb_cyl[0] = cx;
        h->eb_cyl[1] = cy;
        h->eb_cyl[2] = r;
        if (!h->dirty) h->build_eb();
    });
}

MGX_API int mgx_nodal_set_eb_inflow_velocity(mgx_nodal h, const double* eb_vel, int ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        if (!eb_vel) {
            h->has_ueb = false;
            return;
        }
        MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, "EB inflow velocity: bad ghost width");
        MGX_REQUIRE(h->eb_kind == 1, MGX_ERR_UNSUPPORTED,
                    "EB inflow velocity: needs the cylinder geometry (its surface weights)");
        h->ensure_setup();
        const long nc = (long)h->n_cell[0] * h->n_cell[1] * h->n_cell[2];
        h->ueb.alloc(3 * nc * sizeof(double));
        gather_cells_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            kgeom(h->g), h->layout.dev_boxes(), h->layout.tiles(REG_CELL), eb_vel, ng, 3,
            h->ueb.as<double>());
        h->has_ueb = true;
    });
}

MGX_API int mgx_nodal_set_eb_moments(mgx_nodal h, const int* type, const double* moments) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        const long nc = (long)h->n_cell[0] * h->n_cell[1] * h->n_cell[2];
        if (!type) {
            h->eb_kind = 0;
        } else {
            MGX_REQUIRE(moments != nullptr, MGX_ERR_ARG, "EB moments: null moments");
            for (long c = 0; c < nc; ++c)
                MGX_REQUIRE(type[c] >= 0 && type[c] <= 2, MGX_ERR_ARG, "EB moments: bad cell type");
            h->eb_kind = 2;
            h->eb_type.assign(type, type + nc);
            h->eb_mom.assign(moments, moments + 27 * nc);
        }
        if (!h->dirty) h->build_eb();
    });
}

}  // extern "C"

extern "C" MGX_API int mgx_eb_cylinder_geometry(const mgx_level_desc* lev, double cx, double cy,
                                               double r, int* type, double* moments) {
    if (!lev || !type || !moments || r <= 0.0) return MGX_ERR_ARG;
    try {
        int per[3], bc[3] = {0, 0, 0};
        double dx[3];
        for (int d = 0; d < 3; ++d) {
            per[d] = lev->is_periodic[d];
            dx[d] = (lev->prob_hi[d] - lev->prob_lo[d]) / lev->n_cell[d];
        }
        NodalGeom g = make_nodal_geom(lev->n_cell, per, bc, bc, dx);
        std::vector<int> t;
        std::vector<double> m;
        eb_cylinder_moments(g, lev->prob_lo, cx, cy, r, t, m);
        std::memcpy(type, t.data(), t.size() * sizeof(int));
        std::memcpy(moments, m.data(), m.size() * sizeof(double));
        return MGX_OK;
    } catch (...) {
        return MGX_ERR_ARG;
    }
}

extern "C" MGX_API int mgx_eb_cylinder_surface_weights(const mgx_level_desc* lev, double cx,
                                                      double cy, double r, double* weights) {
    if (!lev || !weights || r <= 0.0) return MGX_ERR_ARG;
    try {
        int per[3], bc[3] = {0, 0, 0};
        double dx[3];
        for (int d = 0; d < 3; ++d) {
            per[d] = lev->is_periodic[d];
            dx[d] = (lev->prob_hi[d] - lev->prob_lo[d]) / lev->n_cell[d];
        }
        NodalGeom g = make_nodal_geom(lev->n_cell, per, bc, bc, dx);
        std::vector<int> t;
        std::vector<double> m, b;
        eb_cylinder_moments(g, lev->prob_lo, cx, cy, r, t, m);
        eb_cylinder_surface_weights(g, lev->prob_lo, cx, cy, r, t, b);
        std::memcpy(weights, b.data(), b.size() * sizeof(double));
        return MGX_OK;
    } catch (...) {
        return MGX_ERR_ARG;
    }
}

extern "C" MGX_API int mgx_eb_cylinder_apertures(const mgx_level_desc* lev, double cx, double cy,
                                                double r, double* ax, double* ay, double* az) {
    if (!lev || !ax || !ay || !az || r <= 0.0) return MGX_ERR_ARG;
    try {
        double dx[3];
        for (int d = 0; d < 3; ++d) dx[d] = (lev->prob_hi[d] - lev->prob_lo[d]) / lev->n_cell[d];
        std::vector<double> ap[3];
        std::vector<unsigned char> cov;
        eb_cylinder_apertures(lev->n_cell, lev->prob_lo, dx, cx, cy, r, ap, cov);
        double* out[3] = {ax, ay, az};
        for (int d = 0; d < 3; ++d) std::memcpy(out[d], ap[d].data(), ap[d].size() * sizeof(double));
        return MGX_OK;
    } catch (...) {
        return MGX_ERR_ARG;
    }
}

// adv_api.hip -- C ABI of the Godunov advection pieces on incflo's predictor
// path: ExtrapVelToFaces and the edge states / fluxes / advection terms of
// the velocity, density and tracers (single level, no EB; periodic or
// bounded domains with amrex::BCType records per component).
#include <algorithm>
#include <memory>
#include <vector>

#include "godunov_kernels.h"
#include "mgx_ctx.h"
#include "nodal_mg.h"

using namespace mgx;

struct mgx_adv_s {
    mgx_ctx_s* ctx = nullptr;
    GdGeom g{};
    BoxLayout layout;
    long N = 0;   // cells
    long NF = 0;  // entries of the largest face array
    long NG = 0;  // ghosted cells
    bool bounded = false;       // some direction is not periodic
    int vel_bc[3][6] = {};      // velocity BCRec (get_velocity_bcrec)
    bool vel_bc_set = false;
    DevBuf work;  // all stage arrays
    DevBuf qbuf, fbuf, abuf, cbuf, flbuf;
    DevBuf sbuf, sfbuf, fsbuf;  // scalars, their forces, their face states
    hipStream_t stream() const { return ctx->stream; }
    long nface(int d) const {
        return (long)(g.n[0] + (d == 0)) * (g.n[1] + (d == 1)) * (g.n[2] + (d == 2));
    }
};

namespace {

// 18 (Im, Ip) cell arrays; 3 (ad) + 9 (E) + 27 (T slots) + 9 (out) face arrays
constexpr int kCellArrays = 18, kFaceArrays = 3 + 9 + 27 + 9;

GdArrays bind(mgx_adv_s* h, const double* qg, const double* f, const double* adv, double dt,
              bool extrap, int nc, const int* cons, const int (*bc)[6]) {
    const long N = h->N, NF = h->NF, NG = h->NG;
    h->work.alloc(((size_t)kCellArrays * N + (size_t)kFaceArrays * NF) * sizeof(double));
    double* w = h->work.as<double>();
    GdArrays a{};
    for (int n = 0; n < 3; ++n) {
        a.q[n] = n < nc ? qg + n * NG : nullptr;
        a.f[n] = (f && n < nc) ? f + n * N : nullptr;
        a.adv[n] = adv ? adv + n * NF : nullptr;
        a.cons[n] = (cons && n < nc) ? (cons[n] ? 1 : 0) : 0;
        for (int q = 0; q < 6; ++q) a.bc[n][q] = (n < nc && bc) ? bc[n][q] : GBC_INT;
    }
    a.nc = nc;
    long s = 0;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) {
            a.Im[d][n] = w + (s++) * N;
            a.Ip[d][n] = w + (s++) * N;
        }
    double* wf = w + kCellArrays * N;
    s = 0;
    for (int d = 0; d < 3; ++d) a.ad[d] = wf + (s++) * NF;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) a.E[d][n] = wf + (s++) * NF;
    for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e)
            for (int n = 0; n < 3; ++n) a.T[d][e][n] = wf + (s++) * NF;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) a.out[d][n] = wf + (s++) * NF;
    a.dt = dt;
    a.extrap = extrap ? 1 : 0;
    return a;
}

void run_stages(mgx_adv_s* h, const GdArrays& a) {
    const dim3 gc = grid3(h->g.n[0], h->g.n[1], h->g.n[2]);
    const dim3 gf = grid3(h->g.n[0] + 1, h->g.n[1] + 1, h->g.n[2] + 1), bl = block3();
    hipStream_t s = h->stream();
    gd_ppm_kernel<<<gc, bl, 0, s>>>(h->g, a);
    gd_edge_kernel<<<gf, bl, 0, s>>>(h->g, a);
    gd_trans_kernel<<<gf, bl, 0, s>>>(h->g, a);
    gd_final_kernel<<<gf, bl, 0, s>>>(h->g, a);
}

// ghosted canonical cells (2 layers) from a boxed field
void cells_in_ghosted(mgx_adv_s* h, const double* boxed, int ng, int ncomp, DevBuf& dst) {
    dst.alloc((size_t)ncomp * h->NG * sizeof(double));
    gd_cells_in_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
        h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), boxed, ng, ncomp, dst.as<double>());
    gd_ghosts_kernel<<<grid3(h->g.n[0] + 2 * kGh, h->g.n[1] + 2 * kGh, h->g.n[2] + 2 * kGh), block3(),
                       0, h->stream()>>>(h->g, h->layout.dev_boxes(), h->layout.nboxes, boxed, ng,
                                         ncomp, dst.as<double>());
}

void cells_in(mgx_adv_s* h, const double* boxed, int ng, int ncomp, DevBuf& dst) {
    dst.alloc((size_t)ncomp * h->N * sizeof(double));
    gd_plain_in_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
        h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), boxed, ng, ncomp, dst.as<double>());
}

void faces_in(mgx_adv_s* h, const double* const src[3], int ng) {
    h->abuf.alloc(3 * h->NF * sizeof(double));
    for (int d = 0; d < 3; ++d)
        gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
            h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d,
            const_cast<double*>(src[d]), ng, 1, h->abuf.as<double>() + d * h->NF, h->NF, 0);
}

void check_ng(int ng, const char* what) {
    MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, std::string(what) + ": ghost width out of range");
}

// state ghost cells: 2 layers where a direction is not periodic (PPM's stencil)
void check_state_ng(const mgx_adv_s* h, int ng, const char* what) {
    check_ng(ng, what);
    MGX_REQUIRE(!h->bounded || ng >= 2, MGX_ERR_ARG,
                std::string(what) + ": bounded domains need 2 filled ghost cells");
}

// a BC record: periodic directions must say so, the others take a physical type
void check_bcrec(const mgx_adv_s* h, const int* bc, int ncomp, const char* what) {
    for (int n = 0; n < ncomp; ++n)
        for (int side = 0; side < 2; ++side)
            for (int d = 0; d < 3; ++d) {
                const int t = bc[6 * n + 3 * side + d];
                if (h->g.per[d])
                    MGX_REQUIRE(t == GBC_INT, MGX_ERR_ARG,
                                std::string(what) + ": periodic directions take BCType int_dir (0)");
                else
                    MGX_REQUIRE(t == GBC_EXT || t == GBC_FOE || t == GBC_HOE || t == GBC_REVEN ||
                                    t == GBC_RODD,
                                MGX_ERR_ARG,
                                std::string(what) + ": unknown BCType on a domain face");
            }
}

}  // namespace

extern "C" {

MGX_API int mgx_adv_create(mgx_ctx ctx, const mgx_level_desc* lev, mgx_adv* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        MGX_REQUIRE(lev && lev->boxes && lev->nboxes > 0, MGX_ERR_ARG, "advection: bad level");
        auto h = std::make_unique<mgx_adv_s>();
        h->ctx = ctx;
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(lev->n_cell[d] >= 4, MGX_ERR_ARG, "advection: at least 4 cells per direction");
            h->g.n[d] = lev->n_cell[d];
            h->g.per[d] = lev->is_periodic[d] ? 1 : 0;
            h->g.dx[d] = (lev->prob_hi[d] - lev->prob_lo[d]) / lev->n_cell[d];
            if (!h->g.per[d]) h->bounded = true;
        }
        h->N = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        h->NF = (long)(h->g.n[0] + 1) * (h->g.n[1] + 1) * (h->g.n[2] + 1);
        h->NG = (long)(h->g.n[0] + 2 * kGh) * (h->g.n[1] + 2 * kGh) * (h->g.n[2] + 2 * kGh);
        int per[3] = {h->g.per[0], h->g.per[1], h->g.per[2]}, bc[3] = {0, 0, 0};
        NodalGeom ng = make_nodal_geom(lev->n_cell, per, bc, bc, h->g.dx);
        h->layout.build(lev->boxes, lev->nboxes, ng);
        // periodic directions: int_dir; a bounded domain must set its record
        for (int n = 0; n < 3; ++n)
            for (int q = 0; q < 6; ++q) h->vel_bc[n][q] = GBC_INT;
        h->vel_bc_set = !h->bounded;
        *out = h.release();
    });
}

MGX_API int mgx_adv_destroy(mgx_adv h) {
    if (!h) return MGX_OK;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_adv_set_velocity_bcrec(mgx_adv h, const int* bcrec) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(bcrec != nullptr, MGX_ERR_ARG, "velocity BC record: null");
        check_bcrec(h, bcrec, 3, "velocity BC record");
        for (int n = 0; n < 3; ++n)
            for (int q = 0; q < 6; ++q) h->vel_bc[n][q] = bcrec[6 * n + q];
        h->vel_bc_set = true;
    });
}

MGX_API int mgx_adv_extrap_vel_to_faces(mgx_adv h, const double* vel, int vel_ng,
                                        const double* force, int f_ng, double dt, double* umac,
                                        double* vmac, double* wmac, int umac_ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(vel && umac && vmac && wmac && dt > 0.0, MGX_ERR_ARG,
                    "ExtrapVelToFaces: null field or non-positive dt");
        MGX_REQUIRE(h->vel_bc_set, MGX_ERR_STATE,
                    "ExtrapVelToFaces: a bounded domain needs the velocity BC record");
        check_state_ng(h, vel_ng, "velocity");
        check_ng(f_ng, "forces");
        check_ng(umac_ng, "umac");
        cells_in_ghosted(h, vel, vel_ng, 3, h->qbuf);
        if (force) cells_in(h, force, f_ng, 3, h->fbuf);
        GdArrays a = bind(h, h->qbuf.as<double>(), force ? h->fbuf.as<double>() : nullptr, nullptr,
                          dt, true, 3, nullptr, h->vel_bc);
        run_stages(h, a);
        double* dst[3] = {umac, vmac, wmac};
        for (int d = 0; d < 3; ++d)
            gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, dst[d], umac_ng, 1,
                a.out[d][d], h->NF, 1);
        MGX_HIP(hipStreamSynchronize(h->stream()));
    });
}

MGX_API int mgx_adv_convective_term(mgx_adv h, const double* vel, int vel_ng, const double* force,
                                    int f_ng, const double* umac, const double* vmac,
                                    const double* wmac, int umac_ng, double dt, double* conv,
                                    int conv_ng, double* face_x, double* face_y, double* face_z,
                                    double* flux_x, double* flux_y, double* flux_z) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(vel && umac && vmac && wmac && conv && dt > 0.0, MGX_ERR_ARG,
                    "convective term: null field or non-positive dt");
        MGX_REQUIRE(h->vel_bc_set, MGX_ERR_STATE,
                    "convective term: a bounded domain needs the velocity BC record");
        check_state_ng(h, vel_ng, "velocity");
        check_ng(f_ng, "forces");
        check_ng(umac_ng, "umac");
        check_ng(conv_ng, "conv");
        const long N = h->N;
        cells_in_ghosted(h, vel, vel_ng, 3, h->qbuf);
        if (force) cells_in(h, force, f_ng, 3, h->fbuf);
        const double* src[3] = {umac, vmac, wmac};
        faces_in(h, src, umac_ng);
        GdArrays a = bind(h, h->qbuf.as<double>(), force ? h->fbuf.as<double>() : nullptr,
                          h->abuf.as<double>(), dt, false, 3, nullptr, h->vel_bc);
        run_stages(h, a);
        h->cbuf.alloc(3 * N * sizeof(double));
        double* faces[3] = {face_x, face_y, face_z};
        double* fluxes[3] = {flux_x, flux_y, flux_z};
        GdFluxOut fo{};
        const bool want_flux = flux_x || flux_y || flux_z;
        if (want_flux) {
            h->flbuf.alloc(3 * 3 * h->NF * sizeof(double));
            for (int d = 0; d < 3; ++d) fo.f[d] = h->flbuf.as<double>() + 3 * d * h->NF;
        }
        gd_conv_kernel<<<grid3(h->g.n[0] + 1, h->g.n[1] + 1, h->g.n[2] + 1), block3(), 0,
                         h->stream()>>>(h->g, a, h->cbuf.as<double>(), fo, 0);
        gd_cells_out_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), conv, conv_ng, 3,
            h->cbuf.as<double>());
        for (int d = 0; d < 3; ++d) {
            // out[d][0..2] are consecutive face arrays in the work buffer
            if (faces[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, faces[d], 0, 3,
                    a.out[d][0], h->NF, 1);
            if (fluxes[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, fluxes[d], 0, 3,
                    fo.f[d], h->nface(d), 1);
        }
        MGX_HIP(hipStreamSynchronize(h->stream()));
    });
}

MGX_API int mgx_adv_scalar_term(mgx_adv h, const double* state, int s_ng, int ncomp,
                                const int* iconserv, const int* bcrec, const double* force,
                                int f_ng, const double* umac, const double* vmac,
                                const double* wmac, int umac_ng, double dt, double* conv,
                                int conv_ng, double* face_x, double* face_y, double* face_z,
                                double* flux_x, double* flux_y, double* flux_z) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(state && umac && vmac && wmac && conv && dt > 0.0 && ncomp >= 1 && ncomp <= 64,
                    MGX_ERR_ARG, "scalar advection: null field, bad component count or dt");
        MGX_REQUIRE(bcrec != nullptr || !h->bounded, MGX_ERR_ARG,
                    "scalar advection: a bounded domain needs the BC records");
        check_state_ng(h, s_ng, "state");
        check_ng(f_ng, "forces");
        check_ng(umac_ng, "umac");
        check_ng(conv_ng, "conv");
        std::vector<int> bcv((size_t)6 * ncomp, GBC_INT);
        if (bcrec) {
            check_bcrec(h, bcrec, ncomp, "scalar BC record");
            bcv.assign(bcrec, bcrec + 6 * ncomp);
        }
        const long N = h->N;
        cells_in_ghosted(h, state, s_ng, ncomp, h->sbuf);
        if (force) cells_in(h, force, f_ng, ncomp, h->sfbuf);
        const double* src[3] = {umac, vmac, wmac};
        faces_in(h, src, umac_ng);
        h->cbuf.alloc((size_t)ncomp * N * sizeof(double));
        double* faces[3] = {face_x, face_y, face_z};
        double* fluxes[3] = {flux_x, flux_y, flux_z};
        const bool want_face = face_x || face_y || face_z;
        const bool want_flux = flux_x || flux_y || flux_z;
        if (want_face) h->fsbuf.alloc((size_t)3 * ncomp * h->NF * sizeof(double));
        GdFluxOut fo{};
        if (want_flux) {
            h->flbuf.alloc((size_t)3 * ncomp * h->NF * sizeof(double));
            for (int d = 0; d < 3; ++d) fo.f[d] = h->flbuf.as<double>() + (long)d * ncomp * h->NF;
        }
        // the work arrays hold three components at a time
        for (int n0 = 0; n0 < ncomp; n0 += 3) {
            const int nc = std::min(3, ncomp - n0);
            int cons[3] = {0, 0, 0};
            int bc[3][6];
            for (int q = 0; q < nc; ++q) {
                cons[q] = iconserv ? iconserv[n0 + q] : 0;
                for (int r = 0; r < 6; ++r) bc[q][r] = bcv[6 * (n0 + q) + r];
            }
            GdArrays a = bind(h, h->sbuf.as<double>() + n0 * h->NG,
                              force ? h->sfbuf.as<double>() + n0 * N : nullptr,
                              h->abuf.as<double>(), dt, false, nc, cons, bc);
            run_stages(h, a);
            gd_conv_kernel<<<grid3(h->g.n[0] + 1, h->g.n[1] + 1, h->g.n[2] + 1), block3(), 0,
                             h->stream()>>>(h->g, a, h->cbuf.as<double>(), fo, n0);
            if (want_face)
                for (int d = 0; d < 3; ++d)
                    for (int q = 0; q < nc; ++q)
                        MGX_HIP(hipMemcpyAsync(h->fsbuf.as<double>() + (long)(d * ncomp + n0 + q) * h->NF,
                                               a.out[d][q], h->nface(d) * sizeof(double),
                                               hipMemcpyDeviceToDevice, h->stream()));
        }
        gd_cells_out_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), conv, conv_ng, ncomp,
            h->cbuf.as<double>());
        for (int d = 0; d < 3; ++d) {
            if (faces[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, faces[d], 0,
                    ncomp, h->fsbuf.as<double>() + (long)d * ncomp * h->NF, h->NF, 1);
            if (fluxes[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, fluxes[d], 0,
                    ncomp, fo.f[d], h->nface(d), 1);
        }
        MGX_HIP(hipStreamSynchronize(h->stream()));
    });
}

}  // extern "C"

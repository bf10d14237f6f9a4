// adv_api.hip -- C ABI of the Godunov advection pieces on incflo's predictor
// path: ExtrapVelToFaces and the velocity's edge states / fluxes /
// convective term (periodic, single level, no EB).
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
    KGeom kg{};
    BoxLayout layout;
    long N = 0;
    DevBuf work;  // all stage arrays
    DevBuf qbuf, fbuf, abuf, obuf, cbuf, flbuf;
    DevBuf sbuf, sfbuf, fsbuf;  // scalars, their forces, their face states
    hipStream_t stream() const { return ctx->stream; }
};

namespace {

// 18 (Im, Ip) + 3 (ad) + 9 (E) + 27 (T slots) + 9 (out) arrays of N doubles
constexpr int kWorkArrays = 18 + 3 + 9 + 27 + 9;

GdArrays bind(mgx_adv_s* h, const double* q, const double* f, const double* adv, double dt,
              bool extrap, int nc = 3, const int* cons = nullptr) {
    const long N = h->N;
    h->work.alloc((size_t)kWorkArrays * N * sizeof(double));
    double* w = h->work.as<double>();
    GdArrays a{};
    for (int n = 0; n < 3; ++n) {
        a.q[n] = n < nc ? q + n * N : nullptr;
        a.f[n] = (f && n < nc) ? f + n * N : nullptr;
        a.adv[n] = adv ? adv + n * N : nullptr;
        a.cons[n] = (cons && n < nc) ? (cons[n] ? 1 : 0) : 0;
    }
    a.nc = nc;
    long s = 0;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) {
            a.Im[d][n] = w + (s++) * N;
            a.Ip[d][n] = w + (s++) * N;
        }
    for (int d = 0; d < 3; ++d) a.ad[d] = w + (s++) * N;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) a.E[d][n] = w + (s++) * N;
    for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e)
            for (int n = 0; n < 3; ++n) a.T[d][e][n] = w + (s++) * N;
    for (int d = 0; d < 3; ++d)
        for (int n = 0; n < 3; ++n) a.out[d][n] = w + (s++) * N;
    a.dt = dt;
    a.extrap = extrap ? 1 : 0;
    return a;
}

void run_stages(mgx_adv_s* h, const GdArrays& a) {
    const dim3 gr = grid3(h->g.n[0], h->g.n[1], h->g.n[2]), bl = block3();
    hipStream_t s = h->stream();
    gd_ppm_kernel<<<gr, bl, 0, s>>>(h->g, a);
    gd_edge_kernel<<<gr, bl, 0, s>>>(h->g, a);
    gd_trans_kernel<<<gr, bl, 0, s>>>(h->g, a);
    gd_final_kernel<<<gr, bl, 0, s>>>(h->g, a);
}

void cells_in(mgx_adv_s* h, const double* boxed, int ng, int ncomp, DevBuf& dst) {
    dst.alloc((size_t)ncomp * h->N * sizeof(double));
    gd_cells_in_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
        h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), boxed, ng, ncomp, dst.as<double>());
}

void check_ng(int ng, const char* what) {
    MGX_REQUIRE(ng >= 0 && ng <= 8, MGX_ERR_ARG, std::string(what) + ": ghost width out of range");
}

}  // namespace

extern "C" {

MGX_API int mgx_adv_create(mgx_ctx ctx, const mgx_level_desc* lev, mgx_adv* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        MGX_REQUIRE(lev && lev->boxes && lev->nboxes > 0, MGX_ERR_ARG, "advection: bad level");
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(lev->n_cell[d] >= 4, MGX_ERR_ARG, "advection: at least 4 cells per direction");
            MGX_REQUIRE(lev->is_periodic[d], MGX_ERR_UNSUPPORTED,
                        "advection: only periodic domains are supported");
        }
        auto h = std::make_unique<mgx_adv_s>();
        h->ctx = ctx;
        for (int d = 0; d < 3; ++d) {
            h->g.n[d] = lev->n_cell[d];
            h->g.dx[d] = (lev->prob_hi[d] - lev->prob_lo[d]) / lev->n_cell[d];
        }
        h->N = (long)h->g.n[0] * h->g.n[1] * h->g.n[2];
        int per[3] = {1, 1, 1}, bc[3] = {0, 0, 0};
        NodalGeom ng = make_nodal_geom(lev->n_cell, per, bc, bc, h->g.dx);
        h->layout.build(lev->boxes, lev->nboxes, ng);
        *out = h.release();
    });
}

MGX_API int mgx_adv_destroy(mgx_adv h) {
    if (!h) return MGX_OK;
    h->ctx->make_current();
    delete h;
    return MGX_OK;
}

MGX_API int mgx_adv_extrap_vel_to_faces(mgx_adv h, const double* vel, int vel_ng,
                                        const double* force, int f_ng, double dt, double* umac,
                                        double* vmac, double* wmac, int umac_ng) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(vel && umac && vmac && wmac && dt > 0.0, MGX_ERR_ARG,
                    "ExtrapVelToFaces: null field or non-positive dt");
        check_ng(vel_ng, "velocity");
        check_ng(f_ng, "forces");
        check_ng(umac_ng, "umac");
        cells_in(h, vel, vel_ng, 3, h->qbuf);
        if (force) cells_in(h, force, f_ng, 3, h->fbuf);
        GdArrays a = bind(h, h->qbuf.as<double>(), force ? h->fbuf.as<double>() : nullptr, nullptr,
                          dt, true);
        run_stages(h, a);
        double* dst[3] = {umac, vmac, wmac};
        for (int d = 0; d < 3; ++d)
            gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, dst[d], umac_ng, 1,
                a.out[d][d], h->N, 1);
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
        check_ng(vel_ng, "velocity");
        check_ng(f_ng, "forces");
        check_ng(umac_ng, "umac");
        check_ng(conv_ng, "conv");
        const long N = h->N;
        cells_in(h, vel, vel_ng, 3, h->qbuf);
        if (force) cells_in(h, force, f_ng, 3, h->fbuf);
        h->abuf.alloc(3 * N * sizeof(double));
        const double* src[3] = {umac, vmac, wmac};
        for (int d = 0; d < 3; ++d)
            gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d,
                const_cast<double*>(src[d]), umac_ng, 1, h->abuf.as<double>() + d * N, N, 0);
        GdArrays a = bind(h, h->qbuf.as<double>(), force ? h->fbuf.as<double>() : nullptr,
                          h->abuf.as<double>(), dt, false);
        run_stages(h, a);
        h->cbuf.alloc(3 * N * sizeof(double));
        const bool want_flux = flux_x || flux_y || flux_z;
        if (want_flux) h->flbuf.alloc(9 * N * sizeof(double));
        gd_conv_kernel<<<grid3(h->g.n[0], h->g.n[1], h->g.n[2]), block3(), 0, h->stream()>>>(
            h->g, a, h->cbuf.as<double>(), want_flux ? h->flbuf.as<double>() : nullptr, 0, 3);
        gd_cells_out_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), conv, conv_ng, 3,
            h->cbuf.as<double>());
        double* faces[3] = {face_x, face_y, face_z};
        double* fluxes[3] = {flux_x, flux_y, flux_z};
        for (int d = 0; d < 3; ++d) {
            // the 3 components of direction d are contiguous: out[d][0..2] / flux slots 3d..3d+2
            if (faces[d]) {
                // out[d][n] arrays are consecutive in the work buffer
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, faces[d], 0, 3,
                    a.out[d][0], N, 1);
            }
            if (fluxes[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, fluxes[d], 0, 3,
                    h->flbuf.as<double>() + 3 * d * N, N, 1);
        }
        MGX_HIP(hipStreamSynchronize(h->stream()));
    });
}

MGX_API int mgx_adv_scalar_term(mgx_adv h, const double* state, int s_ng, int ncomp,
                                const int* iconserv, const double* force, int f_ng,
                                const double* umac, const double* vmac, const double* wmac,
                                int umac_ng, double dt, double* conv, int conv_ng, double* face_x,
                                double* face_y, double* face_z, double* flux_x, double* flux_y,
                                double* flux_z) {
    if (!h) return MGX_ERR_ARG;
    return guarded(h->ctx, [&] {
        MGX_REQUIRE(state && umac && vmac && wmac && conv && dt > 0.0 && ncomp >= 1 && ncomp <= 64,
                    MGX_ERR_ARG, "scalar advection: null field, bad component count or dt");
        check_ng(s_ng, "state");
        check_ng(f_ng, "forces");
        check_ng(umac_ng, "umac");
        check_ng(conv_ng, "conv");
        const long N = h->N;
        h->sbuf.alloc((size_t)ncomp * N * sizeof(double));
        gd_cells_in_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), state, s_ng, ncomp,
            h->sbuf.as<double>());
        if (force) {
            h->sfbuf.alloc((size_t)ncomp * N * sizeof(double));
            gd_cells_in_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
                h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), force, f_ng, ncomp,
                h->sfbuf.as<double>());
        }
        h->abuf.alloc(3 * N * sizeof(double));
        const double* src[3] = {umac, vmac, wmac};
        for (int d = 0; d < 3; ++d)
            gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d,
                const_cast<double*>(src[d]), umac_ng, 1, h->abuf.as<double>() + d * N, N, 0);
        h->cbuf.alloc((size_t)ncomp * N * sizeof(double));
        double* faces[3] = {face_x, face_y, face_z};
        double* fluxes[3] = {flux_x, flux_y, flux_z};
        const bool want_face = face_x || face_y || face_z;
        const bool want_flux = flux_x || flux_y || flux_z;
        if (want_face) h->fsbuf.alloc((size_t)3 * ncomp * N * sizeof(double));
        if (want_flux) h->flbuf.alloc((size_t)3 * ncomp * N * sizeof(double));
        // the work arrays hold three components at a time
        for (int n0 = 0; n0 < ncomp; n0 += 3) {
            const int nc = std::min(3, ncomp - n0);
            int cons[3] = {0, 0, 0};
            for (int q = 0; q < nc; ++q) cons[q] = iconserv ? iconserv[n0 + q] : 0;
            GdArrays a = bind(h, h->sbuf.as<double>() + n0 * N,
                              force ? h->sfbuf.as<double>() + n0 * N : nullptr,
                              h->abuf.as<double>(), dt, false, nc, cons);
            run_stages(h, a);
            gd_conv_kernel<<<grid3(h->g.n[0], h->g.n[1], h->g.n[2]), block3(), 0, h->stream()>>>(
                h->g, a, h->cbuf.as<double>(), want_flux ? h->flbuf.as<double>() : nullptr, n0,
                ncomp);
            if (want_face)
                for (int d = 0; d < 3; ++d)
                    for (int q = 0; q < nc; ++q)
                        MGX_HIP(hipMemcpyAsync(h->fsbuf.as<double>() + (long)(d * ncomp + n0 + q) * N,
                                               a.out[d][q], N * sizeof(double),
                                               hipMemcpyDeviceToDevice, h->stream()));
        }
        gd_cells_out_kernel<<<h->layout.ntiles[REG_CELL], 256, 0, h->stream()>>>(
            h->g, h->layout.dev_boxes(), h->layout.tiles(REG_CELL), conv, conv_ng, ncomp,
            h->cbuf.as<double>());
        for (int d = 0; d < 3; ++d) {
            if (faces[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, faces[d], 0,
                    ncomp, h->fsbuf.as<double>() + (long)d * ncomp * N, N, 1);
            if (fluxes[d])
                gd_faces_kernel<<<h->layout.ntiles[REG_FACE0 + d], 256, 0, h->stream()>>>(
                    h->g, h->layout.dev_boxes(), h->layout.tiles(REG_FACE0 + d), d, fluxes[d], 0,
                    ncomp, h->flbuf.as<double>() + (long)d * ncomp * N, N, 1);
        }
        MGX_HIP(hipStreamSynchronize(h->stream()));
    });
}

}  // extern "C"

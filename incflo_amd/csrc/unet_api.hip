// unet_api.hip -- C ABI of the learned coarse-level predictor and of the
// BF16-MFMA 3-D convolution operators it is built from.
#include <memory>
#include <vector>

#include "mgx_ctx.h"
#include "unet.h"

using namespace mgx;

static NetShape make_shape(int nx, int ny, int nz, int batch, const int* per) {
    NetShape sh{};
    sh.nx = nx;
    sh.ny = ny;
    sh.nz = nz;
    sh.batch = batch;
    for (int d = 0; d < 3; ++d) sh.per[d] = per ? (per[d] ? 1 : 0) : 1;
    return sh;
}

namespace {

__global__ void to_bf16_padded(const float* __restrict__ w, int cout, int K, int kp,
                               unsigned short* __restrict__ wb) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= cout * kp) return;
    const int co = id / kp, k = id % kp;
    wb[id] = k < K ? bf16_bits(w[(long)co * K + k]) : (unsigned short)0;
}

void conv_launch(const ConvGeom& cg, const float* in, const unsigned short* w, const float* bias,
                 float* out, int relu, hipStream_t s) {
    conv3d_forward_launch(cg, in, w, bias, out, relu, nullptr, 0, s);
}

int pad32(int k) { return (k + 31) / 32 * 32; }

ConvGeom conv_geom(int nx, int ny, int nz, int batch, int cin, int cout, const int* per) {
    MGX_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && batch >= 1, MGX_ERR_ARG, "conv3d: bad sizes");
    MGX_REQUIRE(cin % 8 == 0 && cin >= 8 && cout % 16 == 0 && cout >= 16, MGX_ERR_ARG,
                "conv3d: cin must be a multiple of 8, cout a multiple of 16");
    ConvGeom cg{};
    cg.nx = nx;
    cg.ny = ny;
    cg.nz = nz;
    cg.batch = batch;
    for (int d = 0; d < 3; ++d) cg.per[d] = per ? (per[d] ? 1 : 0) : 1;
    cg.cin = cin;
    cg.cout = cout;
    cg.kp = pad32(27 * cin);
    return cg;
}

}  // namespace

extern "C" {

MGX_API int mgx_unet_create(mgx_ctx ctx, int base_channels, int levels, unsigned long long seed,
                            mgx_unet* out) {
    if (!ctx || !out) return MGX_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto u = std::make_unique<mgx_unet_s>();
        u->ctx = ctx;
        u->net.create(base_channels, levels, seed, ctx->stream);
        *out = u.release();
    });
}

MGX_API int mgx_unet_destroy(mgx_unet u) {
    if (!u) return MGX_OK;
    u->ctx->make_current();
    delete u;
    return MGX_OK;
}

MGX_API int mgx_unet_num_params(mgx_unet u, long long* n) {
    if (!u || !n) return MGX_ERR_ARG;
    *n = u->net.nparam;
    return MGX_OK;
}

MGX_API int mgx_unet_get_params(mgx_unet u, float* host, long long n) {
    if (!u) return MGX_ERR_ARG;
    return guarded(u->ctx, [&] {
        MGX_REQUIRE(host && n == u->net.nparam, MGX_ERR_ARG, "parameter count mismatch");
        MGX_HIP(hipMemcpyAsync(host, u->net.params.p, n * sizeof(float), hipMemcpyDeviceToHost,
                               u->ctx->stream));
        MGX_HIP(hipStreamSynchronize(u->ctx->stream));
    });
}

MGX_API int mgx_unet_set_params(mgx_unet u, const float* host, long long n) {
    if (!u) return MGX_ERR_ARG;
    return guarded(u->ctx, [&] {
        MGX_REQUIRE(host && n == u->net.nparam, MGX_ERR_ARG, "parameter count mismatch");
        MGX_HIP(hipMemcpyAsync(u->net.params.p, host, n * sizeof(float), hipMemcpyHostToDevice,
                               u->ctx->stream));
        MGX_HIP(hipStreamSynchronize(u->ctx->stream));
        u->net.bf_stale = true;
    });
}

MGX_API int mgx_unet_forward(mgx_unet u, const float* in, float* out, int nx, int ny, int nz,
                             int batch, const int periodic[3]) {
    if (!u) return MGX_ERR_ARG;
    return guarded(u->ctx, [&] {
        MGX_REQUIRE(in && out, MGX_ERR_ARG, "null tensors");
        u->net.forward(in, make_shape(nx, ny, nz, batch, periodic), out);
    });
}

MGX_API int mgx_unet_gradients(mgx_unet u, const float* in, const float* dout, int nx, int ny,
                               int nz, int batch, const int periodic[3], float* host_grads,
                               long long n) {
    if (!u) return MGX_ERR_ARG;
    return guarded(u->ctx, [&] {
        MGX_REQUIRE(in && dout && host_grads && n == u->net.nparam, MGX_ERR_ARG, "bad arguments");
        const NetShape sh = make_shape(nx, ny, nz, batch, periodic);
        u->net.ybuf.alloc((long)nx * ny * nz * batch * sizeof(float));
        u->net.forward(in, sh, u->net.ybuf.as<float>());
        u->net.zero_grads();
        u->net.backward(in, dout, sh);
        MGX_HIP(hipMemcpyAsync(host_grads, u->net.grads.p, n * sizeof(float),
                               hipMemcpyDeviceToHost, u->ctx->stream));
        MGX_HIP(hipStreamSynchronize(u->ctx->stream));
    });
}

MGX_API int mgx_unet_train_step(mgx_unet u, const float* in, const float* target, int nx, int ny,
                                int nz, int batch, const int periodic[3], float lr, float* loss) {
    if (!u) return MGX_ERR_ARG;
    return guarded(u->ctx, [&] {
        MGX_REQUIRE(in && target && lr > 0.f, MGX_ERR_ARG, "bad arguments");
        const float l = u->net.train_step(in, target, make_shape(nx, ny, nz, batch, periodic), lr);
        if (loss) *loss = l;
    });
}

MGX_API int mgx_conv3d_forward(mgx_ctx ctx, const float* in, const float* w, const float* bias,
                               float* out, int nx, int ny, int nz, int batch, int cin, int cout,
                               const int periodic[3], int relu) {
    return guarded(ctx, [&] {
        MGX_REQUIRE(in && w && out, MGX_ERR_ARG, "null tensors");
        const ConvGeom cg = conv_geom(nx, ny, nz, batch, cin, cout, periodic);
        DevBuf wb;
        wb.alloc((size_t)cout * cg.kp * sizeof(unsigned short));
        to_bf16_padded<<<(cout * cg.kp + 255) / 256, 256, 0, ctx->stream>>>(w, cout, 27 * cin, cg.kp,
                                                                          wb.as<unsigned short>());
        conv_launch(cg, in, wb.as<unsigned short>(), bias, out, relu, ctx->stream);
        MGX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

MGX_API int mgx_conv3d_backward(mgx_ctx ctx, const float* in, const float* w, const float* dout,
                                float* din, float* dw, float* db, int nx, int ny, int nz,
                                int batch, int cin, int cout, const int periodic[3]) {
    return guarded(ctx, [&] {
        MGX_REQUIRE(in && w && dout, MGX_ERR_ARG, "null tensors");
        const ConvGeom cg = conv_geom(nx, ny, nz, batch, cin, cout, periodic);
        const long nvox = (long)nx * ny * nz * batch;
        hipStream_t s = ctx->stream;
        DevBuf ws;
        if (dw) {
            MGX_HIP(hipMemsetAsync(dw, 0, (size_t)cout * 27 * cin * sizeof(float), s));
            ws.alloc(conv_dw_workspace(cg) * sizeof(float));
            conv_dw_launch(cg, in, dout, dw, 27 * cin, ws.as<float>(), s);
        }
        if (db) {
            MGX_HIP(hipMemsetAsync(db, 0, cout * sizeof(float), s));
            ws.alloc((size_t)bias_grad_blocks(nvox, cout) * cout * sizeof(float));
            bias_grad_launch(dout, nvox, cout, db, ws.as<float>(), s);
        }
        if (din) {
            MGX_REQUIRE(cin % 16 == 0, MGX_ERR_ARG, "conv3d backward-data: cin must be a multiple of 16");
            DevBuf wb, wt;
            wb.alloc((size_t)cout * cg.kp * sizeof(unsigned short));
            to_bf16_padded<<<(cout * cg.kp + 255) / 256, 256, 0, s>>>(w, cout, 27 * cin, cg.kp,
                                                                    wb.as<unsigned short>());
            const int kpt = pad32(27 * cout);
            wt.alloc((size_t)cin * kpt * sizeof(unsigned short));
            weight_transpose_kernel<<<(cin * kpt + 255) / 256, 256, 0, s>>>(
                wb.as<unsigned short>(), cin, cout, cg.kp, wt.as<unsigned short>(), kpt);
            ConvGeom ct = cg;
            ct.cin = cout;
            ct.cout = cin;
            ct.kp = kpt;
            conv_launch(ct, dout, wt.as<unsigned short>(), nullptr, din, 0, s);
            MGX_HIP(hipStreamSynchronize(s));
        }
        MGX_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"

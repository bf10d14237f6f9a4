// unet.hip -- 3-D conv U-net (encoder/decoder with skip connections) for the
// learned coarse-level correction of the nodal projection, trained and run
// with the BF16-MFMA kernels of unet_kernels.h.
//
// Architecture (levels = 4, base = 16): channels C_l = base * 2^l.
//   encoder l = 0..2:  conv(C_{l-1} -> C_l), conv(C_l -> C_l), 2x2x2 average pool
//   bottleneck:        conv(C_2 -> C_3), conv(C_3 -> C_3)
//   decoder l = 2..0:  nearest x2 upsample, concat skip, conv(C_{l+1}+C_l -> C_l), conv(C_l -> C_l)
//   head:              1x1 conv C_0 -> 1
// every 3x3x3 conv is followed by ReLU; input has 8 channels (2 used).
#include <cmath>
#include <cstring>
#include <memory>
#include <random>

#include "mgx_ctx.h"
#include "unet.h"

namespace mgx {

namespace {

__global__ void head_fwd_kernel(const float* __restrict__ x, int c, long nvox,
                                const float* __restrict__ w, const float* __restrict__ b,
                                float* __restrict__ y) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float s = b[0];
    for (int ch = 0; ch < c; ++ch) s += w[ch] * x[v * c + ch];
    y[v] = s;
}

// dx[v][c] = dy[v] w[c]; per block: part[block][c] = sum_v dy[v] x[v][c],
// part[block][c_in] = sum_v dy[v] (the head's dW and db, summed in order
// afterwards)
__global__ void __launch_bounds__(256) head_bwd_kernel(const float* __restrict__ x, int c, long nvox,
                                                       const float* __restrict__ w,
                                                       const float* __restrict__ dy,
                                                       float* __restrict__ dx, float* __restrict__ part) {
    __shared__ float red[256];
    float acc[32];
    for (int ch = 0; ch < c; ++ch) acc[ch] = 0.f;
    float sb = 0.f;
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < nvox; v += (long)gridDim.x * 256) {
        const float g = dy[v];
        sb += g;
        for (int ch = 0; ch < c; ++ch) {
            acc[ch] += g * x[v * c + ch];
            dx[v * c + ch] = g * w[ch];
        }
    }
    for (int ch = 0; ch <= c; ++ch) {
        red[threadIdx.x] = ch < c ? acc[ch] : sb;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(long)blockIdx.x * (c + 1) + ch] = red[0];
        __syncthreads();
    }
}

inline unsigned g1(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void UNet::launch_conv(const ConvGeom& cg, const float* in, const unsigned short* w, const float* bias,
                       float* out, int relu) {
    conv3d_forward_launch(cg, in, w, bias, out, relu, ksws.as<float>(), kKsWsFloats, stream);
}


static int pad32(int k) { return (k + 31) / 32 * 32; }

void UNet::create(int base_channels, int levels_, uint64_t seed, hipStream_t s) {
    MGX_REQUIRE(levels_ == 4 && base_channels % 16 == 0 && base_channels >= 16 && base_channels <= 32,
                MGX_ERR_UNSUPPORTED, "U-net: 4 levels with base channels 16 or 32");
    levels = levels_;
    base = base_channels;
    stream = s;
    ksws.alloc(kKsWsFloats * sizeof(float));
    conv.clear();
    long off = 0, wb = 0, wt = 0;
    auto add = [&](int cin, int cout) {
        ConvLayer L;
        L.cin = cin;
        L.cout = cout;
        L.kp = pad32(27 * cin);
        L.kpt = pad32(27 * cout);
        L.woff = off;
        off += (long)cout * 27 * cin;
        L.boff = off;
        off += cout;
        L.wb_off = wb;
        wb += (long)cout * L.kp;
        L.wt_off = wt;
        wt += (long)cin * L.kpt;
        conv.push_back(L);
    };
    auto C = [&](int l) { return base << l; };
    for (int l = 0; l < levels - 1; ++l) {
        add(l == 0 ? kInC : C(l - 1), C(l));
        add(C(l), C(l));
    }
    add(C(levels - 2), C(levels - 1));
    add(C(levels - 1), C(levels - 1));
    for (int l = levels - 2; l >= 0; --l) {
        add(C(l + 1) + C(l), C(l));
        add(C(l), C(l));
    }
    head_w = off;
    off += base;
    head_b = off;
    off += 1;
    nparam = off;
    nwb = wb;
    nwt = wt;
    // He-normal initialisation (deterministic), zero biases
    std::vector<float> h(nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (const ConvLayer& L : conv) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * L.cin)));
        for (long q = 0; q < (long)L.cout * 27 * L.cin; ++q) h[L.woff + q] = nd(rng);
    }
    {
        std::normal_distribution<float> nd(0.f, std::sqrt(1.f / base));
        for (int q = 0; q < base; ++q) h[head_w + q] = nd(rng);
    }
    params.alloc(nparam * sizeof(float));
    grads.alloc(nparam * sizeof(float));
    adam_m.alloc(nparam * sizeof(float));
    adam_v.alloc(nparam * sizeof(float));
    MGX_HIP(hipMemcpy(params.p, h.data(), nparam * sizeof(float), hipMemcpyHostToDevice));
    MGX_HIP(hipMemset(adam_m.p, 0, nparam * sizeof(float)));
    MGX_HIP(hipMemset(adam_v.p, 0, nparam * sizeof(float)));
    MGX_HIP(hipMemset(grads.p, 0, nparam * sizeof(float)));
    wbf.alloc(nwb * sizeof(unsigned short));
    wtbf.alloc(nwt * sizeof(unsigned short));
    loss_buf.alloc(16 * sizeof(float));
    step = 0;
    bf_stale = true;
    shape = NetShape{};
}

void UNet::refresh_bf16() {
    if (!bf_stale) return;
    for (const ConvLayer& L : conv) {
        const long n = (long)L.cout * L.kp;
        weight_to_bf16_kernel<<<g1(n), 256, 0, stream>>>(params.as<float>() + L.woff, L.cout,
                                                         27 * L.cin, L.kp,
                                                         wbf.as<unsigned short>() + L.wb_off);
        const long nt = (long)L.cin * L.kpt;
        weight_transpose_kernel<<<g1(nt), 256, 0, stream>>>(wbf.as<unsigned short>() + L.wb_off,
                                                            L.cin, L.cout, L.kp,
                                                            wtbf.as<unsigned short>() + L.wt_off, L.kpt);
    }
    bf_stale = false;
}

int UNet::dims_at(int lvl, int d) const {
    const int n = d == 0 ? shape.nx : (d == 1 ? shape.ny : shape.nz);
    return n >> lvl;
}

// Activation slots: for encoder level l: 3l+0 = conv a, 3l+1 = conv b (skip),
// 3l+2 = pooled; then bottleneck a, b; then per decoder level (l = L-2..0):
// concat, conv a, conv b; last slot: head input gradient scratch.
void UNet::build_acts(const NetShape& sh) {
    if (shape == sh && !act.empty()) return;
    MGX_REQUIRE(sh.nx % 8 == 0 && sh.ny % 8 == 0 && sh.nz % 8 == 0 && sh.batch >= 1,
                MGX_ERR_ARG, "U-net input sizes must be multiples of 8");
    drop_graphs();  // they hold the old activation addresses
    shape = sh;
    act.clear();
    act_c.clear();
    act_lvl.clear();
    auto C = [&](int l) { return base << l; };
    auto slot = [&](int c, int lvl) { act_c.push_back(c); act_lvl.push_back(lvl); };
    for (int l = 0; l < levels - 1; ++l) {
        slot(C(l), l);
        slot(C(l), l);
        slot(C(l), l + 1);
    }
    slot(C(levels - 1), levels - 1);
    slot(C(levels - 1), levels - 1);
    for (int l = levels - 2; l >= 0; --l) {
        slot(C(l + 1) + C(l), l);
        slot(C(l), l);
        slot(C(l), l);
    }
    act.resize(act_c.size());
    long maxn = 0;
    for (size_t q = 0; q < act.size(); ++q) {
        const long nvox = (long)dims_at(act_lvl[q], 0) * dims_at(act_lvl[q], 1) *
                          dims_at(act_lvl[q], 2) * sh.batch;
        act[q].alloc(nvox * act_c[q] * sizeof(float));
        maxn = std::max(maxn, nvox * act_c[q]);
    }
    for (auto& d : dbuf) d.alloc(maxn * sizeof(float));
}

void UNet::conv_fwd(const ConvLayer& L, const float* in, int lvl, float* out, bool relu,
                    const unsigned short* w, const float* bias) {
    ConvGeom cg{};
    cg.nx = dims_at(lvl, 0);
    cg.ny = dims_at(lvl, 1);
    cg.nz = dims_at(lvl, 2);
    for (int d = 0; d < 3; ++d) cg.per[d] = shape.per[d];
    cg.batch = shape.batch;
    cg.cin = L.cin;
    cg.cout = L.cout;
    cg.kp = L.kp;
    launch_conv(cg, in, w, bias, out, relu ? 1 : 0);
}

static ConvGeom geom_at(const NetShape& sh, int lvl, int cin, int cout, int kp) {
    ConvGeom cg{};
    cg.nx = sh.nx >> lvl;
    cg.ny = sh.ny >> lvl;
    cg.nz = sh.nz >> lvl;
    for (int d = 0; d < 3; ++d) cg.per[d] = sh.per[d];
    cg.batch = sh.batch;
    cg.cin = cin;
    cg.cout = cout;
    cg.kp = kp;
    return cg;
}

static long nvox_at(const NetShape& sh, int lvl) {
    return (long)(sh.nx >> lvl) * (sh.ny >> lvl) * (sh.nz >> lvl) * sh.batch;
}

void UNet::zero_grads() {
    zero_f32_kernel<<<g1(nparam), 256, 0, stream>>>(grads.as<float>(), nparam);
}

UNet::~UNet() {
    drop_graphs();
    if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

void UNet::drop_graphs() {
    for (FwdGraph& g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

void UNet::forward(const float* in, const NetShape& sh, float* out) {
    build_acts(sh);
    refresh_bf16();
    forward_body(in, sh, out);
}

void UNet::forward_graph(const float* in, const NetShape& sh, float* out, hipStream_t s) {
    const hipStream_t saved = stream;
    stream = s;
    build_acts(sh);  // drops the graphs when the activations move
    refresh_bf16();
    stream = saved;
    const FwdGraph* hit = nullptr;
    for (const FwdGraph& g : graphs_)
        if (g.in == in && g.out == out && g.sh == sh) hit = &g;
    if (!hit && graphs_ok_) {
        if (graphs_.size() >= 8) drop_graphs();
        if (!cap_stream_ && hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            cap_stream_ = nullptr;
            graphs_ok_ = false;
        } else if (hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            graphs_ok_ = false;
        } else {
            bool body_ok = true;
            stream = cap_stream_;
            try {
                forward_body(in, sh, out);
            } catch (...) {
                body_ok = false;  // rethrown by the plain launches below
            }
            stream = saved;
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(cap_stream_, &g);
            hipGraphExec_t exec = nullptr;
            if (e == hipSuccess && body_ok) e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e == hipSuccess && body_ok) {
                graphs_.push_back(FwdGraph{in, out, sh, exec});
                hit = &graphs_.back();
            } else {
                (void)hipGetLastError();
                graphs_ok_ = false;
            }
        }
    }
    if (hit) {
        MGX_HIP(hipGraphLaunch(hit->exec, s));
        return;
    }
    stream = s;
    try {
        forward_body(in, sh, out);
    } catch (...) {
        stream = saved;
        throw;
    }
    stream = saved;
}

void UNet::forward_body(const float* in, const NetShape& sh, float* out) {
    const unsigned short* wb = wbf.as<unsigned short>();
    const float* P = params.as<float>();
    int li = 0, si = 0;
    const float* cur = in;
    auto A = [&](int q) { return act[q].as<float>(); };
    for (int l = 0; l < levels - 1; ++l) {
        const ConvLayer& La = conv[li++];
        const ConvLayer& Lb = conv[li++];
        conv_fwd(La, cur, l, A(si), true, wb + La.wb_off, P + La.boff);
        conv_fwd(Lb, A(si), l, A(si + 1), true, wb + Lb.wb_off, P + Lb.boff);
        const long n = nvox_at(sh, l + 1) * Lb.cout;
        avgpool_fwd_kernel<<<g1(n), 256, 0, stream>>>(A(si + 1), sh.nx >> l, sh.ny >> l, sh.nz >> l,
                                                      Lb.cout, sh.batch, A(si + 2));
        cur = A(si + 2);
        si += 3;
    }
    {
        const ConvLayer& La = conv[li++];
        const ConvLayer& Lb = conv[li++];
        conv_fwd(La, cur, levels - 1, A(si), true, wb + La.wb_off, P + La.boff);
        conv_fwd(Lb, A(si), levels - 1, A(si + 1), true, wb + Lb.wb_off, P + Lb.boff);
        cur = A(si + 1);
        si += 2;
    }
    int cur_c = base << (levels - 1);
    for (int l = levels - 2; l >= 0; --l) {
        const ConvLayer& La = conv[li++];
        const ConvLayer& Lb = conv[li++];
        const int cl = base << l, ctot = cur_c + cl;
        float* cat = A(si);
        upsample_concat_kernel<<<g1(nvox_at(sh, l) * ctot), 256, 0, stream>>>(
            cur, sh.nx >> (l + 1), sh.ny >> (l + 1), sh.nz >> (l + 1), cur_c, A(3 * l + 1), cl, sh.batch, cat);
        conv_fwd(La, cat, l, A(si + 1), true, wb + La.wb_off, P + La.boff);
        conv_fwd(Lb, A(si + 1), l, A(si + 2), true, wb + Lb.wb_off, P + Lb.boff);
        cur = A(si + 2);
        cur_c = cl;
        si += 3;
    }
    const long n0 = nvox_at(sh, 0);
    head_fwd_kernel<<<g1(n0), 256, 0, stream>>>(cur, base, n0, P + head_w, P + head_b, out);
}

// Size the gradient workspace for every reduction of one backward pass and
// the loss before any of them is launched (growing it later would free
// memory that launched kernels may still use).
void UNet::reserve_grad_ws(const NetShape& sh) {
    size_t need = 1;
    const int nl = (int)conv.size();
    for (int li = 0; li < nl; ++li) {
        // encoder level li / 2, bottleneck levels - 1, decoder from the coarsest
        const int lvl = li < 2 * (levels - 1) ? li / 2
                                              : (li < 2 * levels ? levels - 1 : levels - 2 - (li - 2 * levels) / 2);
        const ConvLayer& L = conv[li];
        const ConvGeom cg = geom_at(sh, lvl, L.cin, L.cout, L.kp);
        need = std::max(need, conv_dw_workspace(cg));
        need = std::max(need, (size_t)bias_grad_blocks(nvox_at(sh, lvl), L.cout) * L.cout);
    }
    const long n0 = nvox_at(sh, 0);
    need = std::max(need, (size_t)std::min<long>(1024, g1(n0)) * (base + 1));  // head
    need = std::max(need, (size_t)std::min<long>(1024, g1(n0)));             // loss
    gws.alloc(need * sizeof(float));
}

void UNet::backward(const float* in, const float* dout, const NetShape& sh) {
    MGX_REQUIRE(shape == sh && !act.empty(), MGX_ERR_STATE, "U-net backward without forward");
    reserve_grad_ws(sh);
    const unsigned short* wt = wtbf.as<unsigned short>();
    float* G = grads.as<float>();
    const float* P = params.as<float>();
    auto A = [&](int q) { return act[q].as<float>(); };
    const int nenc = 3 * (levels - 1);
    const int bott = nenc;                       // slots bott, bott+1
    auto dec_slot = [&](int l) { return nenc + 2 + 3 * (levels - 2 - l); };
    // gradient buffers: one per activation slot, same size
    std::vector<DevBuf>& gb = gact_;
    if (gb.size() != act.size()) {
        gb.clear();
        gb.resize(act.size());
    }
    for (size_t q = 0; q < act.size(); ++q) gb[q].alloc(act[q].bytes);
    auto GA = [&](int q) { return gb[q].as<float>(); };

    // layer-level backward: mask dy by relu, dW/db, optional dx
    auto lbwd = [&](const ConvLayer& L, const float* x, const float* y, float* dy, float* dx, int lvl) {
        const long nv = nvox_at(sh, lvl);
        relu_bwd_kernel<<<g1(nv * L.cout), 256, 0, stream>>>(y, dy, nv * L.cout);
        ConvGeom cg = geom_at(sh, lvl, L.cin, L.cout, L.kp);
        conv_dw_launch(cg, x, dy, G + L.woff, 27 * L.cin, gws.as<float>(), stream);
        bias_grad_launch(dy, nv, L.cout, G + L.boff, gws.as<float>(), stream);
        if (dx) {
            ConvGeom ct = geom_at(sh, lvl, L.cout, L.cin, L.kpt);
            launch_conv(ct, dy, wt + L.wt_off, nullptr, dx, 0);
        }
    };

    // head
    const long n0 = nvox_at(sh, 0);
    const int s0 = dec_slot(0);
    {
        // the head's dW (base values) and db follow each other in the parameter vector
        const unsigned hb = (unsigned)std::min<long>(1024, g1(n0));
        head_bwd_kernel<<<hb, 256, 0, stream>>>(A(s0 + 2), base, n0, P + head_w, dout, GA(s0 + 2),
                                                gws.as<float>());
        MGX_REQUIRE(head_b == head_w + base, MGX_ERR_STATE, "U-net: head parameter layout");
        sum_parts_kernel<<<1, 256, 0, stream>>>(gws.as<float>(), (int)hb, base + 1, G + head_w);
    }
    // decoder, finest first
    int li_dec = (int)conv.size() - 2;  // index of dec level 0 conv a
    for (int l = 0; l <= levels - 2; ++l) {
        const ConvLayer& La = conv[li_dec];
        const ConvLayer& Lb = conv[li_dec + 1];
        const int s = dec_slot(l);
        lbwd(Lb, A(s + 1), A(s + 2), GA(s + 2), GA(s + 1), l);
        lbwd(La, A(s), A(s + 1), GA(s + 1), GA(s), l);
        const int cl = base << l, cup = La.cin - cl;
        // skip part -> encoder skip gradient (overwrite; the pool path adds later)
        channel_copy_kernel<<<g1(nvox_at(sh, l) * cl), 256, 0, stream>>>(
            GA(s), La.cin, cup, GA(3 * l + 1), cl, 0, cl, nvox_at(sh, l), 0);
        // upsampled part -> gradient of the coarser decoder output (or bottleneck)
        float* dsrc = (l == levels - 2) ? GA(bott + 1) : GA(dec_slot(l + 1) + 2);
        const long nc = nvox_at(sh, l + 1) * cup;
        upsample_bwd_kernel<<<g1(nc), 256, 0, stream>>>(GA(s), sh.nx >> (l + 1), sh.ny >> (l + 1),
                                                        sh.nz >> (l + 1), cup, sh.batch, La.cin, 0,
                                                        dsrc);
        li_dec -= 2;
    }
    // bottleneck
    {
        const ConvLayer& La = conv[2 * (levels - 1)];
        const ConvLayer& Lb = conv[2 * (levels - 1) + 1];
        lbwd(Lb, A(bott), A(bott + 1), GA(bott + 1), GA(bott), levels - 1);
        lbwd(La, A(3 * (levels - 2) + 2), A(bott), GA(bott), GA(3 * (levels - 2) + 2), levels - 1);
    }
    // encoder, coarsest first
    for (int l = levels - 2; l >= 0; --l) {
        const ConvLayer& La = conv[2 * l];
        const ConvLayer& Lb = conv[2 * l + 1];
        const int s = 3 * l;
        // skip gradient += pool backward of the pooled output's gradient
        {
            const long n = nvox_at(sh, l) * Lb.cout;
            avgpool_bwd_kernel<<<g1(n), 256, 0, stream>>>(GA(s + 2), sh.nx >> l, sh.ny >> l,
                                                          sh.nz >> l, Lb.cout, sh.batch,
                                                          dbuf[0].as<float>());
            channel_copy_kernel<<<g1(n), 256, 0, stream>>>(dbuf[0].as<float>(), Lb.cout, 0,
                                                           GA(s + 1), Lb.cout, 0, Lb.cout,
                                                           nvox_at(sh, l), 1);
        }
        lbwd(Lb, A(s), A(s + 1), GA(s + 1), GA(s), l);
        const float* x = (l == 0) ? in : A(3 * (l - 1) + 2);
        lbwd(La, x, A(s), GA(s), l == 0 ? nullptr : GA(3 * (l - 1) + 2), l);
    }
}

float UNet::train_step(const float* in, const float* target, const NetShape& sh, float lr) {
    const long n0 = nvox_at(sh, 0);
    ybuf.alloc(n0 * 2 * sizeof(float));
    float* y = ybuf.as<float>();
    float* dy = y + n0;
    forward(in, sh, y);
    MGX_HIP(hipMemsetAsync(loss_buf.p, 0, sizeof(float), stream));
    const unsigned lb = (unsigned)std::min<long>(1024, g1(n0));
    reserve_grad_ws(sh);
    mse_kernel<<<lb, 256, 0, stream>>>(y, 1, target, n0, dy, gws.as<float>());
    sum_parts_kernel<<<1, 256, 0, stream>>>(gws.as<float>(), (int)lb, 1, loss_buf.as<float>());
    zero_grads();
    backward(in, dy, sh);
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    adam_kernel<<<g1(nparam), 256, 0, stream>>>(params.as<float>(), grads.as<float>(),
                                                 adam_m.as<float>(), adam_v.as<float>(), nparam, lr,
                                                 b1, b2, 1e-8f, c1, c2);
    bf_stale = true;
    float h = 0.f;
    MGX_HIP(hipMemcpyAsync(&h, loss_buf.p, sizeof(float), hipMemcpyDeviceToHost, stream));
    MGX_HIP(hipStreamSynchronize(stream));
    return h;
}

}  // namespace mgx

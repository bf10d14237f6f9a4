// unet.h -- the learned coarse-level predictor (3-D conv U-net).
#pragma once

#include <cstdint>
#include <vector>

#include "mgx_common.h"
#include "unet_kernels.h"

namespace mgx {

// One 3x3x3 convolution layer (+ReLU); parameters live in the network's flat
// FP32 master buffer at [woff, woff + cout*27*cin) and [boff, boff + cout).
struct ConvLayer {
    int cin, cout, kp;
    long woff, boff;
    long wb_off;   // offset of the BF16 copy [cout][kp]
    long wt_off;   // offset of the BF16 data-gradient weights [cin][kpt]
    int kpt;       // padded K of the transposed weights (27*cout -> multiple of 32)
};

// Geometry of one forward/backward call.
struct NetShape {
    int nx, ny, nz, batch;
    int per[3];
    bool operator==(const NetShape& o) const {
        return nx == o.nx && ny == o.ny && nz == o.nz && batch == o.batch && per[0] == o.per[0] &&
               per[1] == o.per[1] && per[2] == o.per[2];
    }
};

class UNet {
   public:
    static constexpr int kInC = 8;  // input channels (2 used, zero padded to 8)
    int levels = 4, base = 16;
    std::vector<ConvLayer> conv;   // enc (2/level, levels-1), bottleneck (2), dec (2/level)
    long head_w = 0, head_b = 0;   // 1x1 head 'base' -> 1
    long nparam = 0;
    DevBuf params, grads, adam_m, adam_v, wbf, wtbf;
    long nwb = 0, nwt = 0;
    long step = 0;
    bool bf_stale = true;
    hipStream_t stream = nullptr;

    // activations for the current shape
    NetShape shape{};
    std::vector<DevBuf> act;   // see build_acts() for the slot meaning
    std::vector<int> act_c, act_lvl;
    DevBuf dbuf[4];            // gradient scratch
    std::vector<DevBuf> gact_; // gradients of the activation slots
    DevBuf loss_buf, ybuf;
    DevBuf ksws;                    // split-K partial sums of the forward convolutions
    DevBuf gws;                     // per-chunk / per-block gradient partials (summed in a fixed order)
    static constexpr long kKsWsFloats = 4L << 20;

    UNet() = default;
    UNet(const UNet&) = delete;
    UNet& operator=(const UNet&) = delete;
    ~UNet();
    void create(int base_channels, int levels_, uint64_t seed, hipStream_t s);
    void forward(const float* in, const NetShape& sh, float* out);
    // the same forward on stream s, replayed from a HIP graph captured on the
    // first call for these (in, out, shape); plain launches where capture is
    // not available.  Used inside the solver, where the forward's ~30 small
    // launches would otherwise each pay a host launch gap.
    void forward_graph(const float* in, const NetShape& sh, float* out, hipStream_t s);
    // backward from dL/d(out) (one channel, [batch][S]); accumulates into grads
    void backward(const float* in, const float* dout, const NetShape& sh);
    float train_step(const float* in, const float* target, const NetShape& sh, float lr);
    void zero_grads();
    void refresh_bf16();

   private:
    struct FwdGraph {
        const float* in;
        float* out;
        NetShape sh;
        hipGraphExec_t exec;
    };
    std::vector<FwdGraph> graphs_;
    hipStream_t cap_stream_ = nullptr;  // capture stream (the caller's may be the null stream)
    bool graphs_ok_ = true;
    void drop_graphs();
    void reserve_grad_ws(const NetShape& sh);
    void forward_body(const float* in, const NetShape& sh, float* out);
    void build_acts(const NetShape& sh);
    int dims_at(int lvl, int d) const;
    void conv_fwd(const ConvLayer& L, const float* in, int lvl, float* out, bool relu,
                  const unsigned short* w, const float* bias);
    void launch_conv(const ConvGeom& cg, const float* in, const unsigned short* w, const float* bias,
                     float* out, int relu);
};

}  // namespace mgx

struct mgx_ctx_s;
struct mgx_unet_s {
    mgx_ctx_s* ctx = nullptr;
    mgx::UNet net;
};

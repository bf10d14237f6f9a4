// unet_kernels.h -- kernels of the learned coarse-level predictor: a 3-D
// convolutional U-net with BF16 weights and FP32 accumulation on the CDNA4
// matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Tensors are channels-last FP32: t[(b*S + v)*C + c], v = x + nx*(y + ny*z).
// A 3x3x3 convolution is an implicit GEMM
//     out[v][co] = bias[co] + sum_k A[v][k] W[co][k],  k = tap*Cin + ci,
//     A[v][tap*Cin+ci] = in[v + off(tap)][ci]   (periodic wrap or zero pad),
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1).  Weights are stored BF16 as W[co][Kp]
// with K = 27*Cin padded to Kp (a multiple of 32) by zeros, so one MFMA step
// covers 32 consecutive k.  Cin is a multiple of 8: the 8 k values a lane
// feeds per step are 8 consecutive input channels of one tap (one 32-byte
// activation load converted to BF16 in registers).
//
// MFMA 16x16x32 operand layout (lane l of a 64-wide wave, i = 0..7):
//   A[row = l%16][k = 8*(l/16) + i],  B[k = 8*(l/16) + i][col = l%16],
//   D[row = 4*(l/16) + j][col = l%16], j = 0..3.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgx_common.h"

namespace mgx {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

struct ConvGeom {
    int nx, ny, nz;     // spatial size of one sample
    int per[3];         // 1: periodic wrap, 0: zero padding
    int batch;
    int cin, cout;      // cin % 8 == 0, cout % 16 == 0
    int kp;             // padded K (multiple of 32)
};

__device__ __forceinline__ int conv_wrap(int i, int n, int per, bool& valid) {
    if (i >= 0 && i < n) return i;
    if (per) return i < 0 ? i + n : i - n;
    valid = false;
    return 0;
}

__device__ __forceinline__ unsigned short bf16_bits(float f) {
    // round to nearest even (finite inputs)
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ bf16x8 pack_bf16x8(const float4& a, const float4& b) {
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = __builtin_bit_cast(__bf16, bf16_bits(v[i]));
    return r;
}

// Forward 3x3x3 convolution (+bias, optional fused ReLU).  One wave computes
// 16 voxels x all output channels (cout/16 accumulator tiles); a workgroup of
// 4 waves covers 64 consecutive voxels of the batch.  The K loop issues the
// activation loads of U consecutive K steps before their MFMAs, so U loads are
// in flight per lane (the convolutions here are latency-bound: few voxels, few
// waves per CU).
// Split-K (part != nullptr): blockIdx.y selects a contiguous range of the K
// steps and the raw partial sums go to part[split][v][co]; bias and ReLU are
// applied by conv3d_splitk_reduce_kernel, which adds the splits in a fixed
// order.  Small levels (few voxels, long K) then keep many waves in flight
// instead of a few long dependent chains.  (Finishing the sum in the last
// block of each tile instead saves that launch but costs more here: the
// partials must be written back from one XCD's L2 and read from another's,
// and one block then sums a whole tile alone.)
template <int NT, int U>
static __global__ void __launch_bounds__(256) conv3d_fwd_kernel(ConvGeom g, const float* __restrict__ in,
                                                               const unsigned short* __restrict__ w,
                                                               const float* __restrict__ bias,
                                                               float* __restrict__ out, int relu,
                                                               float* __restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long S = (long)g.nx * g.ny * g.nz;
    const long nvox = S * g.batch;
    const long vblk = (long)blockIdx.x * 64;
    const long v0 = vblk + wave * 16;
    const bool active = v0 < nvox;  // whole wave
    const int row = lane & 15, kq = lane >> 4;
    // this lane's A row: voxel v0 + row
    long va = v0 + row;
    const bool rowok = va < nvox;
    if (!rowok) va = nvox - 1;
    const int b = (int)(va / S);
    const int vs = (int)(va - (long)b * S);
    const int x = vs % g.nx, y = (vs / g.nx) % g.ny, z = vs / (g.nx * g.ny);
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = 27 * g.cin;
    const int ksplit = (int)gridDim.y;
    const int steps = g.kp / 32, per = (steps + ksplit - 1) / ksplit;
    const int s0 = (int)blockIdx.y * per, s1 = min(steps, s0 + per);
    const float4 zero4 = {0.f, 0.f, 0.f, 0.f};
    if (active) {
        for (int st = s0; st < s1; st += U) {
            float4 ra[U][2];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = 32 * (st + u) + 8 * kq;
                ra[u][0] = zero4;
                ra[u][1] = zero4;
                if (st + u < s1 && k < K) {
                    const int tap = k / g.cin, ci = k - tap * g.cin;
                    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                    bool valid = rowok;
                    const int xx = conv_wrap(x + dx, g.nx, g.per[0], valid);
                    const int yy = conv_wrap(y + dy, g.ny, g.per[1], valid);
                    const int zz = conv_wrap(z + dz, g.nz, g.per[2], valid);
                    if (valid) {
                        const float* p = in + ((long)b * S + xx + (long)g.nx * (yy + (long)g.ny * zz)) * g.cin + ci;
                        ra[u][0] = *reinterpret_cast<const float4*>(p);
                        ra[u][1] = *reinterpret_cast<const float4*>(p + 4);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (st + u < s1) {  // uniform across the wave
                    const int k = 32 * (st + u) + 8 * kq;
                    const bf16x8 afrag = pack_bf16x8(ra[u][0], ra[u][1]);
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const int co = t * 16 + row;
                        const bf16x8 bfrag = *reinterpret_cast<const bf16x8*>(w + (long)co * g.kp + k);
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag, acc[t], 0, 0, 0);
                    }
                }
            }
        }
    }
    if (!active) return;
    if (part) {
        float* dst = part + (long)blockIdx.y * nvox * g.cout;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + row;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long v = v0 + 4 * kq + j;
                if (v < nvox) dst[v * g.cout + co] = acc[t][j];
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + row;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long v = v0 + 4 * kq + j;
            if (v < nvox) {
                float r = acc[t][j] + bv;
                if (relu) r = fmaxf(r, 0.f);
                out[v * g.cout + co] = r;
            }
        }
    }
}

// LDS-tiled forward convolution for sizes that are multiples of 4: a
// workgroup computes a 4x4x4 voxel tile (wave w: z-plane w of the tile, 16
// voxels) x all output channels.  It first stages the tile's 6x6x6 input halo,
// every input channel, in LDS as BF16 (one pass of independent float4 loads,
// so the activations are read from memory about 3.4x instead of the 27x of
// the implicit GEMM's per-tap loads), then reads the A fragments from LDS
// (one ds_read_b128 per lane and K step).  The weight fragments of U K steps
// are loaded before their MFMAs.  Split-K as in conv3d_fwd_kernel.
constexpr int kConvHaloPts = 216;  // 6 x 6 x 6
constexpr int kConvLdsMax = 64 * 1024;

// BF16 elements per halo point in LDS: the channels plus 8 (a 16-byte pad)
inline int conv_lds_stride(int cin) { return cin + 8; }
inline size_t conv_lds_bytes(int cin) { return (size_t)kConvHaloPts * conv_lds_stride(cin) * sizeof(unsigned short); }
inline bool conv_lds_ok(const ConvGeom& g) {
    return g.nx % 4 == 0 && g.ny % 4 == 0 && g.nz % 4 == 0 && g.cin % 8 == 0 &&
           conv_lds_bytes(g.cin) <= (size_t)kConvLdsMax;
}

template <int NT, int U>
static __global__ void __launch_bounds__(256) conv3d_fwd_lds_kernel(ConvGeom g, const float* __restrict__ in,
                                                                   const unsigned short* __restrict__ w,
                                                                   const float* __restrict__ bias,
                                                                   float* __restrict__ out, int relu,
                                                                   float* __restrict__ part) {
    extern __shared__ __align__(16) unsigned short hal[];
    const int cs = g.cin + 8;
    const int tnx = g.nx >> 2, tny = g.ny >> 2, tnz = g.nz >> 2;
    int tile = (int)blockIdx.x;
    const int tx = tile % tnx;
    tile /= tnx;
    const int ty = tile % tny;
    tile /= tny;
    const int tz = tile % tnz;
    const int b = tile / tnz;
    const long S = (long)g.nx * g.ny * g.nz;
    const long nvox = S * g.batch;
    // stage the halo: point pos = hx + 6 (hy + 6 hz) holds input voxel
    // (4 tx - 1 + hx, ...) (wrapped, or zero outside a non-periodic face)
    const int c4n = g.cin >> 2;
    const int nld = kConvHaloPts * c4n;
    for (int i = threadIdx.x; i < nld; i += 256) {
        const int pos = i / c4n, c4 = i - pos * c4n;
        const int hx = pos % 6, hy = (pos / 6) % 6, hz = pos / 36;
        bool valid = true;
        const int gx = conv_wrap(4 * tx - 1 + hx, g.nx, g.per[0], valid);
        const int gy = conv_wrap(4 * ty - 1 + hy, g.ny, g.per[1], valid);
        const int gz = conv_wrap(4 * tz - 1 + hz, g.nz, g.per[2], valid);
        float4 v = {0.f, 0.f, 0.f, 0.f};
        if (valid)
            v = *reinterpret_cast<const float4*>(in + ((long)b * S + gx + (long)g.nx * (gy + (long)g.ny * gz)) * g.cin +
                                                 4 * c4);
        ushort4 h;
        h.x = bf16_bits(v.x);
        h.y = bf16_bits(v.y);
        h.z = bf16_bits(v.z);
        h.w = bf16_bits(v.w);
        *reinterpret_cast<ushort4*>(hal + pos * cs + 4 * c4) = h;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    // this lane's A row: tile voxel (row & 3, row >> 2, wave)
    const int x = row & 3, y = row >> 2, z = wave;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = 27 * g.cin;
    const int ksplit = (int)gridDim.y;
    const int steps = g.kp / 32, per = (steps + ksplit - 1) / ksplit;
    const int s0 = (int)blockIdx.y * per, s1 = min(steps, s0 + per);
    for (int st = s0; st < s1; st += U) {
        bf16x8 bq[U][NT];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (st + u < s1) {
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    bq[u][t] = *reinterpret_cast<const bf16x8*>(w + (long)(t * 16 + row) * g.kp + 32 * (st + u) +
                                                                8 * kq);
            }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (st + u < s1) {  // uniform across the wave
                const int k = 32 * (st + u) + 8 * kq;
                bf16x8 afrag;
                if (k < K) {
                    const int tap = k / g.cin, ci = k - tap * g.cin;
                    const int hp = ((z + tap / 9) * 6 + (y + (tap / 3) % 3)) * 6 + (x + tap % 3);
                    afrag = *reinterpret_cast<const bf16x8*>(hal + hp * cs + ci);
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i) afrag[i] = __builtin_bit_cast(__bf16, (unsigned short)0);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bq[u][t], acc[t], 0, 0, 0);
            }
        }
    }
    // D[row = tile voxel 4 kq + j][col = output channel t*16 + row]
    float* dst = part ? part + (long)blockIdx.y * nvox * g.cout : out;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int m = 4 * kq + j;
        const long v = (long)b * S + (4 * tx + (m & 3)) + (long)g.nx * ((4 * ty + (m >> 2)) + (long)g.ny * (4 * tz + z));
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + row;
            float r = acc[t][j];
            if (!part) {
                if (bias) r += bias[co];
                if (relu) r = fmaxf(r, 0.f);
            }
            dst[v * g.cout + co] = r;
        }
    }
}

// out[v][co] = sum_s part[s][v][co] + bias[co] (fixed order), optional ReLU
static __global__ void __launch_bounds__(256) conv3d_splitk_reduce_kernel(
    long n, int cout, int ksplit, const float* __restrict__ part, const float* __restrict__ bias,
    float* __restrict__ out, int relu) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float r = 0.f;
    for (int q = 0; q < ksplit; ++q) r += part[(long)q * n + e];
    if (bias) r += bias[e % cout];
    if (relu) r = fmaxf(r, 0.f);
    out[e] = r;
}

// Forward conv launch.  ws (ws_floats floats, may be null) enables split-K
// when the voxel count alone gives too few waves to fill the GPU.
inline void conv3d_forward_launch(const ConvGeom& cg, const float* in, const unsigned short* w,
                                  const float* bias, float* out, int relu, float* ws,
                                  long ws_floats, hipStream_t s) {
    const long nvox = (long)cg.nx * cg.ny * cg.nz * cg.batch;
    const bool lds = conv_lds_ok(cg);
    // LDS form: one 64-voxel tile per workgroup; otherwise 64 consecutive voxels
    const unsigned nb = (unsigned)((nvox + 63) / 64);
    const int steps = cg.kp / 32;
    int ksplit = 1;
    if (ws) {
        // aim at about one workgroup per CU: every split of the LDS form
        // stages the whole halo again, and every split is one more partial
        ksplit = (int)std::min<long>(std::max<long>(1, 256 / (long)nb), std::max(1, steps / 4));
        while (ksplit > 1 && (long)ksplit * nvox * cg.cout > ws_floats) --ksplit;
        if (ksplit > 1) {
            const int per = (steps + ksplit - 1) / ksplit;
            ksplit = (steps + per - 1) / per;  // no empty splits
        }
    }
    float* part = ksplit > 1 ? ws : nullptr;
    const dim3 grid(nb, ksplit);
    const size_t shm = lds ? conv_lds_bytes(cg.cin) : 0;
#define MGX_CONV_FWD(NT, U)                                                                                  \
    if (lds)                                                                                                 \
        conv3d_fwd_lds_kernel<NT, U><<<grid, 256, shm, s>>>(cg, in, w, bias, out, relu, part);               \
    else                                                                                                     \
        conv3d_fwd_kernel<NT, U><<<grid, 256, 0, s>>>(cg, in, w, bias, out, relu, part)
    switch (cg.cout / 16) {
        case 1: MGX_CONV_FWD(1, 8); break;
        case 2: MGX_CONV_FWD(2, 4); break;
        case 3: MGX_CONV_FWD(3, 2); break;
        case 4: MGX_CONV_FWD(4, 2); break;
        case 6: MGX_CONV_FWD(6, 1); break;
        case 8: MGX_CONV_FWD(8, 1); break;
        case 12: MGX_CONV_FWD(12, 1); break;
        default: throw Error(MGX_ERR_UNSUPPORTED, "conv3d: output channels must be 16,32,48,64,96,128 or 192");
    }
#undef MGX_CONV_FWD
    if (part) {
        const long n = nvox * cg.cout;
        conv3d_splitk_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, cg.cout, ksplit, part,
                                                                               bias, out, relu);
    }
}

// Weight gradient: part[chunk][co][k] = sum over the chunk's voxels v of
// dY[v][co] * A[v][k] (row stride kp); conv3d_dw_reduce_kernel adds the
// chunks in a fixed order, so training is reproducible.  One wave owns one
// 16 (co) x 16 (k) tile and one chunk of voxels.
static __global__ void __launch_bounds__(64) conv3d_dw_kernel(ConvGeom g, const float* __restrict__ in,
                                                              const float* __restrict__ dy,
                                                              float* __restrict__ part, int vchunk) {
    const int lane = threadIdx.x;
    const int row = lane & 15, kq = lane >> 4;
    const int ntk = g.kp / 16;
    const int tile = blockIdx.x % (ntk * (g.cout / 16));
    const int chunk = blockIdx.x / (ntk * (g.cout / 16));
    const int co0 = (tile / ntk) * 16, kk0 = (tile % ntk) * 16;
    const long S = (long)g.nx * g.ny * g.nz;
    const long nvox = S * g.batch;
    const long vbeg = (long)chunk * vchunk;
    const long vend = vbeg + vchunk < nvox ? vbeg + vchunk : nvox;
    const int K = 27 * g.cin;
    // this lane's B column: k = kk0 + row
    const int kcol = kk0 + row;
    int tap = 0, ci = 0, dz = 0, dyo = 0, dx = 0;
    const bool kok = kcol < K;
    if (kok) {
        tap = kcol / g.cin;
        ci = kcol - tap * g.cin;
        dz = tap / 9 - 1;
        dyo = (tap / 3) % 3 - 1;
        dx = tap % 3 - 1;
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long vb = vbeg; vb < vend; vb += 32) {
        bf16x8 a, bb;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const long v = vb + 8 * kq + i;
            float av = 0.f, bv = 0.f;
            if (v < vend) {
                av = dy[v * g.cout + co0 + row];
                if (kok) {
                    const int b = (int)(v / S);
                    const int vs = (int)(v - (long)b * S);
                    const int x = vs % g.nx, y = (vs / g.nx) % g.ny, z = vs / (g.nx * g.ny);
                    bool valid = true;
                    const int xx = conv_wrap(x + dx, g.nx, g.per[0], valid);
                    const int yy = conv_wrap(y + dyo, g.ny, g.per[1], valid);
                    const int zz = conv_wrap(z + dz, g.nz, g.per[2], valid);
                    if (valid) bv = in[((long)b * S + xx + (long)g.nx * (yy + (long)g.ny * zz)) * g.cin + ci];
                }
            }
            a[i] = __builtin_bit_cast(__bf16, bf16_bits(av));
            bb[i] = __builtin_bit_cast(__bf16, bf16_bits(bv));
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
    }
    // D[row = co offset 4*kq + j][col = k offset row]
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int co = co0 + 4 * kq + j;
        const int kc = kk0 + row;
        if (kc < K) part[((long)chunk * g.cout + co) * g.kp + kc] = acc[j];
    }
}

// dW[co][k] += sum over chunks c (in order) of part[c][co][k], k < K (row stride ldw)
static __global__ void __launch_bounds__(256) conv3d_dw_reduce_kernel(const float* __restrict__ part,
                                                                     int nchunk, int cout, int kp, int K,
                                                                     float* __restrict__ dw, int ldw) {
    const long id = (long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long)cout * K) return;
    const int co = (int)(id / K), k = (int)(id - (long)co * K);
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[((long)c * cout + co) * kp + k];
    dw[(long)co * ldw + k] += s;
}

// out[i] += sum over parts p (in order) of part[p n + i]
static __global__ void __launch_bounds__(256) sum_parts_kernel(const float* __restrict__ part, int nparts,
                                                              int n, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int p = 0; p < nparts; ++p) s += part[(long)p * n + i];
    out[i] += s;
}

// voxels per weight-gradient wave: enough waves to fill the GPU, at least 512
// voxels each, a multiple of the 32 voxels of one MFMA step
inline long conv_dw_chunk(const ConvGeom& cg) {
    const long nvox = (long)cg.nx * cg.ny * cg.nz * cg.batch;
    const int tiles = (cg.kp / 16) * (cg.cout / 16);
    const long chunk = std::max<long>(512, (nvox * tiles) / (256L * 64));
    return (chunk + 31) / 32 * 32;
}

// floats of workspace conv_dw_launch needs
inline size_t conv_dw_workspace(const ConvGeom& cg) {
    const long nvox = (long)cg.nx * cg.ny * cg.nz * cg.batch;
    const long chunk = conv_dw_chunk(cg);
    return (size_t)((nvox + chunk - 1) / chunk) * cg.cout * cg.kp;
}

// dW[co][k] += dL/dW of the convolution (row stride ldw), through the per-chunk
// partials in ws (conv_dw_workspace floats)
inline void conv_dw_launch(const ConvGeom& cg, const float* in, const float* dy, float* dw, int ldw,
                           float* ws, hipStream_t s) {
    const long nvox = (long)cg.nx * cg.ny * cg.nz * cg.batch;
    const int tiles = (cg.kp / 16) * (cg.cout / 16);
    const long chunk = conv_dw_chunk(cg);
    const long nchunk = (nvox + chunk - 1) / chunk;
    conv3d_dw_kernel<<<(unsigned)(tiles * nchunk), 64, 0, s>>>(cg, in, dy, ws, (int)chunk);
    const long n = (long)cg.cout * 27 * cg.cin;
    conv3d_dw_reduce_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(ws, (int)nchunk, cg.cout, cg.kp,
                                                                       27 * cg.cin, dw, ldw);
}

// bias gradient, per block: part[block][c] = sum over the block's voxels of
// dY[v][c], summed in a fixed order (sum_parts_kernel adds the blocks)
static __global__ void __launch_bounds__(256) bias_grad_kernel(const float* __restrict__ dy, long nvox,
                                                               int c, float* __restrict__ part) {
    __shared__ float red[256];
    const int co = threadIdx.x % c;
    const int per = 256 / c, lanes = per * c;
    float s = 0.f;
    if ((int)threadIdx.x < lanes)
        for (long v = (long)blockIdx.x * per + threadIdx.x / c; v < nvox; v += (long)gridDim.x * per)
            s += dy[v * c + co];
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < c) {
        float t = 0.f;
        for (int r = 0; r < per; ++r) t += red[r * c + threadIdx.x];
        part[(long)blockIdx.x * c + threadIdx.x] = t;
    }
}

inline int bias_grad_blocks(long nvox, int c) { return (int)std::min<long>(1024, (nvox * c + 255) / 256); }

// db[c] += sum_v dY[v][c] through the per-block partials in ws
// (bias_grad_blocks * c floats)
inline void bias_grad_launch(const float* dy, long nvox, int c, float* db, float* ws, hipStream_t s) {
    const int nb = bias_grad_blocks(nvox, c);
    bias_grad_kernel<<<nb, 256, 0, s>>>(dy, nvox, c, ws);
    sum_parts_kernel<<<(unsigned)((c + 255) / 256), 256, 0, s>>>(ws, nb, c, db);
}

// Data-gradient weights: Wt[ci][tap'*Cout + co] = W[co][(26 - tap')*Cin + ci]
// (flipped taps, transposed channels; dX is then a forward conv of dY).
static __global__ void weight_transpose_kernel(const unsigned short* __restrict__ w, int cin, int cout,
                                               int kp, unsigned short* __restrict__ wt, int kpt) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= cin * kpt) return;
    const int ci = id / kpt, k = id % kpt;
    unsigned short v = 0;
    if (k < 27 * cout) {
        const int tp = k / cout, co = k % cout;
        v = w[(long)co * kp + (26 - tp) * cin + ci];
    }
    wt[id] = v;
}

// FP32 master weights [co][K] (K = 27*cin) -> BF16 [co][Kp] (zero padded)
static __global__ void weight_to_bf16_kernel(const float* __restrict__ wm, int cout, int K, int kp,
                                             unsigned short* __restrict__ wb) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= cout * kp) return;
    const int co = id / kp, k = id % kp;
    wb[id] = k < K ? bf16_bits(wm[(long)co * K + k]) : (unsigned short)0;
}

// dY *= (y > 0)
static __global__ void relu_bwd_kernel(const float* __restrict__ y, float* __restrict__ dy, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n && !(y[id] > 0.f)) dy[id] = 0.f;
}

// 2x2x2 average pooling (channels last)
static __global__ void avgpool_fwd_kernel(const float* __restrict__ in, int nx, int ny, int nz, int c,
                                          int batch, float* __restrict__ out) {
    const int ox = nx / 2, oy = ny / 2, oz = nz / 2;
    const long n = (long)ox * oy * oz * batch * c;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ch = id % c;
    long v = id / c;
    const int X = v % ox; v /= ox;
    const int Y = v % oy; v /= oy;
    const int Z = v % oz;
    const int b = (int)(v / oz);
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const long fv = (long)b * nx * ny * nz + (2 * X + (q & 1)) +
                        (long)nx * ((2 * Y + ((q >> 1) & 1)) + (long)ny * (2 * Z + ((q >> 2) & 1)));
        s += in[fv * c + ch];
    }
    out[id] = 0.125f * s;
}

static __global__ void avgpool_bwd_kernel(const float* __restrict__ dout, int nx, int ny, int nz,
                                          int c, int batch, float* __restrict__ din) {
    const long n = (long)nx * ny * nz * batch * c;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ch = id % c;
    long v = id / c;
    const int x = v % nx; v /= nx;
    const int y = v % ny; v /= ny;
    const int z = v % nz;
    const int b = (int)(v / nz);
    const int ox = nx / 2, oy = ny / 2, oz = nz / 2;
    const long cv = (long)b * ox * oy * oz + (x / 2) + (long)ox * ((y / 2) + (long)oy * (z / 2));
    din[id] = 0.125f * dout[cv * c + ch];
}

// backward of the nearest x2 upsampling into channels [coff, coff+c) of a
// ctot-channel tensor: the sum of the 8 children
static __global__ void upsample_bwd_kernel(const float* __restrict__ dout, int nx, int ny, int nz,
                                           int c, int batch, int ctot, int coff,
                                           float* __restrict__ din) {
    const long n = (long)nx * ny * nz * batch * c;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ch = id % c;
    long v = id / c;
    const int X = v % nx; v /= nx;
    const int Y = v % ny; v /= ny;
    const int Z = v % nz;
    const int b = (int)(v / nz);
    const int fx = 2 * nx, fy = 2 * ny, fz = 2 * nz;
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const long fv = (long)b * fx * fy * fz + (2 * X + (q & 1)) +
                        (long)fx * ((2 * Y + ((q >> 1) & 1)) + (long)fy * (2 * Z + ((q >> 2) & 1)));
        s += dout[fv * ctot + coff + ch];
    }
    din[id] = s;
}

// decoder input: channels [0, c) of out (ctot = c + cs channels) are the
// nearest x2 upsampling of the coarse tensor in (c channels), channels
// [c, c + cs) a copy of the skip tensor (cs channels, fine grid)
static __global__ void upsample_concat_kernel(const float* __restrict__ in, int nx, int ny, int nz, int c,
                                              const float* __restrict__ skip, int cs, int batch,
                                              float* __restrict__ out) {
    const int fx = 2 * nx, fy = 2 * ny, fz = 2 * nz;
    const int ctot = c + cs;
    const long n = (long)fx * fy * fz * batch * ctot;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ch = id % ctot;
    const long fv = id / ctot;
    if (ch >= c) {
        out[id] = skip[fv * cs + (ch - c)];
        return;
    }
    long v = fv;
    const int x = v % fx; v /= fx;
    const int y = v % fy; v /= fy;
    const int z = v % fz;
    const int b = (int)(v / fz);
    const long cv = (long)b * nx * ny * nz + (x / 2) + (long)nx * ((y / 2) + (long)ny * (z / 2));
    out[id] = in[cv * c + ch];
}

// copy c channels between tensors with different channel counts / offsets
static __global__ void channel_copy_kernel(const float* __restrict__ src, int csrc, int soff,
                                           float* __restrict__ dst, int cdst, int doff, int c, long nvox,
                                           int accumulate) {
    const long n = nvox * c;
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int ch = id % c;
    const long v = id / c;
    const float val = src[v * csrc + soff + ch];
    float& d = dst[v * cdst + doff + ch];
    d = accumulate ? d + val : val;
}

// MSE loss 1/N sum (y - t)^2 over channel 0 of y (cy channels); gradient into dy
static __global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ y, int cy,
                                                         const float* __restrict__ t, long nvox,
                                                         float* __restrict__ dy, float* __restrict__ loss) {
    __shared__ float red[256];
    float s = 0.f;
    const float inv = 1.f / (float)nvox;
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < nvox; v += (long)gridDim.x * 256) {
        const float e = y[v * cy] - t[v];
        s += e * e;
        dy[v * cy] = 2.f * e * inv;
        for (int c = 1; c < cy; ++c) dy[v * cy + c] = 0.f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[blockIdx.x] = red[0] * inv;  // per block; summed in order afterwards
}

// Adam (Kingma & Ba 2015) on FP32 master parameters
static __global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                   float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                   float b1, float b2, float eps, float c1, float c2) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const float gi = g[id];
    const float mi = b1 * m[id] + (1.f - b1) * gi;
    const float vi = b2 * v[id] + (1.f - b2) * gi * gi;
    m[id] = mi;
    v[id] = vi;
    p[id] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static __global__ void zero_f32_kernel(float* a, long n) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) a[id] = 0.f;
}

}  // namespace mgx

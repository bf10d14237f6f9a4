/* orc_unet.c -- FP64 CPU restatement of the learned coarse-level predictor:
 * the 3x3x3 convolution (forward, data/weight/bias gradients) and the 4-level
 * U-net built from it.  TEST INFRASTRUCTURE: the checker of the HIP
 * convolution and U-net kernels; never linked into the product library.
 *
 * The network has no counterpart in the reference (incflo has no learned
 * component); its definition is the one the product documents:
 *   tensors channels last, x fastest: a[((b*nz + z)*ny + y)*nx + x][c]
 *   conv:  out[v][co] = bias[co] + sum_{tap,ci} w[co][tap*cin+ci] in[v+off(tap)][ci]
 *          off(tap) = (tap%3-1, (tap/3)%3-1, tap/9-1); periodic wrap or zero padding
 *   U-net (levels L, C_l = base*2^l):
 *     encoder l < L-1: a = relu(conv(x)), s_l = relu(conv(a)), x = avgpool2(s_l)
 *     bottleneck:      a = relu(conv(x)), y = relu(conv(a))
 *     decoder l = L-2..0: c = [upsample2(y), s_l], a = relu(conv(c)), y = relu(conv(a))
 *     head: out = sum_c hw[c] y[c] + hb
 *   flat parameters: per conv (creation order) w[cout][27*cin] then b[cout],
 *   then head w[base], head b.
 * bf16 != 0 rounds the products' operands to bfloat16 (round to nearest even)
 * at the points where the matrix-core kernels do: conv inputs and weights in
 * the forward and data-gradient passes, inputs and upstream gradients in the
 * weight-gradient pass.  Sums stay in FP64. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static double rb(double x) {
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static inline double R(double x, int bf16) { return bf16 ? rb(x) : x; }

/* neighbour voxel index of v shifted by (dx,dy,dz); -1 when outside (zero pad) */
static inline long nbr(int x, int y, int z, int b, int dx, int dy, int dz, int nx, int ny, int nz,
                       const int per[3]) {
    int xx = x + dx, yy = y + dy, zz = z + dz;
    if (xx < 0 || xx >= nx) {
        if (!per[0]) return -1;
        xx = (xx + nx) % nx;
    }
    if (yy < 0 || yy >= ny) {
        if (!per[1]) return -1;
        yy = (yy + ny) % ny;
    }
    if (zz < 0 || zz >= nz) {
        if (!per[2]) return -1;
        zz = (zz + nz) % nz;
    }
    return (((long)b * nz + zz) * ny + yy) * nx + xx;
}

void orc_conv3d_fwd(const double* in, const double* w, const double* bias, double* out, int nx,
                    int ny, int nz, int batch, int cin, int cout, const int per[3], int relu,
                    int bf16) {
    const long nv = (long)nx * ny * nz * batch;
    const long K = 27L * cin;
    double* wr = (double*)malloc(sizeof(double) * cout * K);
    for (long q = 0; q < cout * K; ++q) wr[q] = R(w[q], bf16);
    /* threads only for work that outweighs their start-up (tiny finite-difference cases) */
#pragma omp parallel for schedule(static) if (nv * K * cout > 4000000L)
    for (long v = 0; v < nv; ++v) {
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)((v / ((long)nx * ny)) % nz);
        const int b = (int)(v / ((long)nx * ny * nz));
        double xin[27 * 256];
        for (int t = 0; t < 27; ++t) {
            const long u = nbr(x, y, z, b, t % 3 - 1, (t / 3) % 3 - 1, t / 9 - 1, nx, ny, nz, per);
            for (int ci = 0; ci < cin; ++ci) xin[t * cin + ci] = u < 0 ? 0.0 : R(in[u * cin + ci], bf16);
        }
        for (int co = 0; co < cout; ++co) {
            double s = 0.0;
            const double* wc = wr + (long)co * K;
            for (long k = 0; k < K; ++k) s += wc[k] * xin[k];
            s += bias ? bias[co] : 0.0;
            out[v * cout + co] = (relu && s < 0.0) ? 0.0 : s;
        }
    }
    free(wr);
}

void orc_conv3d_bwd(const double* in, const double* w, const double* dout, double* din,
                    double* dw, double* db, int nx, int ny, int nz, int batch, int cin, int cout,
                    const int per[3], int bf16) {
    const long nv = (long)nx * ny * nz * batch;
    const long K = 27L * cin;
    if (din) {
        /* din[u][ci] = sum_{tap,co} w[co][tap,ci] dout[u - off(tap)][co] */
#pragma omp parallel for schedule(static) if (nv * K * cout > 4000000L)
        for (long u = 0; u < nv; ++u) {
            const int x = (int)(u % nx), y = (int)((u / nx) % ny), z = (int)((u / ((long)nx * ny)) % nz);
            const int b = (int)(u / ((long)nx * ny * nz));
            for (int ci = 0; ci < cin; ++ci) din[u * cin + ci] = 0.0;
            for (int t = 0; t < 27; ++t) {
                const long v = nbr(x, y, z, b, -(t % 3 - 1), -((t / 3) % 3 - 1), -(t / 9 - 1), nx, ny, nz, per);
                if (v < 0) continue;
                for (int co = 0; co < cout; ++co) {
                    const double g = R(dout[v * cout + co], bf16);
                    const double* wc = w + (long)co * K + (long)t * cin;
                    for (int ci = 0; ci < cin; ++ci) din[u * cin + ci] += R(wc[ci], bf16) * g;
                }
            }
        }
    }
    if (dw) {
#pragma omp parallel for schedule(static) if (nv * K * cout > 4000000L)
        for (int co = 0; co < cout; ++co) {
            double* acc = dw + (long)co * K;
            for (long k = 0; k < K; ++k) acc[k] = 0.0;
            for (long v = 0; v < nv; ++v) {
                const double g = R(dout[v * cout + co], bf16);
                if (g == 0.0) continue;
                const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)((v / ((long)nx * ny)) % nz);
                const int b = (int)(v / ((long)nx * ny * nz));
                for (int t = 0; t < 27; ++t) {
                    const long u = nbr(x, y, z, b, t % 3 - 1, (t / 3) % 3 - 1, t / 9 - 1, nx, ny, nz, per);
                    if (u < 0) continue;
                    for (int ci = 0; ci < cin; ++ci) acc[t * cin + ci] += g * R(in[u * cin + ci], bf16);
                }
            }
        }
    }
    if (db) {
        for (int co = 0; co < cout; ++co) db[co] = 0.0;
        for (long v = 0; v < nv; ++v)
            for (int co = 0; co < cout; ++co) db[co] += dout[v * cout + co];
    }
}

/* ---- U-net --------------------------------------------------------------- */

typedef struct {
    int cin, cout;
    long woff, boff;
} layer_t;

static int unet_layers(int base, int levels, int cin0, layer_t* Ls, long* nparam) {
    int n = 0;
    long off = 0;
#define ADD(ci, co)                         \
    do {                                    \
        Ls[n].cin = (ci);                   \
        Ls[n].cout = (co);                  \
        Ls[n].woff = off;                   \
        off += (long)(co) * 27 * (ci);      \
        Ls[n].boff = off;                   \
        off += (co);                        \
        ++n;                                \
    } while (0)
    for (int l = 0; l < levels - 1; ++l) {
        ADD(l == 0 ? cin0 : base << (l - 1), base << l);
        ADD(base << l, base << l);
    }
    ADD(base << (levels - 2), base << (levels - 1));
    ADD(base << (levels - 1), base << (levels - 1));
    for (int l = levels - 2; l >= 0; --l) {
        ADD((base << (l + 1)) + (base << l), base << l);
        ADD(base << l, base << l);
    }
#undef ADD
    *nparam = off + base + 1;
    return n;
}

long orc_unet_nparam(int base, int levels, int cin0) {
    layer_t Ls[32];
    long np;
    unet_layers(base, levels, cin0, Ls, &np);
    return np;
}

static void avgpool(const double* in, int nx, int ny, int nz, int c, int batch, double* out) {
    const int mx = nx / 2, my = ny / 2, mz = nz / 2;
    for (int b = 0; b < batch; ++b)
        for (int z = 0; z < mz; ++z)
            for (int y = 0; y < my; ++y)
                for (int x = 0; x < mx; ++x)
                    for (int ch = 0; ch < c; ++ch) {
                        double s = 0.0;
                        for (int q = 0; q < 8; ++q) {
                            const long u = (((long)b * nz + 2 * z + (q >> 2)) * ny + 2 * y + ((q >> 1) & 1)) * nx + 2 * x + (q & 1);
                            s += in[u * c + ch];
                        }
                        out[((((long)b * mz + z) * my + y) * mx + x) * c + ch] = s / 8.0;
                    }
}

/* concat [upsample2(coarse) (cc channels), skip (cs channels)] at fine size */
static void up_concat(const double* coarse, int cc, const double* skip, int cs, int nx, int ny,
                      int nz, int batch, double* out) {
    const int ct = cc + cs;
    for (int b = 0; b < batch; ++b)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const long v = (((long)b * nz + z) * ny + y) * nx + x;
                    const long u = (((long)b * (nz / 2) + z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2;
                    for (int ch = 0; ch < cc; ++ch) out[v * ct + ch] = coarse[u * cc + ch];
                    for (int ch = 0; ch < cs; ++ch) out[v * ct + cc + ch] = skip[v * cs + ch];
                }
}

typedef struct {
    int nl;
    layer_t L[32];
    long nparam;
    double* x[32];  /* input of conv layer i */
    double* y[32];  /* output (post-ReLU) of conv layer i */
    int lvl[32];
    double* pooled[8];
} unet_run;

static long nvox_l(int nx, int ny, int nz, int batch, int l) {
    return (long)(nx >> l) * (ny >> l) * (nz >> l) * batch;
}

static void unet_fwd(unet_run* U, const double* p, int base, int levels, int cin0,
                     const double* in, double* out, int nx, int ny, int nz, int batch,
                     const int per[3], int bf16) {
    U->nl = unet_layers(base, levels, cin0, U->L, &U->nparam);
    int li = 0;
    const double* cur = in;
    for (int l = 0; l < levels - 1; ++l) {
        for (int q = 0; q < 2; ++q, ++li) {
            const layer_t* L = &U->L[li];
            const long nv = nvox_l(nx, ny, nz, batch, l);
            U->x[li] = (double*)cur;
            U->y[li] = (double*)malloc(sizeof(double) * nv * L->cout);
            U->lvl[li] = l;
            orc_conv3d_fwd(cur, p + L->woff, p + L->boff, U->y[li], nx >> l, ny >> l, nz >> l, batch,
                           L->cin, L->cout, per, 1, bf16);
            cur = U->y[li];
        }
        const int c = U->L[li - 1].cout;
        U->pooled[l] = (double*)malloc(sizeof(double) * nvox_l(nx, ny, nz, batch, l + 1) * c);
        avgpool(cur, nx >> l, ny >> l, nz >> l, c, batch, U->pooled[l]);
        cur = U->pooled[l];
    }
    for (int q = 0; q < 2; ++q, ++li) {
        const int l = levels - 1;
        const layer_t* L = &U->L[li];
        U->x[li] = (double*)cur;
        U->y[li] = (double*)malloc(sizeof(double) * nvox_l(nx, ny, nz, batch, l) * L->cout);
        U->lvl[li] = l;
        orc_conv3d_fwd(cur, p + L->woff, p + L->boff, U->y[li], nx >> l, ny >> l, nz >> l, batch,
                       L->cin, L->cout, per, 1, bf16);
        cur = U->y[li];
    }
    for (int l = levels - 2; l >= 0; --l) {
        const int cc = U->L[li - 1].cout, cs = base << l;
        const long nv = nvox_l(nx, ny, nz, batch, l);
        double* cat = (double*)malloc(sizeof(double) * nv * (cc + cs));
        up_concat(cur, cc, U->y[2 * l + 1], cs, nx >> l, ny >> l, nz >> l, batch, cat);
        for (int q = 0; q < 2; ++q, ++li) {
            const layer_t* L = &U->L[li];
            U->x[li] = q == 0 ? cat : (double*)cur;
            U->y[li] = (double*)malloc(sizeof(double) * nv * L->cout);
            U->lvl[li] = l;
            orc_conv3d_fwd(U->x[li], p + L->woff, p + L->boff, U->y[li], nx >> l, ny >> l, nz >> l,
                           batch, L->cin, L->cout, per, 1, bf16);
            cur = U->y[li];
        }
    }
    const long hw = U->nparam - base - 1, hb = U->nparam - 1;
    const long n0 = nvox_l(nx, ny, nz, batch, 0);
    for (long v = 0; v < n0; ++v) {
        double s = p[hb];
        for (int c = 0; c < base; ++c) s += p[hw + c] * cur[v * base + c];
        out[v] = s;
    }
}

static void unet_free(unet_run* U, int levels) {
    for (int i = 0; i < U->nl; ++i) free(U->y[i]);
    /* decoder concat buffers are the x of each decoder level's first conv */
    for (int l = levels - 2; l >= 0; --l) free(U->x[U->nl - 2 - 2 * l]);
    for (int l = 0; l < levels - 1; ++l) free(U->pooled[l]);
}

void orc_unet_forward(const double* p, int base, int levels, int cin0, const double* in,
                      double* out, int nx, int ny, int nz, int batch, const int per[3], int bf16) {
    unet_run U;
    unet_fwd(&U, p, base, levels, cin0, in, out, nx, ny, nz, batch, per, bf16);
    unet_free(&U, levels);
}

/* gradients of sum_v dout[v] out[v] with respect to the parameters */
void orc_unet_grad(const double* p, int base, int levels, int cin0, const double* in,
                   const double* dout, double* grad, int nx, int ny, int nz, int batch,
                   const int per[3], int bf16) {
    unet_run U;
    const long n0 = nvox_l(nx, ny, nz, batch, 0);
    double* out = (double*)malloc(sizeof(double) * n0);
    unet_fwd(&U, p, base, levels, cin0, in, out, nx, ny, nz, batch, per, bf16);
    free(out);
    memset(grad, 0, sizeof(double) * U.nparam);
    const long hw = U.nparam - base - 1, hb = U.nparam - 1;
    /* gy[i]: gradient w.r.t. y[i] */
    double* gy[32] = {0};
    for (int i = 0; i < U.nl; ++i)
        gy[i] = (double*)calloc(nvox_l(nx, ny, nz, batch, U.lvl[i]) * U.L[i].cout, sizeof(double));
    {
        const double* yl = U.y[U.nl - 1];
        double* g = gy[U.nl - 1];
        for (long v = 0; v < n0; ++v) {
            grad[hb] += dout[v];
            for (int c = 0; c < base; ++c) {
                grad[hw + c] += dout[v] * yl[v * base + c];
                g[v * base + c] = dout[v] * p[hw + c];
            }
        }
    }
    for (int i = U.nl - 1; i >= 0; --i) {
        const layer_t* L = &U.L[i];
        const int l = U.lvl[i];
        const long nv = nvox_l(nx, ny, nz, batch, l);
        double* dy = gy[i];
        for (long q = 0; q < nv * L->cout; ++q)
            if (U.y[i][q] <= 0.0) dy[q] = 0.0; /* ReLU */
        double* dx = (i == 0) ? NULL : (double*)calloc(nv * L->cin, sizeof(double));
        double* dw = (double*)malloc(sizeof(double) * L->cout * 27 * L->cin);
        double* db = (double*)malloc(sizeof(double) * L->cout);
        orc_conv3d_bwd(U.x[i], p + L->woff, dy, dx, dw, db, nx >> l, ny >> l, nz >> l, batch,
                       L->cin, L->cout, per, bf16);
        for (long q = 0; q < (long)L->cout * 27 * L->cin; ++q) grad[L->woff + q] += dw[q];
        for (int q = 0; q < L->cout; ++q) grad[L->boff + q] += db[q];
        free(dw);
        free(db);
        if (!dx) continue;
        /* route dx to whatever produced x[i] */
        const int nenc = 2 * (levels - 1);
        if (i >= nenc + 2 && ((i - nenc - 2) % 2 == 0)) {
            /* decoder first conv: x = [upsample(prev y), skip] */
            const int lev = l;
            const int cc = L->cin - (base << lev), cs = base << lev;
            const int src = i - 1;              /* previous layer output (coarser) */
            const int skip = 2 * lev + 1;       /* encoder second conv at this level */
            const int mx = nx >> (lev + 1), my = ny >> (lev + 1), mz = nz >> (lev + 1);
            for (int b = 0; b < batch; ++b)
                for (int z = 0; z < (nz >> lev); ++z)
                    for (int y = 0; y < (ny >> lev); ++y)
                        for (int x = 0; x < (nx >> lev); ++x) {
                            const long v = (((long)b * (nz >> lev) + z) * (ny >> lev) + y) * (nx >> lev) + x;
                            const long u = (((long)b * mz + z / 2) * my + y / 2) * mx + x / 2;
                            for (int ch = 0; ch < cc; ++ch) gy[src][u * cc + ch] += dx[v * L->cin + ch];
                            for (int ch = 0; ch < cs; ++ch) gy[skip][v * cs + ch] += dx[v * L->cin + cc + ch];
                        }
        } else if (i < nenc + 2 && (i % 2 == 0)) {
            /* first conv of an encoder level (or bottleneck): x = pooled of level l-1 */
            const int pl = l - 1;
            const int c = base << pl;
            /* avgpool backward into the skip output of level pl */
            const int fx = nx >> pl, fy = ny >> pl, fz = nz >> pl;
            const int src = 2 * pl + 1;
            for (int b = 0; b < batch; ++b)
                for (int z = 0; z < fz; ++z)
                    for (int y = 0; y < fy; ++y)
                        for (int x = 0; x < fx; ++x) {
                            const long v = (((long)b * fz + z) * fy + y) * fx + x;
                            const long u = (((long)b * (fz / 2) + z / 2) * (fy / 2) + y / 2) * (fx / 2) + x / 2;
                            for (int ch = 0; ch < c; ++ch) gy[src][v * c + ch] += dx[u * c + ch] / 8.0;
                        }
        } else {
            /* second conv of a pair: x = y of the previous layer */
            const int src = i - 1;
            for (long q = 0; q < nv * L->cin; ++q) gy[src][q] += dx[q];
        }
        free(dx);
    }
    for (int i = 0; i < U.nl; ++i) free(gy[i]);
    unet_free(&U, levels);
}

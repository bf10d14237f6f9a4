// learned_kernels.h -- glue between the nodal multigrid hierarchy and the
// learned coarse-level predictor: feature packing, scaling and the training
// target.  Normalisation (identical for training and inference):
//   input  ch0 = b_c / rms(b_c),   ch1 = log(sigma_c) - mean(log sigma_c)
//   output e_c = net * rms(b_c) * h_c^2 / exp(mean(log sigma_c))
// so the network sees a dimensionless problem whatever the grid spacing,
// time step or density scale.  Periodic levels only (nodes == cells).
#pragma once

#include <hip/hip_runtime.h>

#include "nodal_kernels.h"

namespace mgx {

// per-block partial sums of a[i]^2 (mode 0) or log(sigma cell) (mode 1)
template <class T>
static __global__ void __launch_bounds__(256) feature_partials_kernel(KGeom g, const T* __restrict__ b,
                                                                      const T* __restrict__ sig,
                                                                      int mode, double* __restrict__ part) {
    __shared__ double red[256];
    const long n = (long)g.nn[0] * g.nn[1] * g.nn[2];
    double s = 0.0;
    for (long id = (long)blockIdx.x * 256 + threadIdx.x; id < n; id += (long)gridDim.x * 256) {
        if (mode == 0) {
            const double v = (double)b[id];
            s += v * v;
        } else {
            const int i = id % g.nn[0];
            const int j = (id / g.nn[0]) % g.nn[1];
            const int k = id / ((long)g.nn[0] * g.nn[1]);
            s += log((double)sig[(i + 1) + (long)g.sg[0] * ((j + 1) + (long)g.sg[1] * (k + 1))]);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// sc[0] = mean log sigma (mode 1) ; for mode 0: sc[1] = 1/rms, sc[2] = rms*h2/exp(sc[0])
static __global__ void __launch_bounds__(256) feature_scalars_kernel(const double* __restrict__ part,
                                                                     int np, long n, int mode,
                                                                     double h2, double* sc) {
    __shared__ double red[256];
    double s = 0.0;
    for (int id = threadIdx.x; id < np; id += 256) s += part[id];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == 1) {
            sc[0] = red[0] / (double)n;
        } else {
            const double rms = sqrt(red[0] / (double)n);
            sc[1] = rms > 0.0 ? 1.0 / rms : 0.0;
            sc[2] = rms * h2 / exp(sc[0]);
        }
    }
}

// in[v][8]: ch0 = b*sc[1], ch1 = log sigma - sc[0], rest 0
template <class T>
static __global__ void pack_features_kernel(KGeom g, const T* __restrict__ b, const T* __restrict__ sig,
                                            const double* __restrict__ sc, float* __restrict__ in) {
    int i, j, k;
    if (!grid_point(g.nn[0], g.nn[1], i, j, k)) return;
    const long id = i + (long)g.nn[0] * (j + (long)g.nn[1] * k);
    const double s = (double)sig[(i + 1) + (long)g.sg[0] * ((j + 1) + (long)g.sg[1] * (k + 1))];
    float* o = in + id * 8;
    o[0] = (float)((double)b[id] * sc[1]);
    o[1] = (float)(log(s) - sc[0]);
#pragma unroll
    for (int c = 2; c < 8; ++c) o[c] = 0.f;
}

// x_c = net_out * sc[2]
template <class T>
static __global__ void unscale_kernel(const float* __restrict__ out, long n, const double* __restrict__ sc,
                                      T* __restrict__ x) {
    const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) x[id] = (T)((double)out[id] * sc[2]);
}

// target[I] = phi[(2^L) I] / sc[2]   (fine nodes -> coarse level L, periodic)
static __global__ void inject_target_kernel(KGeom gf, KGeom gc, int stride, const double* __restrict__ phi,
                                            const double* __restrict__ sc, float* __restrict__ t) {
    int I, J, K;
    if (!grid_point(gc.nn[0], gc.nn[1], I, J, K)) return;
    const double v = phi[(long)stride * I + (long)gf.nn[0] * ((long)stride * J + (long)gf.nn[1] * (long)stride * K)];
    t[I + (long)gc.nn[0] * (J + (long)gc.nn[1] * K)] = sc[2] > 0.0 ? (float)(v / sc[2]) : 0.f;
}

}  // namespace mgx

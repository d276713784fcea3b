// Device vector algebra of the minimiser (Min_LBFGS, src/min_l-bfgs.cpp:302-787,
// and LineSearch_BackTrack, src/line_search.cpp:244-341): dot products, axpys,
// scalings, max |h| and min alpha over the owned (3+nt)-column rows of x / f / h,
// all FP64.  Dot products are deterministic: fixed per-block trees, then the
// two-stage partial reduction.
#pragma once
#include "force.hpp"

namespace mappdmd {

struct DotArgs {
    const double* a[4];
    const double* b[4];
};

// ND dot products <a_k, b_k> over n entries: per-block partial sums
template <int ND>
__global__ __launch_bounds__(kBlock) void k_dots(DotArgs D, long n, double* __restrict__ part) {
    double acc[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) acc[k] = 0.0;
    const long stride = (long)gridDim.x * kBlock;
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
#pragma unroll
        for (int k = 0; k < ND; ++k) acc[k] += D.a[k][e] * D.b[k][e];
    }
    block_sum_store<ND>(acc, part);
}

// y = a * x + b * y
__global__ void k_axpby(long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = a * x[e] + b * y[e];
}

// y = x0 + a * h
__global__ void k_step(long n, const double* __restrict__ x0, double a, const double* __restrict__ h,
                       double* __restrict__ y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = x0[e] + a * h[e];
}

// s = a * h, y = f - f0
__global__ void k_sy(long n, double a, const double* __restrict__ h, const double* __restrict__ f,
                     const double* __restrict__ f0, double* __restrict__ s, double* __restrict__ y) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) {
        s[e] = a * h[e];
        y[e] = f[e] - f0[e];
    }
}

// per-block max |h| and min alpha (alpha columns 3..xd-1 of the owned rows of x)
__global__ __launch_bounds__(kBlock) void k_extrema(long n, int xd, const double* __restrict__ h,
                                                    const double* __restrict__ x, double* __restrict__ part) {
    __shared__ double smax[kBlock], smin[kBlock];
    double mx = 0.0, mn = INFINITY;
    const long stride = (long)gridDim.x * kBlock;
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
        mx = fmax(mx, fabs(h[e]));
        if ((int)(e % xd) >= 3) mn = fmin(mn, x[e]);
    }
    smax[threadIdx.x] = mx;
    smin[threadIdx.x] = mn;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = smax[0];
        part[blockIdx.x * 2 + 1] = smin[0];
    }
}

}  // namespace mappdmd

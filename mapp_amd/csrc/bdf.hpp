// Device vector algebra of the BDF clock (Clock_BDF, src/clock_bdf.cpp:295-883):
// the interpolation of the predictor and of the residual's constant part,
// the step bounds that keep c inside [0,1], the error sums and the order
// estimates.  FP64, no contraction (same rounding as the host loops they
// replace); sums are deterministic per-block trees.
#pragma once
#include "force.hpp"

namespace mappdmd {

constexpr int kBdfMaxOrder = 7;

struct BdfHist {
    const double* y[kBdfMaxOrder];
};

struct BdfInterp {
    double b_ady0;                // beta * dalpha_dy_0
    double ady0;                  // alpha_dy_0
    double b_ay[kBdfMaxOrder];    // beta * dalpha_y[j]
    double ay[kBdfMaxOrder];      // alpha_y[j]
};

// Clock_BDF::interpolate per dof (src/clock_bdf.cpp:345-362): a, y_0; flag when y_0 leaves [0,1]
__global__ void k_bdf_interp(long n, int q, BdfInterp C, BdfHist Y, const double* __restrict__ dy,
                             double* __restrict__ a, double* __restrict__ y0, int* __restrict__ flag) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double ak = C.b_ady0 * dy[k];
    double yk = C.ady0 * dy[k];
    for (int j = 0; j < q; ++j) {
        ak += C.b_ay[j] * Y.y[j][k];
        yk += C.ay[j] * Y.y[j][k];
    }
    ak -= yk;
    a[k] = ak;
    y0[k] = yk;
    if (yk < 0.0 || yk > 1.0) atomicOr(flag, 1);
}

// largest fraction of the step base + scale*dir that stays inside [0,1]
// (src/clock_bdf.cpp:704-712 predictor, :740-752 CG step): per-block minima
__global__ __launch_bounds__(kBlock) void k_bdf_bound(long n, const double* __restrict__ base,
                                                      const double* __restrict__ dir, double scale,
                                                      double* __restrict__ part) {
    __shared__ double sm[kBlock];
    double r = INFINITY;
    const long stride = (long)gridDim.x * kBlock;
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
        const double b = base[e];
        const double tmp = dir[e] * scale + b;
        if (tmp > 1.0) r = fmin(r, (1.0 - b) / (tmp - b));
        else if (tmp < 0.0) r = fmin(r, (b - 0.0) / (b - tmp));
    }
    sm[threadIdx.x] = r;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// c = y0 + dy * del_t * ratio  (src/clock_bdf.cpp:714-715)
__global__ void k_bdf_pred(long n, const double* __restrict__ y, const double* __restrict__ dy, double del_t,
                           double ratio, double* __restrict__ c) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) c[k] = y[k] + dy[k] * del_t * ratio;
}

// sums of (y_0 - c)^2 and c_d^2 (src/clock_bdf.cpp:857-868)
__global__ __launch_bounds__(kBlock) void k_bdf_err(long n, const double* __restrict__ y0,
                                                    const double* __restrict__ c, const double* __restrict__ cd,
                                                    double* __restrict__ part) {
    double acc[2] = {0.0, 0.0};
    const long stride = (long)gridDim.x * kBlock;
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
        const double d = y0[e] - c[e];
        acc[0] += d * d;
        acc[1] += cd[e] * cd[e];
    }
    block_sum_store<2>(acc, part);
}

struct BdfHiLo {
    double Q_n, hi_pre;                          // order q+1
    double la_y, la_dy, la[kBdfMaxOrder], c1c0;  // order q-1
    int hi, lo, q;
};

// Clock_BDF::hi_lo_err_est per dof (src/clock_bdf.cpp:444-468): sums for
// hi_err and lo_err, then e_n = c - y_0
__global__ __launch_bounds__(kBlock) void k_bdf_hilo(long n, BdfHiLo H, BdfHist Y, const double* __restrict__ c,
                                                     const double* __restrict__ cd, const double* __restrict__ y0,
                                                     double* __restrict__ e_n, double* __restrict__ part) {
    double acc[2] = {0.0, 0.0};  // lo, hi
    const long stride = (long)gridDim.x * kBlock;
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        if (H.hi) {
            double t0 = (c[i] - y0[i]) - H.Q_n * e_n[i];
            t0 *= H.hi_pre;
            acc[1] += t0 * t0;
        }
        if (H.lo) {
            double t0 = H.la_y * c[i];
            t0 += H.la_dy * cd[i];
            for (int j = 0; j < H.q - 1; ++j) t0 += H.la[j] * Y.y[j][i];
            t0 *= H.c1c0;
            acc[0] += t0 * t0;
        }
        e_n[i] = c[i] - y0[i];
    }
    block_sum_store<2>(acc, part);
}

// per-block min and max of v (the domain check of src/clock_bdf.cpp:875-877)
__global__ __launch_bounds__(kBlock) void k_minmax(long n, const double* __restrict__ v, double* __restrict__ part) {
    __shared__ double smin[kBlock], smax[kBlock];
    double mn = INFINITY, mx = -INFINITY;
    const long stride = (long)gridDim.x * kBlock;
    for (long e = (long)blockIdx.x * kBlock + threadIdx.x; e < n; e += stride) {
        mn = fmin(mn, v[e]);
        mx = fmax(mx, v[e]);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = smin[0];
        part[blockIdx.x * 2 + 1] = smax[0];
    }
}

}  // namespace mappdmd

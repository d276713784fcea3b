// Learned surrogate of the Gauss-Hermite averaged primitives.
//
// One MLP per ordered type pair (it, jt) maps (r, alpha_i^it, alpha_j^jt) to
// {rho_bar, d rho_bar/dr, d rho_bar/dalpha, phi_bar, d phi_bar/dr,
// d phi_bar/dalpha} -- exactly what pass 1 of the force computes per
// (pair, type pair) by quadrature (src/ff_dmd_I.cpp:180-242).  Shape
// 3 -> 64 -> 64 -> 64 -> 6 with SiLU, inputs standardised.  The networks do
// not learn the averages themselves but their residual against a cheap
// baseline read from one table row (second-order Gaussian smoothing of the
// tabulated functions at t = r, sur_baseline), divided by an r-envelope of
// that residual (so the networks' error is relative to the residual's local
// size, which falls by orders of magnitude towards the cutoff):
//     y = baseline(r, alpha) + s(r) * net(r, alpha_i, alpha_j).
//
//   k_sur_samples   training / validation samples labelled by the FP32
//                   quadrature itself (the mixed path's tables and rule),
//                   with the baseline of each sample
//   k_sur_train     one micro-batch of 64 samples of one type pair per block:
//                   forward, loss, backward in FP32 with everything in LDS;
//                   writes the block's gradient (fixed-order sums, no atomics)
//   k_sur_adam      sums block gradients in block order, Adam update
//   k_sur_judge     the inference path itself (BF16 MFMA networks, baseline,
//                   envelope) on validation samples: absolute error of every
//                   output against the quadrature, from which the host marks
//                   the (r, alpha_i, alpha_j) cells that meet the per-pair
//                   tolerances derived from the force bound
//   k_pass1_sur     pass 1 of the force with the networks on BF16 MFMA
//                   (v_mfma_f32_16x16x32_bf16, FP32 accumulate): one wavefront
//                   per site, lane = type pair * 16 + pair slot, so each MFMA
//                   M tile holds the 16 samples of one type pair; samples
//                   outside the trained box or in a cell that failed the
//                   tolerance fall back to the quadrature, and a network whose
//                   16 samples all fall back is not run.
#pragma once
#include "force.hpp"
#include "surrogate_types.hpp"
#include <type_traits>

#pragma clang fp contract(fast)

namespace mappdmd {

constexpr int kHid = 64, kIn = 3, kOut = 6;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kIn * kHid;
constexpr int kW2 = kB1 + kHid;
constexpr int kB2 = kW2 + kHid * kHid;
constexpr int kW3 = kB2 + kHid;
constexpr int kB3 = kW3 + kHid * kHid;
constexpr int kW4 = kB3 + kHid;
constexpr int kB4 = kW4 + kHid * kOut;
constexpr int kNParam = kB4 + kOut;  // 8966 per network
constexpr int kMB = 64;              // training micro-batch per block
constexpr int kCellR = 64, kCellA = 4;  // fallback cells: r x alpha_i x alpha_j


__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + __expf(-z)); }
// SiLU for BF16 activations: the hardware reciprocal (1 ulp) in place of the
// correctly rounded division, far below the 2^-9 rounding of the BF16 result
__device__ __forceinline__ float silu_bf16(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
// Inference carries every hidden pre-activation scaled by kKappa = -log2(e):
// with z' = kKappa z, SiLU(z) kKappa = z' / (1 + 2^z'), one exp2 and one
// reciprocal.  The scale is folded into the layer-1 weights and the layer-2/3
// biases; the 64 -> 64 weights see kKappa-scaled inputs and give
// kKappa-scaled outputs unchanged, and the linear output layer divides once.
constexpr float kKappa = -1.44269504088896341f, kInvKappa = -0.693147180559945309f;
__device__ __forceinline__ float silu_scaled(float zs) { return zs * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(zs)); }
__device__ __forceinline__ float silu_df(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double u01(unsigned long long h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// Baseline of the six outputs from the table row at t = r (the middle node of an
// odd rule).  For a tabulated g (g(u) = u rho(u) for the density, the stored
// r phi for the pair term) with derivatives g1, g2, g3, the Gaussian average
// with variance 1/(2 alpha) is g + g2/(4 alpha) + O(alpha^-2), so
//     b = (g + g2/(4 alpha)) / r,   db/dr = (g1 + g3/(4 alpha)) / r - b / r,
//     db/dalpha = -g2 / (4 alpha^2 r),
// with the derivatives of the row's cubic in p = r/dr - m.  Zero at and beyond
// the cutoff.  ia = 1/alpha.
__device__ __forceinline__ void sur_baseline(const FFParams<float>& P, const float* __restrict__ tab, int stride,
                                             float r, float ia, float* B) {
    if (!(r < P.rc)) {
#pragma unroll
        for (int c = 0; c < 6; ++c) B[c] = 0.0f;
        return;
    }
    float p = r * P.dr_inv;
    int m = (int)p;
    m = m < P.nr - 2 ? m : P.nr - 2;
    m = m > 0 ? m : 0;
    p -= (float)m;
    p = p < 1.0f ? p : 1.0f;
    const float h = P.dr_inv;
    const float4 cr = *reinterpret_cast<const float4*>(tab + (size_t)m * stride);
    const float4 cp = *reinterpret_cast<const float4*>(tab + (size_t)m * stride + 4);
    // value and first three r-derivatives of a cubic row
    auto d4 = [&](float4 c, float* d) {
        d[0] = ((c.w * p + c.z) * p + c.y) * p + c.x;
        d[1] = ((3.0f * c.w * p + 2.0f * c.z) * p + c.y) * h;
        d[2] = (6.0f * c.w * p + 2.0f * c.z) * h * h;
        d[3] = 6.0f * c.w * h * h * h;
    };
    float v[4], w[4];
    d4(cr, v);
    d4(cp, w);
    const float g0 = r * v[0], g1 = v[0] + r * v[1], g2 = 2.0f * v[1] + r * v[2], g3 = 3.0f * v[2] + r * v[3];
    const float q = 0.25f * ia, ri = 1.0f / r;
    B[0] = (g0 + q * g2) * ri;
    B[1] = (g1 + q * g3) * ri - B[0] * ri;
    B[2] = -4.0f * q * q * g2 * ri;
    B[3] = (w[0] + q * w[2]) * ri;
    B[4] = (w[1] + q * w[3]) * ri - B[3] * ri;
    B[5] = -4.0f * q * q * w[2] * ri;
}

// envelope s(r) of output c of type pair tp: log2 s interpolated linearly
// between the centres of kEnvBins equal r bins of the trained box
__host__ __device__ inline float sur_env_log2(const float* __restrict__ lenv, const SurNorm& N, int tp, int c,
                                              float r) {
    const float* L = lenv + ((size_t)tp * kSurOut + c) * kEnvBins;
    float u = (r - N.r_lo) / (N.r_hi - N.r_lo) * (float)kEnvBins - 0.5f;
    int b = (int)floorf(u);
    b = b < 0 ? 0 : (b > kEnvBins - 2 ? kEnvBins - 2 : b);
    float f = u - (float)b;
    f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
    return L[b] + f * (L[b + 1] - L[b]);
}

// samples n_per_tp per type pair: X[s] = {r, alpha_i, alpha_j}, T[s] = quadrature
// outputs, B[s] = their baseline
template <int NT, int NQ>
__global__ void k_sur_samples(FFParams<float> PQ, SurNorm N, const float* __restrict__ quad, long n_per_tp,
                              unsigned long long seed, float* __restrict__ X, float* __restrict__ T,
                              float* __restrict__ B) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_per_tp * NT * NT) return;
    const int tp = (int)(s / n_per_tp);
    const unsigned long long h = splitmix64(seed ^ (unsigned long long)s * 0x632BE59BD9B4E019ull);
    const double r = N.r_lo + (N.r_hi - N.r_lo) * u01(h);
    const double ai = N.a_lo + (N.a_hi - N.a_lo) * u01(splitmix64(h + 1));
    const double aj = N.a_lo + (N.a_hi - N.a_lo) * u01(splitmix64(h + 2));
    const double ia = 1.0 / ai + 1.0 / aj;
    const float asq = (float)rsqrt(ia);
    float g[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    const float* tab = quad + (size_t)tp * kQuadRow;
    gh_avg<float, NQ, true>(PQ, (float)r, (float)(1.0 / r), asq, (float)ia, tab, quad_stride(NT), g, f);
    X[s * 3 + 0] = (float)r;
    X[s * 3 + 1] = (float)ai;
    X[s * 3 + 2] = (float)aj;
    for (int k = 0; k < 3; ++k) {
        T[s * kOut + k] = g[k];
        T[s * kOut + 3 + k] = f[k];
    }
    float b[6];
    sur_baseline(PQ, tab, quad_stride(NT), (float)r, (float)ia, b);
    for (int k = 0; k < 6; ++k) B[s * kOut + k] = b[k];
}

// one dense layer over the block's micro-batch: out[s][o] = b[o] + sum_k in[s][k] W[k][o]
template <int K>
__device__ __forceinline__ void dense_fwd(const float* W, const float* b, const float* in, int in_stride, float* z,
                                          float* h) {
    for (int e = threadIdx.x; e < kMB * kHid; e += blockDim.x) {
        const int s = e / kHid, o = e % kHid;
        float acc = b[o];
#pragma unroll 8
        for (int k = 0; k < K; ++k) acc += in[s * in_stride + k] * W[k * kHid + o];
        z[s * kHid + o] = acc;
        h[s * kHid + o] = silu_f(acc);
    }
}

// forward + backward of one micro-batch (64 samples of type pair tp); the
// block's gradient goes to gpart[block][kNParam], its loss sum to lpart[block]
__global__ __launch_bounds__(256) void k_sur_train(const float* __restrict__ params, SurNorm N,
                                                   const float* __restrict__ X, const float* __restrict__ T,
                                                   const int* __restrict__ idx, int blocks_per_tp, float inv_count,
                                                   float* __restrict__ gpart, float* __restrict__ lpart) {
    extern __shared__ float sm[];
    float* W = sm;                       // kNParam
    float* H0 = W + kNParam + 2;         // [64][4]
    float* Z1 = H0 + kMB * 4;            // [64][64] each
    float* Z2 = Z1 + kMB * kHid;
    float* Z3 = Z2 + kMB * kHid;
    float* H1 = Z3 + kMB * kHid;
    float* H2 = H1 + kMB * kHid;
    float* H3 = H2 + kMB * kHid;
    float* DY = H3 + kMB * kHid;         // [64][8]
    float* red = DY + kMB * 8;           // [256]
    const int tp = blockIdx.x / blocks_per_tp;
    const float* P = params + (size_t)tp * kNParam;
    for (int e = threadIdx.x; e < kNParam; e += blockDim.x) W[e] = P[e];
    for (int s = threadIdx.x; s < kMB; s += blockDim.x) {
        const int id = idx[(size_t)blockIdx.x * kMB + s];
        H0[s * 4 + 0] = (X[(size_t)id * 3 + 0] - N.r_c) / N.r_s;
        H0[s * 4 + 1] = (X[(size_t)id * 3 + 1] - N.a_c) / N.a_s;
        H0[s * 4 + 2] = (X[(size_t)id * 3 + 2] - N.a_c) / N.a_s;
        H0[s * 4 + 3] = 0.0f;
    }
    __syncthreads();
    dense_fwd<kIn>(W + kW1, W + kB1, H0, 4, Z1, H1);
    __syncthreads();
    dense_fwd<kHid>(W + kW2, W + kB2, H1, kHid, Z2, H2);
    __syncthreads();
    dense_fwd<kHid>(W + kW3, W + kB3, H2, kHid, Z3, H3);
    __syncthreads();
    // output layer, loss and dL/dy
    float lsum = 0.0f;
    for (int e = threadIdx.x; e < kMB * kOut; e += blockDim.x) {
        const int s = e / kOut, o = e % kOut;
        float y = W[kB4 + o];
#pragma unroll 8
        for (int k = 0; k < kHid; ++k) y += H3[s * kHid + k] * W[kW4 + k * kOut + o];
        const int id = idx[(size_t)blockIdx.x * kMB + s];
        const float t = (T[(size_t)id * kOut + o] - N.y_mean[tp][o]) / N.y_std[tp][o];
        const float d = y - t;
        lsum += d * d;
        DY[s * 8 + o] = 2.0f * d * inv_count;
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
    float* G = gpart + (size_t)blockIdx.x * kNParam;
    // layer 4 gradients
    for (int e = threadIdx.x; e < kHid * kOut; e += blockDim.x) {
        const int k = e / kOut, o = e % kOut;
        float acc = 0.0f;
        for (int s = 0; s < kMB; ++s) acc += H3[s * kHid + k] * DY[s * 8 + o];
        G[kW4 + e] = acc;
    }
    if (threadIdx.x < kOut) {
        float acc = 0.0f;
        for (int s = 0; s < kMB; ++s) acc += DY[s * 8 + threadIdx.x];
        G[kB4 + threadIdx.x] = acc;
    }
    // dZ3 = (dY W4^T) * silu'(Z3), written over Z3
    for (int e = threadIdx.x; e < kMB * kHid; e += blockDim.x) {
        const int s = e / kHid, k = e % kHid;
        float acc = 0.0f;
#pragma unroll
        for (int o = 0; o < kOut; ++o) acc += DY[s * 8 + o] * W[kW4 + k * kOut + o];
        Z3[e] = acc * silu_df(Z3[e]);
    }
    __syncthreads();
    // hidden layers 3 and 2: dW = H^T dZ, db = sum dZ, dZ_prev = (dZ W^T) * silu'(Z_prev)
    float* Zs[3] = {Z1, Z2, Z3};
    float* Hs[3] = {H1, H2, H3};
    const int wofs[3] = {kW1, kW2, kW3}, bofs[3] = {kB1, kB2, kB3};
    for (int l = 2; l >= 1; --l) {
        const float* dZ = Zs[l];
        const float* Hin = Hs[l - 1];
        for (int e = threadIdx.x; e < kHid * kHid; e += blockDim.x) {
            const int k = e / kHid, o = e % kHid;
            float acc = 0.0f;
            for (int s = 0; s < kMB; ++s) acc += Hin[s * kHid + k] * dZ[s * kHid + o];
            G[wofs[l] + e] = acc;
        }
        if (threadIdx.x < kHid) {
            float acc = 0.0f;
            for (int s = 0; s < kMB; ++s) acc += dZ[s * kHid + threadIdx.x];
            G[bofs[l] + threadIdx.x] = acc;
        }
        float* Zp = Zs[l - 1];
        const float* Wl = W + wofs[l];
        // each (s, k) of the previous layer is read then overwritten by one thread
        for (int e = threadIdx.x; e < kMB * kHid; e += blockDim.x) {
            const int s = e / kHid, k = e % kHid;
            float acc = 0.0f;
#pragma unroll 8
            for (int o = 0; o < kHid; ++o) acc += dZ[s * kHid + o] * Wl[k * kHid + o];
            Zp[e] = acc * silu_df(Zp[e]);
        }
        __syncthreads();
    }
    // layer 1
    for (int e = threadIdx.x; e < kIn * kHid; e += blockDim.x) {
        const int k = e / kHid, o = e % kHid;
        float acc = 0.0f;
        for (int s = 0; s < kMB; ++s) acc += H0[s * 4 + k] * Z1[s * kHid + o];
        G[kW1 + e] = acc;
    }
    if (threadIdx.x < kHid) {
        float acc = 0.0f;
        for (int s = 0; s < kMB; ++s) acc += Z1[s * kHid + threadIdx.x];
        G[kB1 + threadIdx.x] = acc;
    }
}

constexpr size_t kSurTrainLds = sizeof(float) * (kNParam + 2 + kMB * 4 + 6 * kMB * kHid + kMB * 8 + 256);

// Adam: gradient of each network = sum of its blocks' gradients in block order
__global__ void k_sur_adam(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ gpart, int ntp, int blocks_per_tp, float lr, float b1, float b2,
                           float eps, float bc1, float bc2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntp * kNParam) return;
    const int tp = e / kNParam, k = e % kNParam;
    float g = 0.0f;
    for (int b = 0; b < blocks_per_tp; ++b) g += gpart[(size_t)(tp * blocks_per_tp + b) * kNParam + k];
    m[e] = b1 * m[e] + (1.0f - b1) * g;
    v[e] = b2 * v[e] + (1.0f - b2) * g * g;
    const float mh = m[e] / bc1, vh = v[e] / bc2;
    params[e] -= lr * mh / (sqrtf(vh) + eps);
}

// sums of block gradients (for the gradient check): G[tp][k]
__global__ void k_sur_sum_grad(const float* __restrict__ gpart, int ntp, int blocks_per_tp, float* __restrict__ G) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ntp * kNParam) return;
    const int tp = e / kNParam, k = e % kNParam;
    float g = 0.0f;
    for (int b = 0; b < blocks_per_tp; ++b) g += gpart[(size_t)(tp * blocks_per_tp + b) * kNParam + k];
    G[e] = g;
}

__device__ __forceinline__ int sur_cell(const SurNorm& N, float r, float ai, float aj) {
    int cr = (int)((r - N.r_lo) / (N.r_hi - N.r_lo) * kCellR);
    int ci = (int)((ai - N.a_lo) / (N.a_hi - N.a_lo) * kCellA);
    int cj = (int)((aj - N.a_lo) / (N.a_hi - N.a_lo) * kCellA);
    cr = cr < 0 ? 0 : (cr >= kCellR ? kCellR - 1 : cr);
    ci = ci < 0 ? 0 : (ci >= kCellA ? kCellA - 1 : ci);
    cj = cj < 0 ? 0 : (cj >= kCellA ? kCellA - 1 : cj);
    return (cr * kCellA + ci) * kCellA + cj;
}

// ---------------------------------------------------------------- BF16 MFMA inference
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// Layers 2-4 chain on v_mfma_f32_16x16x32_bf16 with the weights as the A
// operand (M = output features, K = input features) and 16 samples as the N
// columns.  The D layout (lane l: rows (l/16)*4 + e of the M tile, column
// l%16) is the next layer's B operand (lane l: column l%16, K positions
// (l/16)*8 + [0,8) of each 32-wide k step) once K position (ks, l/16, e) is
// read as input feature kfeat(ks, l/16, e) = (2 ks + e/4)*16 + (l/16)*4 + e%4:
// the activations never leave registers, and the weight fragments carry the
// same permutation.  frag[tp][f][lane][8], f = frag_index(layer, mt, ks):
// lane l holds W[kfeat(ks, l/16, e)][mt*16 + l%16], e = 0..7; layer 2 (64->6)
// has one M tile whose rows >= 6 are zero (layer 0 = W2, 1 = W3, 2 = W4).
__host__ __device__ constexpr int kfeat(int ks, int g, int e) { return (2 * ks + e / 4) * 16 + g * 4 + e % 4; }
constexpr int kFragPerTp = 18;  // layers 0 and 1 (64->64): 4 M tiles x 2 k steps; layer 2 (64->6): 2 k steps
__host__ __device__ constexpr int frag_index(int layer, int mt, int ks) { return layer < 2 ? layer * 8 + mt * 2 + ks : 16 + ks; }

__global__ void k_sur_pack_frags(const float* __restrict__ params, int ntp, __bf16* __restrict__ frag) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (tp, fragment, lane)
    if (idx >= ntp * kFragPerTp * 64) return;
    const int lane = idx % 64, fr = (idx / 64) % kFragPerTp, tp = idx / (64 * kFragPerTp);
    const int layer = fr < 16 ? fr / 8 : 2, mt = fr < 16 ? (fr % 8) / 2 : 0, ks = fr < 16 ? fr % 2 : fr - 16;
    const float* P = params + (size_t)tp * kNParam;
    const int out = mt * 16 + lane % 16;
    for (int e = 0; e < 8; ++e) {
        const int in = kfeat(ks, lane / 16, e);
        float w = 0.0f;
        if (layer == 0) w = P[kW2 + in * kHid + out];
        else if (layer == 1) w = P[kW3 + in * kHid + out];
        else w = out < kOut ? P[kW4 + in * kOut + out] : 0.0f;
        frag[(size_t)idx * 8 + e] = (__bf16)w;
    }
}

struct SurDev {
    const float* params;   // [ntp][kNParam] FP32 (layer 1, biases)
    const __bf16* frag;    // MFMA weight fragments
    const unsigned* mask;  // fallback bit per cell, [ntp][cells/32]
    const float* lenv;     // log2 of the residual envelope, [ntp][kSurOut][kEnvBins]
    SurNorm N;
};

// One block of kSurWaves wavefronts per CU holds every network's MFMA weight
// fragments (72 KB), the layer-1 weights, the biases and the envelope in LDS.
// Per chunk of 16 samples of each type pair, lane tp*16 + k stages the
// standardised input of sample k of type pair tp in the wavefront's LDS slot;
// the wavefront runs the networks one after the other, 16 samples per MFMA N
// tile (lane l: sample l%16, feature group l/16), and lane tp*16 + k reads its
// six outputs back.
constexpr int kSurWaves = 16;
constexpr int kSurBlock = 64 * kSurWaves;
constexpr int kSurBlocks = 256;  // one block per CU
struct SurLds {
    __bf16 frag[4][kFragPerTp][64][8];  // MFMA A operands
    float4 w1[4][kHid];                 // layer 1: {w_r, w_ai, w_aj, bias} per output
    float bias[4][3][kHid];             // biases of layers 2, 3 and 4
    float lenv[4][kSurOut][kEnvBins];   // log2 of the residual envelope
    float4 u[kSurWaves][64];            // standardised inputs, lane tp*16 + k
    float y[kSurWaves][64][8];          // standardised outputs, lane tp*16 + k
};
constexpr size_t kSurLds = sizeof(SurLds);

__device__ __forceinline__ void sur_load_lds(SurLds& sh, const SurDev& S) {
    const uint4* src = reinterpret_cast<const uint4*>(S.frag);
    uint4* dst = reinterpret_cast<uint4*>(&sh.frag[0][0][0][0]);
    for (int e = threadIdx.x; e < 4 * kFragPerTp * 64; e += kSurBlock) dst[e] = src[e];
    for (int e = threadIdx.x; e < 4 * kHid; e += kSurBlock) {
        const int m = e / kHid, o = e % kHid;
        const float* Pm = S.params + (size_t)m * kNParam;
        sh.w1[m][o] = make_float4(kKappa * Pm[kW1 + o], kKappa * Pm[kW1 + kHid + o], kKappa * Pm[kW1 + 2 * kHid + o],
                                  kKappa * Pm[kB1 + o]);
        sh.bias[m][0][o] = kKappa * Pm[kB2 + o];
        sh.bias[m][1][o] = kKappa * Pm[kB3 + o];
        sh.bias[m][2][o] = o < kOut ? Pm[kB4 + o] : 0.0f;
    }
    for (int e = threadIdx.x; e < 4 * kSurOut * kEnvBins; e += kSurBlock) (&sh.lenv[0][0][0])[e] = S.lenv[e];
    __syncthreads();
}

// the networks of the type pairs whose bit is set in need (wave-uniform) on the
// staged inputs sh.u[w][tp*16 + n] -> standardised outputs sh.y[w][tp*16 + n][0..5]
__device__ __forceinline__ void sur_nets(SurLds& sh, int w, int lane, unsigned need) {
    const int g = lane / 16, n = lane % 16;  // feature group, sample column
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
        if (!((need >> t) & 1u)) continue;
        const float4 uu = sh.u[w][t * 16 + n];
        bf16x8_t b[2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float4 wt = sh.w1[t][kfeat(ks, g, e)];
                b[ks][e] = (__bf16)silu_scaled(wt.w + uu.x * wt.x + uu.y * wt.y + uu.z * wt.z);
            }
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            floatx4_t d[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                // the accumulator starts from the (scaled) biases of the lane's rows
                floatx4_t c = *reinterpret_cast<const floatx4_t*>(&sh.bias[t][layer][mt * 16 + g * 4]);
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const bf16x8_t a =
                        *reinterpret_cast<const bf16x8_t*>(&sh.frag[t][frag_index(layer, mt, ks)][lane][0]);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[ks], c, 0, 0, 0);
                }
                d[mt] = c;
            }
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int e = 0; e < 4; ++e) b[mt / 2][(mt % 2) * 4 + e] = (__bf16)silu_scaled(d[mt][e]);
        }
        floatx4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&sh.frag[t][frag_index(2, 0, ks)][lane][0]);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[ks], c, 0, 0, 0);
        }
        // rows g*4 + e are outputs 0..3 (g = 0) and 4, 5 (g = 1)
        if (g < 2) {
            const float4 bs = *reinterpret_cast<const float4*>(&sh.bias[t][2][g * 4]);
            float* y = &sh.y[w][t * 16 + n][g * 4];
            y[0] = c[0] * kInvKappa + bs.x;
            y[1] = c[1] * kInvKappa + bs.y;
            if (g == 0) {
                y[2] = c[2] * kInvKappa + bs.z;
                y[3] = c[3] * kInvKappa + bs.w;
            }
        }
    }
}

// which of the four 16-lane groups hold a lane with v set (wave-uniform)
__device__ __forceinline__ unsigned group_any(bool v) {
    const unsigned long long bal = __ballot(v);
    unsigned need = 0u;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if ((bal >> (16 * t)) & 0xFFFFull) need |= 1u << t;
    return need;
}

// lane's prediction from the networks' standardised outputs y and the baseline B:
// out = B + s(r) * (y * std + mean)
__device__ __forceinline__ void sur_predict(const SurLds& sh, const SurNorm& N, int tp, float r, const float* y,
                                            const float* B, float* out) {
#pragma unroll
    for (int c = 0; c < kSurOut; ++c) {
        const float sc = __builtin_amdgcn_exp2f(sur_env_log2(&sh.lenv[0][0][0], N, tp, c, r));
        out[c] = B[c] + sc * (y[c] * N.y_std[tp][c] + N.y_mean[tp][c]);
    }
}

// The inference path on validation samples: err[(tp * n_val + s) * 6 + c] =
// |prediction - quadrature| of output c.  Persistent wavefronts over chunks of
// 16 samples of every type pair.
__global__ __launch_bounds__(kSurBlock) void k_sur_judge(FFParams<float> PQ, SurDev S, const float* __restrict__ quad,
                                                         const float* __restrict__ X, const float* __restrict__ T,
                                                         long n_val, float* __restrict__ err) {
    constexpr int NT = 2;
    extern __shared__ __attribute__((aligned(16))) unsigned char sur_lds_raw[];
    SurLds& sh = *reinterpret_cast<SurLds*>(sur_lds_raw);
    sur_load_lds(sh, S);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tp = lane / 16, k = lane % 16;
    const long nchunk = (n_val + 15) / 16;
    for (long ch = (long)blockIdx.x * kSurWaves + w; ch < nchunk; ch += (long)gridDim.x * kSurWaves) {
        const long s = ch * 16 + k;
        const bool valid = s < n_val;
        const long id = (long)tp * n_val + (valid ? s : 0);
        const float r = X[id * 3], ai = X[id * 3 + 1], aj = X[id * 3 + 2];
        const float ia = 1.0f / ai + 1.0f / aj;
        float B[6];
        sur_baseline(PQ, quad + (size_t)tp * kQuadRow, quad_stride(NT), r, ia, B);
        sh.u[w][lane] = make_float4((r - S.N.r_c) / S.N.r_s, (ai - S.N.a_c) / S.N.a_s, (aj - S.N.a_c) / S.N.a_s, 0.0f);
        __builtin_amdgcn_wave_barrier();
        sur_nets(sh, w, lane, 0xFu);
        __builtin_amdgcn_wave_barrier();
        float out[6];
        sur_predict(sh, S.N, tp, r, sh.y[w][lane], B, out);
        __builtin_amdgcn_wave_barrier();
        if (valid)
            for (int c = 0; c < kSurOut; ++c) err[id * kSurOut + c] = fabsf(out[c] - T[id * kSurOut + c]);
    }
}

// pass 1 with the surrogate (DMD force, NT == 2).  The block's wavefronts are
// persistent over the sites of their XCD range and never synchronise after the
// set-up.  Per chunk of 16 pairs, lane tp*16 + k prepares sample (pair k, type
// pair tp) (partner record, distance, trained box, fallback cell, baseline row);
// the networks of the type pairs that serve at least one sample of the chunk
// run; samples outside the trained box or in a masked cell take the FP32
// quadrature instead.
template <int NQ>
__global__ __launch_bounds__(kSurBlock) void k_pass1_sur(FFParams<double> P, FFParams<float> PQ, SurDev S,
                                                         const double4* __restrict__ pos,
                                                         const double4* __restrict__ arec, PairBufs L,
                                                         const int2* nlrp, const float* __restrict__ quad,
                                                         const double* __restrict__ rho7,
                                                         const double* __restrict__ phi7, double* __restrict__ Eown,
                                                         double* __restrict__ pairE, double* __restrict__ deriv,
                                                         double* __restrict__ part, int* __restrict__ nfall,
                                                         CutState C) {
    constexpr int NT = 2;
    extern __shared__ __attribute__((aligned(16))) unsigned char sur_lds_raw[];
    SurLds& sh = *reinterpret_cast<SurLds*>(sur_lds_raw);
    sur_load_lds(sh, S);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tp = lane / 16, k = lane % 16, it = tp / NT, jt = tp % NT;
    const bool want_phi = it >= jt;
    const float* tab = quad + (size_t)tp * kQuadRow;
    double en = 0.0, npair = 0.0;
    const SiteLoop sl(P.natms);
    for (int i0 = sl.first; i0 < sl.end; i0 += sl.stride) {
        const int i = __builtin_amdgcn_readfirstlane(i0);
        int fallbacks = 0;
        const double4 xi = pos[i];
        double ci[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) ci[t] = arec[(size_t)i * NT + t].y;
        const double4 ra = arec[(size_t)i * NT + it];
        const double ai = ra.x, iai = ra.z;
        double Ei = 0.0;
        const int p0 = L.off[i], p1 = L.off[i + 1];
        int done = 0;  // inside entries of the earlier chunks (records at off[i] + rank)
        for (int chunk = p0; chunk < p1; chunk += 64) {
            int packed;
            int cnt;
            {
                const int pc = chunk + lane;
                bool in = false;
                if (pc < p1) {
                    const int2 e = nlrp[pc];
                    const int jc = list_partner(e);
                    const double4 xj = pos[jc];
                    const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
                    double r2;
                    {
#pragma clang fp contract(off)
                        r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
                    }
                    in = r2 < P.cut_sq_mod;
                    C.update<NT>(pc, e, in, jc, P.natms, pairE);
                }
                cnt = pack_in_range(pc, in, &packed);
            }
        for (int base = 0; base < cnt; base += 16) {
            const int slot = base + k;
            const bool valid = slot < cnt;
            const int p = __shfl(packed, valid ? slot : 0, 64);
            int j = 0;
            double rsq = 0.0;
            double aj = 1.0, cj[NT] = {0.0, 0.0}, r = 1.0, r_inv = 1.0, ia = 1.0;
            int br = 0;
            if (valid) {  // inside the cut by construction
                j = list_partner(nlrp[p]);
                const double4 xj = pos[j];
                const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
                {
#pragma clang fp contract(off)
                    rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
                }
                const double4 rj = arec[(size_t)j * NT + jt];
#pragma unroll
                for (int t = 0; t < NT; ++t) cj[t] = t == jt ? rj.y : arec[(size_t)j * NT + t].y;
                aj = rj.x;
                r_inv = rsqrt(rsq);
                r = rsq * r_inv;
                ia = iai + rj.z;
                br = alpha_branch(P, ia, ai, aj);
            }
            const float rf = (float)r, aif = (float)ai, ajf = (float)aj;
            bool use_net = valid && br == 1 && rf >= S.N.r_lo && rf <= S.N.r_hi && aif >= S.N.a_lo &&
                           aif <= S.N.a_hi && ajf >= S.N.a_lo && ajf <= S.N.a_hi;
            if (use_net) {
                const int cell = sur_cell(S.N, rf, aif, ajf);
                const int per_tp = kCellR * kCellA * kCellA / 32;
                if ((S.mask[tp * per_tp + cell / 32] >> (cell % 32)) & 1u) use_net = false;
            }
            const unsigned need = group_any(use_net);
            float out[6];
            if (need) {
                float B[6];
                sur_baseline(PQ, tab, quad_stride(NT), rf, (float)ia, B);
                sh.u[w][lane] = make_float4((rf - S.N.r_c) / S.N.r_s, (aif - S.N.a_c) / S.N.a_s,
                                            (ajf - S.N.a_c) / S.N.a_s, 0.0f);
                __builtin_amdgcn_wave_barrier();
                sur_nets(sh, w, lane, need);
                __builtin_amdgcn_wave_barrier();
                sur_predict(sh, S.N, tp, rf, sh.y[w][lane], B, out);
                __builtin_amdgcn_wave_barrier();
            }
            double gv[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
            if (use_net) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    gv[c] = out[c];
                    fv[c] = out[3 + c];
                }
            } else if (valid && br == 1) {
                ++fallbacks;
                float gq[3] = {0, 0, 0}, fq[3] = {0, 0, 0};
                gh_avg<float, NQ, true>(PQ, rf, (float)r_inv, (float)rsqrt(ia), (float)ia, tab, quad_stride(NT), gq,
                                        fq);
                for (int c = 0; c < 3; ++c) {
                    gv[c] = gq[c];
                    fv[c] = fq[c];
                }
            } else if (valid && br == 2 && rsq < P.cut_sq) {
                classical_pair(P, rsq, r_inv, rho7 + (size_t)type2rho(NT, it, jt) * P.nr * 7,
                               phi7 + (size_t)type2phi(it, jt) * P.nr * 7, true, gv, fv);
            }
            double ej = 0.0;
            if (valid) {
                if (tp == 0) npair += 1.0;
                Ei += gv[0] * cj[jt];
                ej = gv[0] * ci[it];
                if (want_phi) {
                    const double half = (j < P.natms ? 0.5 : 0.25) * (it != jt ? P.xw : 1.0);
                    en += half * (cj[jt] * ci[it] + cj[it] * ci[jt]) * fv[0];
                }
                double* d = deriv + (size_t)(p0 + done + slot) * DerivRec<NT>::kLen;
                __builtin_nontemporal_store(gv[1], d + DerivRec<NT>::rho(tp));
                __builtin_nontemporal_store(gv[2], d + DerivRec<NT>::rho(tp) + 1);
                if (want_phi) {
                    __builtin_nontemporal_store(fv[1], d + DerivRec<NT>::phi(it, jt));
                    __builtin_nontemporal_store(fv[2], d + DerivRec<NT>::phi(it, jt) + 1);
                }
            }
            ej += __shfl_xor(ej, 32, 64);  // sum over it: type pairs (0,jt) and (1,jt) are 32 lanes apart
            if (valid && it == 0 && j < P.natms) pairE[(size_t)nlrp[p].y * NT + jt] = ej;  // scattered: L2 merges lines
        }
        done += cnt;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double s = wave_sum(it == t ? Ei : 0.0);
            if (lane == 0) Eown[(size_t)i * NT + t] = s;
        }
        const double sf = wave_sum((double)fallbacks);
        if (lane == 0) nfall[i] = (int)sf;
    }
    double acc[2] = {wave_sum(en), wave_sum(npair)};
    block_store_waves<2, kSurWaves>(acc, part);
}

}  // namespace mappdmd

#pragma clang fp contract(off)

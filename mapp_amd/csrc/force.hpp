// Force kernels of the DMD force field (ForceField_DMD_I::force_calc,
// src/ff_dmd_I.cpp:88-450, and energy_calc, :454-668).
//
// The reference scatters every half-list pair into both ends (E[j] +=,
// f[j] -=).  Here one wavefront owns one site and never writes another
// site's row; what a pair owes its partner goes to a per-pair share slot at
// the pair's position in the partner's reverse list:
//   pass 1   the site's list entries inside the quadrature cut are staged in
//            LDS; lanes (pair, type pair, part) evaluate the Gauss-Hermite
//            averaged density and pair terms, the site's own density sums
//            stay in registers, the partner's density share (nt doubles)
//            goes to its share slot, and the r- and alpha-derivatives go to
//            the pair's derivative record (DerivRec);
//   embed    a lane group per site adds the density shares of the sites that
//            list it (reverse CSR, ascending pair index), then F(rho_bar), F',
//            the per-site energy terms and the alpha entropic force;
//   pass 2   lanes (pair, type pair) combine the derivative records with F'
//            of both ends into the site's force row and the partner's force
//            share (3 + nt doubles);
//   gather   a lane group per site adds its force shares (reverse CSR).
// No floating-point atomics: every sum has a fixed order, so results are
// bitwise reproducible run to run.  Energy and virial partials are written per
// wavefront and reduced in two fixed-shape stages.
//
// Q is the arithmetic of the quadrature (double, or float for the mixed
// mode whose point sums run in FP32 with FP32 accumulation); site-level sums,
// F(rho_bar), energies and forces are FP64 in both.
#pragma once
#include "dmd_common.hpp"

// The force kernels do not need the reference's bitwise operation order (the
// bar is 1e-12 relative), so a*b+c may fuse into FMA here; the geometry
// kernels, which must reproduce the reference's list bit for bit, are
// compiled before this region with contraction off.
#pragma clang fp contract(fast)

namespace mappdmd {

constexpr int kBlock = 256;
constexpr int kQuadRow = 8;  // rho c0..c3, r*phi c0..c3 of one ordered type pair
// The quadrature table is bin-major, quad[m][tp][kQuadRow]: the rows of all
// type pairs at one grid row are adjacent, so the type pairs of a pair that
// evaluate at the same distance (the middle node of an odd rule, t = r) share
// cache lines.  Type pair tp starts at quad + tp * kQuadRow; consecutive rows
// of one type pair are quad_stride(nt) elements apart.
__host__ __device__ constexpr int quad_stride(int nt) { return nt * nt * kQuadRow; }

__host__ __device__ constexpr int type2rho(int nt, int i, int j) { return j * nt + i; }
__host__ __device__ constexpr int type2phi(int i, int j) {
    return i < j ? (j + 1) * j / 2 + i : (i + 1) * i / 2 + j;
}

template <typename T>
__device__ __forceinline__ T horner3(const T* c, T p) {
    return fma(fma(fma(c[3], p, c[2]), p, c[1]), p, c[0]);
}

// Gauss-Hermite average of rho (odd extension u*rho(|u|)) and r*phi (odd
// extension sgn(u) r*phi(|u|)) of one ordered type pair; src/ff_dmd_I.cpp:180-242.
// g, f receive {value, d/dr, d/dalpha} (derivatives when DERIV).
// Returns false when the whole rule lies beyond the cutoff (lower >= xi_max).
template <typename Q, int NQ, bool DERIV, bool PHI>
__device__ __forceinline__ bool gh_average(const FFParams<Q>& P, Q r, Q r_inv, Q alpha,
                                           const Q* __restrict__ tab, int stride, Q* g, Q* f) {
    const int nq = NQ > 0 ? NQ : P.no_i;
    const Q asq = sqrt(alpha);
    const Q inv = Q(1) / asq;
    const Q upper = (r + P.rc) * asq;
    const Q lower = (r - P.rc) * asq;
    if (!(lower < P.xi[nq - 1])) return false;
    Q g0 = 0, g1 = 0, g2 = 0, f0 = 0, f1 = 0, f2 = 0;
#pragma unroll
    for (int k = 0; k < (NQ > 0 ? NQ : kMaxPoints); ++k) {
        if (NQ == 0 && k >= nq) break;
        const Q xk = P.xi[k];
        if (xk > lower && xk < upper) {
            const Q t = fma(-xk, inv, r);
            Q p = fabs(t) * P.dr_inv;
            int m = (int)p;
            m = m < P.nr - 2 ? m : P.nr - 2;
            p -= (Q)m;
            p = p < Q(1) ? p : Q(1);
            const Q* row = tab + (size_t)m * stride;
            const Q rv = horner3(row, p) * t;
            g0 = fma(P.w0[k], rv, g0);
            if (DERIV) {
                g1 = fma(P.w1[k], rv, g1);
                g2 = fma(P.w2[k], rv, g2);
            }
            if (PHI) {
                Q pv = horner3(row + 4, p);
                pv = t < Q(0) ? -pv : pv;
                f0 = fma(P.w0[k], pv, f0);
                if (DERIV) {
                    f1 = fma(P.w1[k], pv, f1);
                    f2 = fma(P.w2[k], pv, f2);
                }
            }
        }
    }
    const Q sc = Q(kPiInSq) * r_inv;
    g0 *= sc;
    g[0] = g0;
    if (DERIV) {
        g1 *= sc;
        g2 *= sc;
        g[1] = -g0 * r_inv - Q(2) * asq * g1;
        g[2] = (Q(0.5) * g0 - g2) / alpha;
    }
    if (PHI) {
        f0 *= sc;
        f[0] = f0;
        if (DERIV) {
            f1 *= sc;
            f2 *= sc;
            f[1] = -f0 * r_inv - Q(2) * asq * f1;
            f[2] = (Q(0.5) * f0 - f2) / alpha;
        }
    }
    return true;
}

// classical EAM values for alpha >= alpha_max (src/ff_dmd_I.cpp:261-310), FP64
__device__ __forceinline__ void classical_pair(const FFParams<double>& P, double rsq, double r_inv,
                                               const double* __restrict__ rho7, const double* __restrict__ phi7,
                                               bool want_phi, double* g, double* f) {
    double r = sqrt(rsq);
    double p = r * P.dr_inv;
    int m = (int)p;
    m = m < P.nr - 2 ? m : P.nr - 2;
    p -= m;
    p = p < 1.0 ? p : 1.0;
    double p2 = p * p, p3 = p2 * p;
    if (want_phi) {
        const double* c = phi7 + (size_t)m * 7;
        double z2p = c[6] * p2 + c[5] * p + c[4];
        double z2 = c[3] * p3 + c[2] * p2 + c[1] * p + c[0];
        f[0] = z2 * r_inv;
        f[2] = 0.0;
        f[1] = (z2p - z2 * r_inv) * r_inv;
    }
    const double* c = rho7 + (size_t)m * 7;
    g[0] = c[3] * p3 + c[2] * p2 + c[1] * p + c[0];
    g[2] = 0.0;
    g[1] = c[6] * p2 + c[5] * p + c[4];
}

__device__ __forceinline__ double calc_ent(double x) {
    if (x < 1.0e-8 || 0.99999999 < x) return 0.0;
    return x * log(x) + (1 - x) * log(1 - x);
}

// F(rho_bar), F', F'' by the quartic spline with linear extrapolation above
// rho_max (src/ff_dmd_I.cpp:322-341, :2484-2491)
__device__ __forceinline__ void embed_F(const FFParams<double>& P, const double* __restrict__ F9, int type,
                                        double e, double* Fv, double* dF, double* ddF) {
    double p = e * P.drho_inv;
    int m = (int)p;
    m = m < P.nr - 2 ? m : P.nr - 2;
    m = m < 0 ? 0 : m;  // negative density: outside the reference's defined range
    p -= m;
    p = p < 1.0 ? p : 1.0;
    double p2 = p * p, p3 = p2 * p, p4 = p3 * p;
    const double* c = F9 + ((size_t)type * (P.nrho + 1) + m) * 9;
    double t0 = c[4] * p4 + c[3] * p3 + c[2] * p2 + c[1] * p + c[0];
    double t1 = c[8] * p3 + c[7] * p2 + c[6] * p + c[5];
    if (e > P.rho_max) t0 += t1 * (e - P.rho_max);
    *Fv = t0;
    *dF = t1;
    if (ddF) *ddF = (3.0 * c[8] * p2 + c[7] * p + c[6]) * P.drho_inv;
}

// deterministic block sum of NV doubles (fixed tree), thread 0 writes out[blockIdx.x*NV + k]
template <int NV>
__device__ __forceinline__ void block_sum_store(double* v, double* __restrict__ out) {
    __shared__ double sh[NV][kBlock];
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][tid] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int k = 0; k < NV; ++k) sh[k][tid] += sh[k][tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[(size_t)blockIdx.x * NV + k] = sh[k][0];
    }
}

// XCD-aware block order: consecutive atoms (spatial neighbours) land on the
// same XCD, whose L2 then holds their shared neighbour rows.  Dispatch places
// block b on XCD b % 8.
__device__ __forceinline__ int xcd_block(int nblocks) {
    const int b = blockIdx.x;
    const int per = (nblocks + 7) / 8;
    const int xcd = b & 7, k = b >> 3;
    const int full = nblocks - 8 * (per - 1);  // XCDs that receive 'per' blocks
    int start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return start + k;
}

struct PairBufs {
    const int* off;
    const int* nl;
    const int* roff;
    const int* rev;
};

// Partner force shares of the DMD passes, one slot per reverse-list entry, in
// two arrays: sh4[slot] = {fx, fy, fz, alpha force of type 0} (32 bytes, so a
// pair's share is written with 16-byte stores) and shx[slot] = the alpha
// forces of types 1..nt-1.
struct ForceShares {
    double* sh4;
    double* shx;
};

// Which half-list entries lay inside the quadrature cut at the last pass 1:
// the top bit of the partner index in the packed list {partner, reverse
// position}, read with the entry at no extra cost.  The partner shares (pairE
// of pass 1, the force shares of pass 2) of an entry outside the cut must read
// zero in the embedding and the force gather: they are cleared when the list
// is built and again only when an entry leaves the cut, instead of being
// rewritten with zeros at every evaluation.
constexpr int kInsideBit = (int)0x80000000u;
__device__ __forceinline__ int list_partner(int2 e) { return e.x & 0x7fffffff; }
struct CutState {
    int2* nlrp;
    ForceShares F;
    // entry pc (as read: e, partner jc) is inside the cut now: record a change; on
    // leaving, clear both shares of an owned partner
    template <int NT>
    __device__ __forceinline__ void update(int pc, int2 e, bool in, int jc, int natms, double* pairE) const {
        const bool was = (e.x & kInsideBit) != 0;
        if (was != in) {
            nlrp[pc].x = in ? (jc | kInsideBit) : jc;
            if (!in && jc < natms) {
                const int rp = e.y;
#pragma unroll
                for (int t = 0; t < NT; ++t) pairE[(size_t)rp * NT + t] = 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c) F.sh4[(size_t)rp * 4 + c] = 0.0;
#pragma unroll
                for (int c = 0; c < NT - 1; ++c) F.shx[(size_t)rp * (NT - 1) + c] = 0.0;
            }
        }
    }
};

// ---------------------------------------------------------------- lane mapping
// NT*NT lanes per site, lane q = it*NT + jt owns the ordered type pair
// (it, jt) of every pair of the site's list.  For NT = 2 the four lanes of a
// site form a DPP quad: the transposed pair (jt, it) is quad_perm [0,2,1,3],
// sums over jt are xor 1, sums over it are xor 2.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
constexpr int kDppTranspose = 0xD8;  // quad_perm [0,2,1,3]
constexpr int kDppXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kDppRowRor8 = 0x128;   // row_ror:8 (rotate by 8 lanes within a row of 16)
constexpr int kDppRowRor4 = 0x124;   // row_ror:4
constexpr int kDppRowRor12 = 0x12C;  // row_ror:12: lane l reads lane (l + 4) mod 16 of its row
constexpr int kDppQuad0022 = 0xA0;   // quad_perm [0,0,2,2]: lane q = 1 reads lane q = 0
constexpr int kDppQuad0000 = 0x00;   // quad_perm [0,0,0,0]: every lane of a quad reads q = 0
constexpr int kDppQuad0033 = 0xF0;   // quad_perm [0,0,3,3]: lane 1 reads lane 0, lane 2 reads lane 3

// v summed with the same lane of the paired row (rows 0+1, 2+3: permlane16
// swap) or of the other half of the wavefront (permlane32 swap); both lanes of
// a pair add in the same order, so they hold the same bits
template <bool HALVES>
__device__ __forceinline__ double row_pair_sum(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const auto a = HALVES ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                          : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = HALVES ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                          : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double(b[0], a[0]) + __hiloint2double(b[1], a[1]);
}

// Sums over the 16 quads of a wavefront, per position q in the quad, of two
// registers at once: afterwards lane q of the wavefront holds, in A and B, the
// totals of A and B over the lanes 4k + q (fixed tree: two rotations within each
// row of 16, then the rows in pairs and the halves; the result is read from lane q).
__device__ __forceinline__ void quad_slot_sums(double& A, double& B) {
    A += dpp_f64<kDppRowRor4>(A);
    B += dpp_f64<kDppRowRor4>(B);
    A += dpp_f64<kDppRowRor8>(A);
    B += dpp_f64<kDppRowRor8>(B);
    A = row_pair_sum<false>(A);
    B = row_pair_sum<false>(B);
    A = row_pair_sum<true>(A);
    B = row_pair_sum<true>(B);
}

template <int NT>
__device__ __forceinline__ double transpose_tp(double v) {
    if constexpr (NT == 2) {
        return dpp_f64<kDppTranspose>(v);
    } else if constexpr (NT == 1) {
        return v;
    } else {  // lane (it, jt) of the pair's NT*NT lanes reads lane (jt, it)
        const int lane = threadIdx.x & 63, q = lane % (NT * NT);
        return __shfl(v, lane - q + (q % NT) * NT + q / NT, 64);
    }
}
template <int NT>
__device__ __forceinline__ double sum_over_jt(double v) {
    if constexpr (NT == 2) {
        return v + dpp_f64<kDppXor1>(v);
    } else if constexpr (NT == 1) {
        return v;
    } else {  // the NT lanes (it, 0..NT-1) in a fixed order
        const int lane = threadIdx.x & 63, q = lane % (NT * NT), row = lane - q + (q / NT) * NT;
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < NT; ++t) s += __shfl(v, row + t, 64);
        return s;
    }
}
template <int NT>
__device__ __forceinline__ double sum_over_it(double v) {
    if constexpr (NT == 2) {
        return v + dpp_f64<kDppXor2>(v);
    } else if constexpr (NT == 1) {
        return v;
    } else {  // the NT lanes (0..NT-1, jt) in a fixed order
        const int lane = threadIdx.x & 63, q = lane % (NT * NT), col = lane - q + q % NT;
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < NT; ++t) s += __shfl(v, col + t * NT, 64);
        return s;
    }
}

// which branch of the reference applies to the pair width alpha = 1/ia:
// 1 quadrature (alpha_min < alpha < alpha_max), 2 classical (alpha >= alpha_max),
// 0 nothing.  ia = 1/alpha_i + 1/alpha_j equals the reference's
// alpha_i alpha_j / (alpha_i + alpha_j) to within two ulp; within 1e-12 of a
// bound the reference's own expression decides.
__device__ __forceinline__ int alpha_branch(const FFParams<double>& P, double ia, double ai, double aj) {
    const double lo = P.inv_amax, hi = P.inv_amin;
    if (fabs(ia - hi) <= 1e-12 * hi || fabs(ia - lo) <= 1e-12 * lo) {
        const double a = (ai * aj) / (ai + aj);
        return (P.alpha_min < a && a < P.alpha_max) ? 1 : (P.alpha_max <= a ? 2 : 0);
    }
    return (ia > lo && ia < hi) ? 1 : (ia <= lo ? 2 : 0);
}

// Gauss-Hermite average with sqrt(alpha) = asq and 1/alpha = ia given
// (src/ff_dmd_I.cpp:180-242); no divisions and no branches: a point outside
// the window (xi_k <= lower or xi_k >= upper) or a rule beyond the cutoff
// (lower >= xi_max) contributes zero, its row index being clamped in range.
// g, f = {value, d/dr, d/dalpha}; returns whether the rule is inside.
template <typename Q, int NQ, bool DERIV>
__device__ __forceinline__ bool gh_avg(const FFParams<Q>& P, Q r, Q r_inv, Q asq, Q ia, const Q* __restrict__ tab,
                                       int stride, Q* g, Q* f) {
    const int nq = NQ > 0 ? NQ : P.no_i;
    const Q inv = ia * asq;
    const Q upper = (r + P.rc) * asq;
    const Q lower = (r - P.rc) * asq;
    const bool rule = lower < P.xi[nq - 1];
    Q g0 = 0, g1 = 0, g2 = 0, f0 = 0, f1 = 0, f2 = 0;
#pragma unroll
    for (int k = 0; k < (NQ > 0 ? NQ : kMaxPoints); ++k) {
        if (NQ == 0 && k >= nq) break;
        const Q xk = P.xi[k];
        const bool inside = rule && xk > lower && xk < upper;
        const Q t = r - xk * inv;
        Q p = fabs(t) * P.dr_inv;
        int m = (int)p;
        m = m < P.nr - 2 ? m : P.nr - 2;
        m = m > 0 ? m : 0;
        p -= (Q)m;
        p = p < Q(1) ? p : Q(1);
        const Q* row = tab + (size_t)m * stride;
        Q rv = horner3(row, p) * t;
        Q pv = horner3(row + 4, p);
        pv = t < Q(0) ? -pv : pv;
        rv = inside ? rv : Q(0);
        pv = inside ? pv : Q(0);
        g0 += P.w0[k] * rv;
        f0 += P.w0[k] * pv;
        if (DERIV) {
            g1 += P.w1[k] * rv;
            g2 += P.w2[k] * rv;
            f1 += P.w1[k] * pv;
            f2 += P.w2[k] * pv;
        }
    }
    const Q sc = Q(kPiInSq) * r_inv;
    g0 *= sc;
    f0 *= sc;
    g[0] = g0;
    f[0] = f0;
    if (DERIV) {
        g1 *= sc;
        g2 *= sc;
        f1 *= sc;
        f2 *= sc;
        g[1] = -g0 * r_inv - Q(2) * asq * g1;
        g[2] = (Q(0.5) * g0 - g2) * ia;
        f[1] = -f0 * r_inv - Q(2) * asq * f1;
        f[2] = (Q(0.5) * f0 - f2) * ia;
    }
    return rule;
}

// 1/alpha per (site, type), ghosts included
__global__ void k_inv_alpha(const double* __restrict__ alpha, long n, double* __restrict__ inva) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) inva[k] = 1.0 / alpha[k];
}

// pair values (and r-, alpha-derivatives when DERIV) of the lane's type pair
template <int NT, typename Q, int NQ, bool DERIV>
__device__ __forceinline__ void lane_pair(const FFParams<double>& P, const FFParams<Q>& PQ, int it, int jt,
                                          double rsq, double r, double r_inv, double ai, double aj, double ia,
                                          double asq, const Q* __restrict__ quad, const double* __restrict__ rho7,
                                          const double* __restrict__ phi7, double* g, double* f) {
    const int br = alpha_branch(P, ia, ai, aj);
    if (br == 1) {
        Q gq[3], fq[3];
        const Q* tab = quad + (size_t)(it * NT + jt) * kQuadRow;
        if (gh_avg<Q, NQ, DERIV>(PQ, (Q)r, (Q)r_inv, (Q)asq, (Q)ia, tab, quad_stride(NT), gq, fq)) {
            g[0] = (double)gq[0];
            f[0] = (double)fq[0];
            if (DERIV) {
                g[1] = (double)gq[1];
                g[2] = (double)gq[2];
                f[1] = (double)fq[1];
                f[2] = (double)fq[2];
            }
        }
    } else if (br == 2 && rsq < P.cut_sq) {
        classical_pair(P, rsq, r_inv, rho7 + (size_t)type2rho(NT, it, jt) * P.nr * 7,
                       phi7 + (size_t)type2phi(it, jt) * P.nr * 7, true, g, f);
    }
}

// ---------------------------------------------------------------- wave per site
// One wavefront per site i: lane l = k*LPP + q works on pair k of the current
// chunk of i's half list (PPW = 64/LPP pairs per chunk) and on ordered type
// pair q = it*NT + jt.  Streams indexed by the pair (list, rpos, derivative
// records) are then contiguous across the wave, the site's own data is
// wave-uniform, and each neighbour's attributes are one 32-byte record per type.
template <int NT>
struct WaveMap {
    static constexpr int LPP = NT * NT;
    static constexpr int PPW = 64 / LPP;
};

// per (site, type) record {alpha, c, 1/alpha, dE}, ghosts included
__global__ void k_pack_atoms(const double* __restrict__ alpha, const double* __restrict__ conc, long n,
                             double4* __restrict__ rec) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const double a = alpha[k];
        rec[k] = make_double4(a, conc[k], 1.0 / a, 0.0);
    }
}

// per half-list entry {partner, reverse position}: the force passes read both
// with one load (every entry starts outside the quadrature cut)
__global__ void k_pack_nlrp(const int* __restrict__ nl, const int* __restrict__ rpos, long n,
                            int2* __restrict__ nlrp) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) nlrp[k] = make_int2(nl[k], rpos[k]);
}

// sum over the wavefront, the same value in every lane: a DPP rotation tree
// inside each row of 16 lanes, then the four row sums in a fixed order
__device__ __forceinline__ double read_lane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<kDppRowRor8>(v);
    v += dpp_f64<kDppRowRor4>(v);
    v += dpp_f64<kDppXor2>(v);
    v += dpp_f64<kDppXor1>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// deterministic per-block sum of NV wave partials (lane 0 of each wave holds them)
template <int NV, int MAXW = kBlock / 64>
__device__ __forceinline__ void block_store_waves(const double* v, double* __restrict__ out) {
    __shared__ double sh[MAXW][NV];  // blockDim.x <= 64 * MAXW
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[w][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int q = 0; q < nw; ++q) s += sh[q][k];
            out[(size_t)blockIdx.x * NV + k] = s;
        }
    }
}

// Per-wavefront partials without a block barrier: lane 0 of wavefront w of
// block b stores the wave's NV totals at out[(b * waves per block + w) * NV].
// A wave that finishes early leaves at once instead of holding its slot until
// the block's slowest wave reaches a barrier.
template <int NV>
__device__ __forceinline__ void wave_store(const double* v, double* __restrict__ out) {
    if ((threadIdx.x & 63) == 0) {
        const size_t w = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
#pragma unroll
        for (int k = 0; k < NV; ++k) out[w * NV + k] = v[k];
    }
}

// Sites of one wavefront.  Dispatch places block b on XCD b % 8; the sites are
// split into 8 contiguous ranges, one per XCD, so spatial neighbours share an
// L2, and the waves of an XCD stride through its range.  With one wave per
// site (the default grid) each wave takes one site; a smaller grid makes the
// launch persistent.  Every wave of a block runs the same number of
// iterations or fewer; kernels that synchronise a block inside the loop must
// not use it.
struct SiteLoop {
    int first, end, stride;
    __device__ __forceinline__ explicit SiteLoop(int n) {
        const int wpb = blockDim.x >> 6, w = threadIdx.x >> 6;
        const int nb = gridDim.x, b = blockIdx.x;
        if (nb >= 8) {
            const int x = b & 7, lb = b >> 3;
            const int nbx = (nb - x + 7) >> 3;
            first = (int)((long long)n * x / 8) + lb * wpb + w;
            end = (int)((long long)n * (x + 1) / 8);
            stride = nbx * wpb;
        } else {
            first = b * wpb + w;
            end = n;
            stride = nb * wpb;
        }
    }
};

// Pairs inside the cut, packed.  Each lane holds one list entry of a 64-entry
// chunk and whether it lies inside the cut; pack_dest gives the lane the entry
// moves to (inside entries to their rank in list order, the others after
// them), and packed() moves any per-entry value there through ds_permute, so
// lane l < count then holds the l-th inside entry's values.
__device__ __forceinline__ int pack_dest(bool in, int* count) {
    const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
    const int cnt = __popcll(m);
    const int lane = threadIdx.x & 63;
    const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
    *count = cnt;
    return in ? rank : cnt + (lane - rank);
}
__device__ __forceinline__ int packed(int dest, int v) { return __builtin_amdgcn_ds_permute(dest << 2, v); }
__device__ __forceinline__ double packed(int dest, double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_permute(dest << 2, (int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_ds_permute(dest << 2, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ int pack_in_range(int p, bool in, int* out) {
    int cnt;
    *out = packed(pack_dest(in, &cnt), p);
    return cnt;
}

// ---------------------------------------------------------------- pass 1
// Wave per site, PARTS lanes per ordered type pair: lane l = k*LPP + q*PARTS + part.
// A spline row of one type pair holds rho c0..c3 and r*phi c0..c3.
// PARTS = 4: the four part lanes load one quarter each (16 bytes in FP64), so
// each table row costs one cache-line access; each evaluates half of a cubic
// (c_lo + c_hi p) and one DPP exchange joins the halves (val = h0 + p^2 h1) in
// lane part 0 (rho) and lane part 2 (phi).
// PARTS = 2: part 0 loads the rho half of the row and evaluates its cubic,
// part 1 the phi half: twice the pairs per wavefront iteration and half the
// lanes repeating the point set-up, for one more cache-line access per row in
// FP64 (in FP32 a half row is one 16-byte load).
// PARTS = 1: one lane loads the whole row and evaluates both cubics.
// Every variant evaluates a cubic as (c0 + c1 p) + p^2 (c2 + c3 p), so the
// results do not depend on the variant.
// The rho and phi lanes then carry the weighted Gauss-Hermite sums of their
// function.
//
// The derivative records stream out with nontemporal stores (whole lines, read
// back only after the whole pass, so they are kept from displacing the spline
// tables in L2); the partner shares, 16 bytes at scattered reverse positions,
// use ordinary stores so that L2 merges them into lines before they leave.
//
// ENERGY_ONLY restates energy_calc's triangle (jtype starts at itype, so phi
// only for itype == jtype; src/ff_dmd_I.cpp:508).  The partner's density share
// goes to pairE in reverse-CSR order (rpos), so the embedding reads it
// contiguously.  Unless ENERGY_ONLY, the rho lane stores {drho/dr, drho/dalpha}
// and the phi lane {dphi/dr, dphi/dalpha} in the record (DerivRec) of the k-th
// pair of site i inside the quadrature cut, at off[i] + k (the reference's
// per-pair caches, src/ff_dmd_I.cpp:227-242, restricted to what the back-pass
// reads): a site's records are contiguous, and pass 2 ranks its inside pairs
// the same way.  The phi lanes of the type pairs it < jt neither evaluate nor
// store: the back-pass takes their phi from the transposed pair.
template <int NT, int PARTS = 4>
struct Pass1Map {
    static constexpr int LPP = NT * NT * PARTS;
    static constexpr int PPW = 64 / LPP;
};

// The derivative record of one pair inside the cut, kLen values of the
// quadrature type: {drho/dr, drho/dalpha} of every ordered type pair
// q = it*NT + jt, then {dphi/dr, dphi/dalpha} of the type pairs it >= jt only.
// The back-pass uses the phi of (it, jt) for it >= jt on both (it, jt) and
// (jt, it), as ForceField_DMD_I::force_calc reads one phi cache slot for both
// orders (src/ff_dmd_I.cpp:244-248, :376), so the other half is never read.
template <int NT>
struct DerivRec {
    static constexpr int kLen = 2 * NT * NT + NT * (NT + 1);
    __host__ __device__ static constexpr int rho(int q) { return 2 * q; }
    __host__ __device__ static constexpr int phi(int it, int jt) { return 2 * NT * NT + it * (it + 1) + 2 * jt; }
};

template <typename Q>
struct Vec2;
template <>
struct Vec2<double> {
    using T = double2;
};
template <>
struct Vec2<float> {
    using T = float2;
};

template <typename Q>
__device__ __forceinline__ Q dpp_xor1_q(Q v);
template <>
__device__ __forceinline__ double dpp_xor1_q<double>(double v) {
    return dpp_f64<kDppXor1>(v);
}
template <>
__device__ __forceinline__ float dpp_xor1_q<float>(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kDppXor1, 0xF, 0xF, false));
}


// Per-wavefront LDS stage of pass 1 (dynamic shared memory, kPass1StageBytes per
// wavefront): for the inside entries of the current chunk, in rank order,
// {list entry, partner, reverse position} and per type {alpha, 1/alpha, c, r^2}.
template <int NT>
struct Pass1Stage {
    static constexpr int kBytes = 64 * (16 + 32 * NT);
    int4* idx;
    double4* rec;
    __device__ __forceinline__ explicit Pass1Stage(int wave) {
        extern __shared__ __attribute__((aligned(32))) unsigned char pass1_lds[];
        unsigned char* base = pass1_lds + (size_t)wave * kBytes;
        rec = reinterpret_cast<double4*>(base);
        idx = reinterpret_cast<int4*>(base + 64 * 32 * NT);
    }
};

// Quadrature-table rows through buffer loads: each point costs a 32-bit byte
// offset (one multiply-add) instead of a 64-bit address, and the resource's
// size bounds every access.
template <typename Q2>
struct RowLoad;
template <>
struct RowLoad<double2> {
    static __device__ __forceinline__ double2 get(__amdgpu_buffer_rsrc_t r, int off) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        return make_double2(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
    }
};
template <>
struct RowLoad<float2> {
    static __device__ __forceinline__ float2 get(__amdgpu_buffer_rsrc_t r, int off) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        return make_float2(__int_as_float((int)v[0]), __int_as_float((int)v[1]));
    }
};
// raw buffer (stride 0, 32-bit data format) over n bytes
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* p, int nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, nbytes, 0x00020000);
}

// ENERGY_ONLY: the loop of ForceField_DMD_I::energy_calc (src/ff_dmd_I.cpp:454-668),
// type pairs jt >= it only (:508) unless FULL, which sums every ordered pair so the
// energy equals force_calc's.
template <int NT, typename Q, int NQ, bool ENERGY_ONLY, bool FULL = false, int PARTS = 4>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_pass1(FFParams<double> P, FFParams<Q> PQ,
                                                  const double4* __restrict__ pos,
                                                  const double4* __restrict__ arec, PairBufs L,
                                                  const int2* nlrp, const Q* __restrict__ quad,
                                                  const double* __restrict__ rho7, const double* __restrict__ phi7,
                                                  double* __restrict__ Eown, double* __restrict__ pairE,
                                                  Q* __restrict__ deriv, double* __restrict__ part,
                                                  CutState C) {
    static_assert(PARTS == 1 || PARTS == 2 || PARTS == 4, "one, two or four lanes per (pair, type pair)");
    using Q2 = typename Vec2<Q>::T;
    constexpr int LPP = Pass1Map<NT, PARTS>::LPP, PPW = Pass1Map<NT, PARTS>::PPW;
    const int lane = threadIdx.x & 63;
    const int k = lane / LPP, q = (lane % LPP) / PARTS, prt = lane % PARTS;
    const int it = q / NT, jt = q % NT;
    const bool active = !ENERGY_ONLY || FULL || jt >= it;
    const bool want_phi = it >= jt;
    // four parts: quarters of the row, rho in parts 0/1 and phi in 2/3 (results in 0 and 2);
    // two parts: part 0 the rho cubic, part 1 the phi cubic; one part: both
    const bool rho_lane = prt == 0, phi_lane = prt == PARTS / 2;
    double en = 0.0, npair = 0.0;  // energy, pairs inside the quadrature cut (this lane's share)
    const Pass1Stage<NT> st(threadIdx.x >> 6);
    const SiteLoop sl(P.natms);
    for (int i0 = sl.first; i0 < sl.end; i0 += sl.stride) {
        const int i = __builtin_amdgcn_readfirstlane(i0);
        const double4 xi = pos[i];
        double ci[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) ci[t] = arec[(size_t)i * NT + t].y;
        const double4 ra = arec[(size_t)i * NT + it];
        const double ai = ra.x, iai = ra.z;
        const __amdgpu_buffer_rsrc_t qrs = raw_rsrc(quad, P.nr * quad_stride(NT) * (int)sizeof(Q));
        const int tab_off = (q * kQuadRow + (PARTS == 1 ? 0 : prt * (8 / PARTS))) * (int)sizeof(Q);
        double Ei = 0.0;
        const int p0 = L.off[i], p1 = L.off[i + 1];
        int done = 0;  // inside entries of the earlier chunks: records go to off[i] + rank
        for (int chunk = p0; chunk < p1; chunk += 64) {
            // one list entry per lane: the cut state of each entry is kept current
            // (CutState); each entry inside the cut is staged in LDS at its rank
            // among the inside entries, with its partner's width, inverse width
            // and concentration of every type, read from HBM once per entry
            int cnt;
            {
                const int pc = chunk + lane;
                bool in = false;
                int jc = 0, rp = 0;
                double r2 = 0.0;
                if (pc < p1) {
                    const int2 e = nlrp[pc];
                    jc = list_partner(e);
                    rp = e.y;  // meaningful for owned partners only
                    const double4 xj = pos[jc];
                    const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
                    {
#pragma clang fp contract(off)
                        r2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
                    }
                    in = r2 < P.cut_sq_mod;
                    C.update<NT>(pc, e, in, jc, P.natms, pairE);
                }
                const int dest = pack_dest(in, &cnt);
                if (in) {
                    st.idx[dest] = make_int4(pc, jc, rp, 0);
                    // the partner's {alpha, c} of every type in one 16-byte load each; the
                    // division gives the same bits as the stored 1/alpha (both correctly rounded)
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const double2 ac = *reinterpret_cast<const double2*>(arec + (size_t)jc * NT + t);
                        st.rec[dest * NT + t] = make_double4(ac.x, 1.0 / ac.x, ac.y, r2);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
        for (int base = 0; base < cnt; base += PPW) {
            const int slot = base + k;
            const bool valid = slot < cnt && (64 % LPP == 0 || k < PPW);  // lanes past PPW pairs idle
            int p = 0, j = 0, rp = 0;
            double rsq = 1.0, aj = 1.0, iaj = 1.0;
            double cj[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) cj[t] = 0.0;
            if (valid) {
                const int4 e = st.idx[slot];
                p = e.x;
                j = e.y;
                rp = e.z;
                const double4 rj = st.rec[slot * NT + jt];
                aj = rj.x;
                iaj = rj.y;
                rsq = rj.w;
#pragma unroll
                for (int t = 0; t < NT; ++t) cj[t] = t == jt ? rj.z : st.rec[slot * NT + t].z;
            }
            const bool inr = valid;  // inside the cut by construction
            double ej = 0.0;
            if (inr) {
                const double r_inv = rsqrt(rsq), r = rsq * r_inv;
                const double ia = iai + iaj, asq = rsqrt(ia);
                // the phi lanes of it < jt are idle (DerivRec)
                const bool live = PARTS == 1 || prt < PARTS / 2 || want_phi;
                const int br = active && live ? alpha_branch(P, ia, ai, aj) : 0;
                // v1: the lane's function (rho on the rho lane, phi on the phi lane) or,
                // with one lane per pair, rho; v2: phi with one lane per pair.
                // Each holds value, d/dr, d/dalpha.
                double v1[3] = {0.0, 0.0, 0.0};
                double v2[3] = {0.0, 0.0, 0.0};
                if (br == 1) {
                    const Q rq = (Q)r, iq = (Q)ia, aq = (Q)asq;
                    const Q inv = iq * aq;
                    const Q upper = (rq + PQ.rc) * aq, lower = (rq - PQ.rc) * aq;
                    const int nq = NQ > 0 ? NQ : PQ.no_i;
                    const bool rule = lower < PQ.xi[nq - 1];
                    Q s0 = 0, s1 = 0, s2 = 0;  // rho sums (PARTS 1) or this lane's function
                    Q u0 = 0, u1 = 0, u2 = 0;  // phi sums (PARTS 1)
                    // all rows of the rule are requested before any is used, so their
                    // latencies overlap (a point outside the window reads a clamped row)
                    constexpr int KP = NQ > 0 ? NQ : kMaxPoints;
                    constexpr int NV = PARTS == 4 ? 1 : (PARTS == 2 ? 2 : 4);
                    Q2 rowv[KP][NV];
                    Q ppv[KP];
#pragma unroll
                    for (int kp = 0; kp < KP; ++kp) {
                        if (NQ == 0 && kp >= nq) break;
                        const Q t = rq - PQ.xi[kp] * inv;
                        Q pp = fabs(t) * PQ.dr_inv;
                        int m = (int)pp;
                        m = m < P.nr - 2 ? m : P.nr - 2;
                        m = m > 0 ? m : 0;
                        pp -= (Q)m;
                        ppv[kp] = pp < Q(1) ? pp : Q(1);
                        const int row = tab_off + m * (quad_stride(NT) * (int)sizeof(Q));
#pragma unroll
                        for (int v = 0; v < NV; ++v)  // one lane per pair: the phi half only where it >= jt
                            rowv[kp][v] = PARTS != 1 || v < 2 || want_phi
                                              ? RowLoad<Q2>::get(qrs, row + v * (int)sizeof(Q2))
                                              : Q2{Q(0), Q(0)};
                    }
                    // keep the requests ahead of the arithmetic where the rows fit the
                    // register budget of four waves per SIMD (at most 24 registers)
                    if constexpr (KP * NV * sizeof(Q2) <= 96) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int kp = 0; kp < KP; ++kp) {
                        if (NQ == 0 && kp >= nq) break;
                        const Q xk = PQ.xi[kp];
                        const bool inside = rule && xk > lower && xk < upper;
                        const Q t = rq - xk * inv, pp = ppv[kp];
                        const Q p2 = pp * pp;
                        if constexpr (PARTS == 4) {
                            const Q2 c = rowv[kp][0];
                            const Q h = c.x + c.y * pp;
                            const Q hb = dpp_xor1_q<Q>(h);
                            const Q val = h + p2 * hb;
                            Q x = prt == 0 ? val * t : (t < Q(0) ? -val : val);
                            x = inside ? x : Q(0);
                            s0 += PQ.w0[kp] * x;
                            if (!ENERGY_ONLY) {
                                s1 += PQ.w1[kp] * x;
                                s2 += PQ.w2[kp] * x;
                            }
                        } else if constexpr (PARTS == 2) {
                            const Q2 lo = rowv[kp][0];
                            const Q2 hi = rowv[kp][1];
                            const Q val = (lo.x + lo.y * pp) + p2 * (hi.x + hi.y * pp);
                            Q x = prt == 0 ? val * t : (t < Q(0) ? -val : val);
                            x = inside ? x : Q(0);
                            s0 += PQ.w0[kp] * x;
                            if (!ENERGY_ONLY) {
                                s1 += PQ.w1[kp] * x;
                                s2 += PQ.w2[kp] * x;
                            }
                        } else {
                            const Q2 a0 = rowv[kp][0];
                            const Q2 a1 = rowv[kp][1];
                            const Q2 b0 = rowv[kp][2];
                            const Q2 b1 = rowv[kp][3];
                            const Q vrho = (a0.x + a0.y * pp) + p2 * (a1.x + a1.y * pp);
                            const Q vphi = (b0.x + b0.y * pp) + p2 * (b1.x + b1.y * pp);
                            const Q xr = inside ? vrho * t : Q(0);
                            const Q xp = inside ? (t < Q(0) ? -vphi : vphi) : Q(0);
                            s0 += PQ.w0[kp] * xr;
                            u0 += PQ.w0[kp] * xp;
                            if (!ENERGY_ONLY) {
                                s1 += PQ.w1[kp] * xr;
                                s2 += PQ.w2[kp] * xr;
                                u1 += PQ.w1[kp] * xp;
                                u2 += PQ.w2[kp] * xp;
                            }
                        }
                    }
                    if (rule) {
                        const Q sc = Q(kPiInSq) * (Q)r_inv;
                        auto fin = [&](Q a0, Q a1, Q a2, double* v) {
                            const Q v0 = a0 * sc;
                            v[0] = (double)v0;
                            if (!ENERGY_ONLY) {
                                v[1] = (double)(-v0 * (Q)r_inv - Q(2) * aq * (a1 * sc));
                                v[2] = (double)((Q(0.5) * v0 - a2 * sc) * iq);
                            }
                        };
                        fin(s0, s1, s2, v1);
                        if constexpr (PARTS == 1) fin(u0, u1, u2, v2);
                    }
                } else if (br == 2 && rsq < P.cut_sq) {
                    double g[3], f[3];
                    classical_pair(P, rsq, r_inv, rho7 + (size_t)type2rho(NT, it, jt) * P.nr * 7,
                                   phi7 + (size_t)type2phi(it, jt) * P.nr * 7, true, g, f);
                    for (int c = 0; c < 3; ++c) {
                        if constexpr (PARTS == 1) {
                            v1[c] = g[c];
                            v2[c] = f[c];
                        } else {
                            v1[c] = rho_lane ? g[c] : f[c];
                        }
                    }
                }
                if constexpr (PARTS == 1) {
                    Ei += v1[0] * cj[jt];
                    ej = v1[0] * ci[it];
                    if (want_phi && active) {
                        const double half = (j < P.natms ? 0.5 : 0.25) * (it != jt ? P.xw : 1.0);
                        en += half * (cj[jt] * ci[it] + cj[it] * ci[jt]) * v2[0];
                    }
                    if (!ENERGY_ONLY) {
                        Q* d = deriv + (size_t)(p0 + done + slot) * DerivRec<NT>::kLen;
                        __builtin_nontemporal_store((Q)v1[1], d + DerivRec<NT>::rho(q));
                        __builtin_nontemporal_store((Q)v1[2], d + DerivRec<NT>::rho(q) + 1);
                        if (want_phi) {
                            __builtin_nontemporal_store((Q)v2[1], d + DerivRec<NT>::phi(it, jt));
                            __builtin_nontemporal_store((Q)v2[2], d + DerivRec<NT>::phi(it, jt) + 1);
                        }
                    }
                } else {
                    if (rho_lane) {
                        Ei += v1[0] * cj[jt];
                        ej = v1[0] * ci[it];
                    }
                    if (phi_lane && want_phi && active) {
                        const double half = (j < P.natms ? 0.5 : 0.25) * (it != jt ? P.xw : 1.0);
                        en += half * (cj[jt] * ci[it] + cj[it] * ci[jt]) * v1[0];
                    }
                    if (!ENERGY_ONLY && (rho_lane || (phi_lane && want_phi))) {
                        Q* d = deriv + (size_t)(p0 + done + slot) * DerivRec<NT>::kLen +
                               (rho_lane ? DerivRec<NT>::rho(q) : DerivRec<NT>::phi(it, jt));
                        __builtin_nontemporal_store((Q)v1[1], d);
                        __builtin_nontemporal_store((Q)v1[2], d + 1);
                    }
                }
            }
            // sum over it: the (0, jt) and (1, jt) lanes of a pair are NT * PARTS apart
            if constexpr (NT == 2 && PARTS == 4) ej += dpp_f64<kDppRowRor8>(ej);
            if constexpr (NT == 2 && PARTS == 2) ej += dpp_f64<kDppRowRor12>(ej);  // lane l + 4 for l % 8 < 4
            if constexpr (NT == 2 && PARTS == 1) ej += __shfl_xor(ej, 2, 64);
            if constexpr (NT > 2) {  // the rho lanes (0..NT-1, jt) of the pair, in a fixed order
                const int col = lane - lane % LPP + jt * PARTS;
                double s = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t) s += __shfl(ej, col + t * NT * PARTS, 64);
                ej = s;
            }
            // an ordinary store: the 16-byte shares land scattered, and L2 merges them into lines
            if (valid && rho_lane && it == 0 && j < P.natms) pairE[(size_t)rp * NT + jt] = ej;
        }
        done += cnt;
        }
        if (lane == 0) npair += (double)done;  // the site's pairs inside the cut
        if constexpr (NT == 2 && PARTS == 2) {
            // Ei[it] lives on the rho lanes 8k + 2q (q = 2 it + jt): join the jt lanes
            // (then lanes 8k + 4 it hold slot k's share), add the two slots of each row
            // and the four rows
            double e = rho_lane ? Ei : 0.0;
            e += dpp_f64<kDppXor2>(e);
            e += dpp_f64<kDppRowRor8>(e);
            e = row_pair_sum<false>(e);
            e = row_pair_sum<true>(e);
            const double e0 = read_lane(e, 0), e1 = read_lane(e, 4);
            if (lane == 0) {
                Eown[(size_t)i * NT] = e0;
                Eown[(size_t)i * NT + 1] = e1;
            }
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double s = wave_sum(rho_lane && it == t ? Ei : 0.0);
                if (lane == 0) Eown[(size_t)i * NT + t] = s;
            }
        }
    }
    double acc[2] = {wave_sum(en), npair};  // npair: lane 0 only
    wave_store<2>(acc, part);
}

// ---------------------------------------------------------------- lane groups
// G consecutive lanes per site (a DPP row for G = 16, a quad for G = 4);
// blocks are remapped so that each XCD takes a contiguous range of sites.
template <int G>
struct SiteGroup {
    int site, l;
    __device__ __forceinline__ SiteGroup() {
        const int b = xcd_block(gridDim.x);
        site = (b * kBlock + (int)threadIdx.x) / G;
        l = threadIdx.x & (G - 1);
    }
};

// sum over the G lanes of a group (fixed tree; every lane of the group gets it)
template <int G>
__device__ __forceinline__ double group_sum(double v) {
    static_assert(G == 16 || G == 4, "lane groups of 4 or 16");
    if constexpr (G == 16) {
        v += dpp_f64<kDppRowRor8>(v);
        v += dpp_f64<kDppRowRor4>(v);
    }
    v += dpp_f64<kDppXor2>(v);
    v += dpp_f64<kDppXor1>(v);
    return v;
}

// value v[t] of the lane whose group index is t (t < N)
template <int N>
__device__ __forceinline__ double pick(const double* v, int t) {
    double r = v[0];
#pragma unroll
    for (int x = 1; x < N; ++x)
        if (t == x) r = v[x];
    return r;
}

constexpr int kGatherG = 16;  // lanes per site of the reverse-CSR gathers
constexpr int kEmbedG = 4;    // lanes per site of the embedding (its per-site arithmetic runs on NT lanes of each group)

// ---------------------------------------------------------------- embed
// rho_bar = own share + partner shares (kEmbedG lanes per site over the site's
// contiguous reverse range); F, F', energy terms, alpha entropic force
template <int NT, bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_embed(FFParams<double> P, const double* __restrict__ alpha,
                                                  const double* __restrict__ conc, PairBufs L,
                                                  const double* __restrict__ Eown,
                                                  const double* __restrict__ pairE,
                                                  const double* __restrict__ F9, double* __restrict__ E,
                                                  double* __restrict__ dE, double* __restrict__ fa,
                                                  double4* __restrict__ arec, double* __restrict__ part) {
    const SiteGroup<kEmbedG> g;
    const int i = g.site;
    double acc[1] = {0.0};
    if (i < P.natms) {
        double e[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) e[t] = 0.0;
        const int q1 = L.roff[i + 1];
        for (int q = L.roff[i] + g.l; q < q1; q += kEmbedG) {
#pragma unroll
            for (int t = 0; t < NT; ++t) e[t] += pairE[(size_t)q * NT + t];
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) e[t] = group_sum<kEmbedG>(e[t]);
        if (g.l < NT) {
            const int t = g.l;
            const size_t k = (size_t)i * NT + t;
            const double et = Eown[k] + pick<NT>(e, t);
            double Fv, dF;
            embed_F(P, F9, t, et, &Fv, &dF, nullptr);
            const double c = conc[k];
            const double a = alpha[k];
            E[k] = Fv;
            if (!ENERGY_ONLY) {
                dE[k] = dF;
                arec[k].w = dF;
                fa[k] = -(1.5 * P.kbT * c / a);
            }
            acc[0] += c * Fv;
            acc[0] += P.kbT * calc_ent(c);
            acc[0] += 1.5 * P.kbT * c * log(a) + c * P.c0[t];
        }
    }
    acc[0] = wave_sum(acc[0]);
    wave_store<1>(acc, part);
}

// ---------------------------------------------------------------- pass 2
// src/ff_dmd_I.cpp:357-436: fpair/apair of each ordered type pair from the
// derivative records of pass 1 (slot type2rho[jt][it], i.e. the transposed
// lane's record, pairs with dE_i; phi for it < jt is the phi record of
// (jt, it), as in the reference's shared COMP slot).  No table lookups.
template <int NT, typename Q>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_pass2(FFParams<double> P, const double4* __restrict__ pos,
                                                  const double4* __restrict__ arec, PairBufs L,
                                                  const int2* __restrict__ nlrp, const Q* __restrict__ deriv,
                                                  double* __restrict__ fown, ForceShares F,
                                                  double* __restrict__ part) {
    constexpr int LPP = WaveMap<NT>::LPP, PPW = WaveMap<NT>::PPW;
    constexpr int XD = 3 + NT;
    const int lane = threadIdx.x & 63;
    const int k = lane / LPP, q = lane % LPP, it = q / NT, jt = q % NT;
    const int qT = jt * NT + it;
    const bool want_phi = it >= jt;
    double vir[6] = {0, 0, 0, 0, 0, 0};
    const SiteLoop sl(P.natms);
    for (int i0 = sl.first; i0 < sl.end; i0 += sl.stride) {
        const int i = __builtin_amdgcn_readfirstlane(i0);
        const double4 xi = pos[i];
        const double4 ra = arec[(size_t)i * NT + it];
        const double ai = ra.x, ci = ra.y, iai = ra.z, dEi = ra.w;
        double fx = 0.0, fy = 0.0, fz = 0.0, fa = 0.0;
        const int p0 = L.off[i], p1 = L.off[i + 1];
        int done = 0;  // inside entries of the earlier chunks (records at off[i] + rank)
        for (int chunk = p0; chunk < p1; chunk += 64) {
            // partner and separation of the inside pairs, packed
            int pk_j, pk_rp, cnt;
            double pk_d0, pk_d1, pk_d2, pk_r2;
            {
                const int pc = chunk + lane;
                bool in = false;
                int jc = 0, rp = 0;
                double e0 = 0.0, e1 = 0.0, e2 = 0.0, r2 = 0.0;
                if (pc < p1) {
                    // pass 1 of this evaluation left each entry's cut state in its
                    // partner's top bit (same positions): only the entries inside
                    // need a position; the shares of the others stay zero (CutState)
                    const int2 e = nlrp[pc];
                    in = (e.x & kInsideBit) != 0;
                    if (in) {
                        jc = list_partner(e);
                        rp = e.y;  // meaningful for owned partners only
                        const double4 xj = pos[jc];
                        e0 = xi.x - xj.x;
                        e1 = xi.y - xj.y;
                        e2 = xi.z - xj.z;
                        {
#pragma clang fp contract(off)
                            r2 = e0 * e0 + e1 * e1 + e2 * e2;
                        }
                    }
                }
                const int dest = pack_dest(in, &cnt);
                pk_j = packed(dest, jc);
                pk_rp = packed(dest, rp);
                pk_d0 = packed(dest, e0);
                pk_d1 = packed(dest, e1);
                pk_d2 = packed(dest, e2);
                pk_r2 = packed(dest, r2);
            }
        for (int base = 0; base < cnt; base += PPW) {
            const int slot = base + k;
            const bool valid = slot < cnt && (64 % LPP == 0 || k < PPW);  // lanes past PPW pairs idle
            const int src = valid ? slot : 0;
            const int j = __shfl(pk_j, src, 64);
            const int rp = __shfl(pk_rp, src, 64);
            const double dx0 = __shfl(pk_d0, src, 64), dx1 = __shfl(pk_d1, src, 64), dx2 = __shfl(pk_d2, src, 64);
            const double rsq = __shfl(pk_r2, src, 64);
            const bool inr = valid;  // inside the cut by construction
            const bool owned = j < P.natms;
            double S = 0.0, aJ = 0.0;
            if (inr) {
                double aj, iaj, cj, dEj;
                if constexpr (NT == 2) {
                    // lanes (0, jt) and (1, jt) share the partner's record of type jt:
                    // each loads one half ({alpha, c} or {1/alpha, F'}) and the quad
                    // exchange (xor 2) brings the other
                    const double2 h = reinterpret_cast<const double2*>(arec + (size_t)j * NT + jt)[it];
                    const double o0 = dpp_f64<kDppXor2>(h.x), o1 = dpp_f64<kDppXor2>(h.y);
                    aj = it == 0 ? h.x : o0;
                    cj = it == 0 ? h.y : o1;
                    iaj = it == 0 ? o0 : h.x;
                    dEj = it == 0 ? o1 : h.y;
                } else {
                    const double4 rj = arec[(size_t)j * NT + jt];
                    aj = rj.x;
                    cj = rj.y;
                    iaj = rj.z;
                    dEj = rj.w;
                }
                const double r_inv = rsqrt(rsq);
                // the lane's own density record; the transposed type pair's comes from
                // its lane (transpose_tp)
                const Q* rec = deriv + (size_t)(p0 + done + slot) * DerivRec<NT>::kLen;
                const typename Vec2<Q>::T rg =
                    *reinterpret_cast<const typename Vec2<Q>::T*>(rec + DerivRec<NT>::rho(q));
                // one phi slot serves (it, jt) and (jt, it)
                const typename Vec2<Q>::T rf = *reinterpret_cast<const typename Vec2<Q>::T*>(
                    rec + DerivRec<NT>::phi(want_phi ? it : jt, want_phi ? jt : it));
                const double g_dr = (double)rg.x, g_da = (double)rg.y;
                const double f_dr = (double)rf.x, f_da = (double)rf.y;
                const double gT_dr = transpose_tp<NT>(g_dr), gT_da = transpose_tp<NT>(g_da);
                const double fr_dr = f_dr, fr_da = f_da;
                const double cc = ci * cj;
                // 1/(alpha_i + alpha_j) = (1/alpha_i)(1/alpha_j) / (1/alpha_i + 1/alpha_j)
                const double ia = iai + iaj;
                double ria = __builtin_amdgcn_rcp(ia);
                ria = ria * (2.0 - ia * ria);
                ria = ria * (2.0 - ia * ria);
                const double is = iai * iaj * ria;
                const double fpair = -(gT_dr * dEi + g_dr * dEj + fr_dr) * cc * r_inv;
                const double apair = -(gT_da * dEi + g_da * dEj + fr_da) * cc * (is * is);
                fa += apair * aj * aj;
                aJ = apair * ai * ai;
                S = fpair;
            }
            S = sum_over_it<NT>(sum_over_jt<NT>(S));
            aJ = sum_over_it<NT>(aJ);
            if (inr && q == 0) {
                fx += dx0 * S;
                fy += dx1 * S;
                fz += dx2 * S;
                const double wf = owned ? S : 0.5 * S;
                vir[0] -= wf * dx0 * dx0;
                vir[1] -= wf * dx1 * dx1;
                vir[2] -= wf * dx2 * dx2;
                vir[3] -= wf * dx1 * dx2;
                vir[4] -= wf * dx2 * dx0;
                vir[5] -= wf * dx0 * dx1;
            }
            if constexpr (NT == 2) {
                // lanes 0 and 1 of the pair store {fx, fy} and {fz, alpha force 0} (32
                // contiguous bytes), lane 2 the alpha force of type 1: aJ of type jt
                // lives on the lanes it = 0, 1 of that jt, moved to lanes 1 and 2
                const double aJx = dpp_f64<kDppQuad0033>(aJ);
                if (valid && owned) {
                    if (q < 2) {
                        const double2 v = q == 0 ? make_double2(-dx0 * S, -dx1 * S) : make_double2(-dx2 * S, aJx);
                        reinterpret_cast<double2*>(F.sh4 + (size_t)rp * 4)[q] = v;
                    } else if (q == 2) {
                        F.shx[rp] = aJx;
                    }
                }
            } else {
                if (valid && owned) {
                    double* dst = F.sh4 + (size_t)rp * 4;
                    if (q == 0) {
                        dst[0] = -dx0 * S;
                        dst[1] = -dx1 * S;
                        dst[2] = -dx2 * S;
                    }
                    if (it == 0) {
                        if (jt == 0) dst[3] = aJ;
                        else F.shx[(size_t)rp * (NT - 1) + jt - 1] = aJ;
                    }
                }
            }
        }
        done += cnt;
        }
        double sx, sy, sz, sa[NT];
        if constexpr (NT == 2) {
            // fx, fy, fz live on the lanes q = 0 and fa[it] on q = 2 it + jt: join the
            // jt lanes (q = 0 and 2 then hold fa[0], fa[1]), pack the five sums into two
            // registers -- q = 0: (fx, fy), q = 1: (fz, fa[0]), q = 2: (fa[1], 0) -- and
            // sum both over the 16 pair slots at once
            const double fa2 = fa + dpp_f64<kDppXor1>(fa);
            const double fz1 = dpp_f64<kDppQuad0022>(fz), fa1 = dpp_f64<kDppQuad0022>(fa2);
            double A = q == 0 ? fx : (q == 1 ? fz1 : (q == 2 ? fa2 : 0.0));
            double B = q == 0 ? fy : (q == 1 ? fa1 : 0.0);
            quad_slot_sums(A, B);
            sx = read_lane(A, 0);
            sy = read_lane(B, 0);
            sz = read_lane(A, 1);
            sa[0] = read_lane(B, 1);
            sa[1] = read_lane(A, 2);
        } else {
            sx = wave_sum(fx);
            sy = wave_sum(fy);
            sz = wave_sum(fz);
#pragma unroll
            for (int t = 0; t < NT; ++t) sa[t] = wave_sum(it == t ? fa : 0.0);
        }
        if (lane == 0) {
            fown[(size_t)i * XD + 0] = sx;
            fown[(size_t)i * XD + 1] = sy;
            fown[(size_t)i * XD + 2] = sz;
#pragma unroll
            for (int t = 0; t < NT; ++t) fown[(size_t)i * XD + 3 + t] = sa[t];
        }
    }
    if constexpr (NT == 2) {
        // the virial lives on the lanes q = 0: q = 0 packs (v0, v1), q = 1 (v2, v3), q = 2 (v4, v5)
        double w[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) w[c] = dpp_f64<kDppQuad0000>(vir[2 + c]);
        double A = q == 0 ? vir[0] : (q == 1 ? w[0] : (q == 2 ? w[2] : 0.0));
        double B = q == 0 ? vir[1] : (q == 1 ? w[1] : (q == 2 ? w[3] : 0.0));
        quad_slot_sums(A, B);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vir[2 * c] = read_lane(A, c);
            vir[2 * c + 1] = read_lane(B, c);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) vir[c] = wave_sum(vir[c]);
    }
    wave_store<6>(vir, part);
}


// ---------------------------------------------------------------- force gather
template <int NT>
__global__ __launch_bounds__(kBlock) void k_fgather(int natms, PairBufs L, const double* __restrict__ fown,
                                                    ForceShares F,
                                                    const double* __restrict__ fa, double* __restrict__ f) {
    constexpr int XD = 3 + NT;
    const SiteGroup<kGatherG> g;
    const int i = g.site;
    if (i >= natms) return;
    double acc[XD];
#pragma unroll
    for (int k = 0; k < XD; ++k) acc[k] = 0.0;
    const int q1 = L.roff[i + 1];
    for (int q = L.roff[i] + g.l; q < q1; q += kGatherG) {
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] += F.sh4[(size_t)q * 4 + k];
#pragma unroll
        for (int k = 4; k < XD; ++k) acc[k] += F.shx[(size_t)q * (NT - 1) + k - 4];
    }
#pragma unroll
    for (int k = 0; k < XD; ++k) acc[k] = group_sum<kGatherG>(acc[k]);
    if (g.l < XD) {
        const int k = g.l;
        double v = fown[(size_t)i * XD + k];
        if (k >= 3) v += fa[(size_t)i * NT + k - 3];
        f[(size_t)i * XD + k] = v + pick<XD>(acc, k);
    }
}

// Two-stage deterministic sum of per-block partials: stage 1 block b sums the
// contiguous range [b*chunk, (b+1)*chunk) of partials in a fixed tree; stage 2
// (one block) sums the stage-1 results the same way.
__global__ __launch_bounds__(kBlock) void k_reduce_stage(const double* __restrict__ part, long nparts, int nv,
                                                         int stride, int offset, long chunk,
                                                         double* __restrict__ out, int out_stride) {
    __shared__ double sh[kBlock];
    const long lo = (long)blockIdx.x * chunk;
    long hi = lo + chunk;
    hi = hi < nparts ? hi : nparts;
    for (int k = 0; k < nv; ++k) {
        double s = 0.0;
        for (long b = lo + threadIdx.x; b < hi; b += kBlock) s += part[(size_t)b * stride + offset + k];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * out_stride + k] = sh[0];
        __syncthreads();
    }
}

// The same two-stage sums for up to kMaxSegs arrays of partials (at most 8
// values each) in one pair of launches: segment s takes stage-1 blocks
// [first[s], first[s] + nb[s]) and stage-2 block s; each sum keeps the fixed
// shape of k_reduce_stage.
constexpr int kMaxSegs = 4;
struct ReduceSegs {
    const double* part[kMaxSegs];
    long nparts[kMaxSegs];
    int nv[kMaxSegs], stride[kMaxSegs], out_off[kMaxSegs], first[kMaxSegs], nb[kMaxSegs];
    int nseg;
};
__device__ __forceinline__ double block_tree_sum(double s, double* sh) {
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
__global__ __launch_bounds__(kBlock) void k_reduce_segs1(ReduceSegs R, long chunk, double* __restrict__ stage) {
    __shared__ double sh[kBlock];
    int sg = 0;
    while (sg + 1 < R.nseg && (int)blockIdx.x >= R.first[sg + 1]) ++sg;
    const long lo = (long)(blockIdx.x - R.first[sg]) * chunk;
    long hi = lo + chunk;
    hi = hi < R.nparts[sg] ? hi : R.nparts[sg];
    const double* part = R.part[sg];
    const int nv = R.nv[sg], stride = R.stride[sg];
    // every value of a partial in one sweep (they share its cache line), then one tree per value
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (long b = lo + threadIdx.x; b < hi; b += kBlock) {
        const double* pb = part + (size_t)b * stride;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nv) s[k] += pb[k];
    }
    for (int k = 0; k < nv; ++k) {
        const double t = block_tree_sum(s[k], sh);
        if (threadIdx.x == 0) stage[(size_t)blockIdx.x * 8 + k] = t;
    }
}
__global__ __launch_bounds__(kBlock) void k_reduce_segs2(ReduceSegs R, const double* __restrict__ stage,
                                                         double* __restrict__ out) {
    __shared__ double sh[kBlock];
    const int sg = blockIdx.x;
    for (int k = 0; k < R.nv[sg]; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < R.nb[sg]; b += kBlock) s += stage[(size_t)(R.first[sg] + b) * 8 + k];
        s = block_tree_sum(s, sh);
        if (threadIdx.x == 0) out[R.out_off[sg] + k] = s;
    }
}

}  // namespace mappdmd

#pragma clang fp contract(off)

// Force kernels of the DMD force field (ForceField_DMD_I::force_calc,
// src/ff_dmd_I.cpp:88-450, and energy_calc, :454-668).
//
// The reference scatters every half-list pair into both ends (E[j] +=,
// f[j] -=).  Here one lane owns one site and never writes another site's row:
//   pass 1   lane i walks its half list, evaluates the Gauss-Hermite averaged
//            primitives, keeps its own density sum in registers and writes
//            the partner's share of the pair (nt doubles) to a pair buffer;
//   embed    lane i adds the shares of the lower-index sites that list i
//            (reverse CSR, ascending pair index), then F(rho_bar), F', the
//            per-site energy terms and the alpha entropic force;
//   pass 2   lane i re-evaluates the primitives with their r- and alpha-
//            derivatives, accumulates its own force row and writes the
//            partner's share (3+nt doubles) to a pair buffer;
//   gather   lane i adds the partner shares (reverse CSR).
// No floating-point atomics: every sum has a fixed order, so results are
// bitwise reproducible run to run.  Energy and virial partials are reduced
// per block in a fixed tree and then over blocks in block order.
//
// Q is the arithmetic of the quadrature (double, or float for the mixed
// mode whose point sums run in FP32 with FP32 accumulation); site-level sums,
// F(rho_bar), energies and forces are FP64 in both.
#pragma once
#include "dmd_common.hpp"

namespace mappdmd {

constexpr int kBlock = 256;
constexpr int kQuadRow = 8;  // rho c0..c3, r*phi c0..c3 of one ordered type pair

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
                                           const Q* __restrict__ tab, Q* g, Q* f) {
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
            const Q* row = tab + (size_t)m * kQuadRow;
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

// ---------------------------------------------------------------- pass 1
// ENERGY_ONLY restates energy_calc's triangle (jtype starts at itype, so phi
// only for itype == jtype; src/ff_dmd_I.cpp:508)
template <int NT, typename Q, int NQ, bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_pass1(FFParams<double> P, FFParams<Q> PQ,
                                                  const double4* __restrict__ pos,
                                                  const double* __restrict__ alpha,
                                                  const double* __restrict__ conc, PairBufs L,
                                                  const Q* __restrict__ quad, const double* __restrict__ rho7,
                                                  const double* __restrict__ phi7, double* __restrict__ Eown,
                                                  double* __restrict__ pairE, double* __restrict__ part) {
    const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};  // energy, pairs inside the quadrature cut
    if (i < P.natms) {
        const double4 xi = pos[i];
        double ai[NT], ci[NT], Ei[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            ai[t] = alpha[(size_t)i * NT + t];
            ci[t] = conc[(size_t)i * NT + t];
            Ei[t] = 0.0;
        }
        const int p1 = L.off[i + 1];
        for (int p = L.off[i]; p < p1; ++p) {
            const int j = L.nl[p];
            const double4 xj = pos[j];
            const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
            const double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            double Ej[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) Ej[t] = 0.0;
            if (rsq < P.cut_sq_mod) {
                acc[1] += 1.0;
                double aj[NT], cj[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    aj[t] = alpha[(size_t)j * NT + t];
                    cj[t] = conc[(size_t)j * NT + t];
                }
                const double r = sqrt(rsq);
                const double r_inv = 1.0 / r;
                const double half = j < P.natms ? 0.5 : 0.25;
#pragma unroll
                for (int it = 0; it < NT; ++it) {
#pragma unroll
                    for (int jt = (ENERGY_ONLY ? it : 0); jt < NT; ++jt) {
                        const bool want_phi = it >= jt;
                        const double a = (ai[it] * aj[jt]) / (ai[it] + aj[jt]);
                        double g0 = 0.0, f0 = 0.0;
                        bool hit = false;
                        if (P.alpha_min < a && a < P.alpha_max) {
                            Q g[3], f[3];
                            const Q* tab = quad + (size_t)(it * NT + jt) * P.nr * kQuadRow;
                            if (want_phi)
                                hit = gh_average<Q, NQ, false, true>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, g, f);
                            else
                                hit = gh_average<Q, NQ, false, false>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, g, f);
                            if (hit) {
                                g0 = (double)g[0];
                                if (want_phi) f0 = (double)f[0];
                            }
                        } else if (P.alpha_max <= a && rsq < P.cut_sq) {
                            double g[3], f[3];
                            classical_pair(P, rsq, r_inv, rho7 + (size_t)type2rho(NT, it, jt) * P.nr * 7,
                                           phi7 + (size_t)type2phi(it, jt) * P.nr * 7, want_phi, g, f);
                            g0 = g[0];
                            if (want_phi) f0 = f[0];
                            hit = true;
                        }
                        if (hit) {
                            Ei[it] += g0 * cj[jt];
                            Ej[jt] += g0 * ci[it];
                            if (want_phi) acc[0] += half * (cj[jt] * ci[it] + cj[it] * ci[jt]) * f0;
                        }
                    }
                }
            }
            if (j < P.natms) {
#pragma unroll
                for (int t = 0; t < NT; ++t) pairE[(size_t)p * NT + t] = Ej[t];
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) Eown[(size_t)i * NT + t] = Ei[t];
    }
    block_sum_store<2>(acc, part);
}

// ---------------------------------------------------------------- embed
// rho_bar = own share + partner shares; F, F', energy terms, alpha entropic force
template <int NT, bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_embed(FFParams<double> P, const double* __restrict__ alpha,
                                                  const double* __restrict__ conc, PairBufs L,
                                                  const double* __restrict__ Eown,
                                                  const double* __restrict__ pairE,
                                                  const double* __restrict__ F9, double* __restrict__ E,
                                                  double* __restrict__ dE, double* __restrict__ fa,
                                                  double* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < P.natms) {
        double e[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) e[t] = Eown[(size_t)i * NT + t];
        const int q1 = L.roff[i + 1];
        for (int q = L.roff[i]; q < q1; ++q) {
            const size_t p = (size_t)L.rev[q];
#pragma unroll
            for (int t = 0; t < NT; ++t) e[t] += pairE[p * NT + t];
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            double Fv, dF;
            embed_F(P, F9, t, e[t], &Fv, &dF, nullptr);
            const double c = conc[(size_t)i * NT + t];
            const double a = alpha[(size_t)i * NT + t];
            E[(size_t)i * NT + t] = Fv;
            if (!ENERGY_ONLY) {
                dE[(size_t)i * NT + t] = dF;
                fa[(size_t)i * NT + t] = -(1.5 * P.kbT * c / a);
            }
            acc[0] += c * Fv;
            acc[0] += P.kbT * calc_ent(c);
            acc[0] += 1.5 * P.kbT * c * log(a) + c * P.c0[t];
        }
    }
    block_sum_store<1>(acc, part);
}

// ---------------------------------------------------------------- pass 2
template <int NT, typename Q, int NQ>
__global__ __launch_bounds__(kBlock) void k_pass2(FFParams<double> P, FFParams<Q> PQ,
                                                  const double4* __restrict__ pos,
                                                  const double* __restrict__ alpha,
                                                  const double* __restrict__ conc, PairBufs L,
                                                  const Q* __restrict__ quad, const double* __restrict__ rho7,
                                                  const double* __restrict__ phi7, const double* __restrict__ dE,
                                                  double* __restrict__ fown, double* __restrict__ pairF,
                                                  double* __restrict__ part) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    constexpr int XD = 3 + NT;
    const int i = xcd_block(gridDim.x) * kBlock + threadIdx.x;
    double vir[6] = {0, 0, 0, 0, 0, 0};
    if (i < P.natms) {
        const double4 xi = pos[i];
        double ai[NT], ci[NT], dEi[NT], fi[XD];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            ai[t] = alpha[(size_t)i * NT + t];
            ci[t] = conc[(size_t)i * NT + t];
            dEi[t] = dE[(size_t)i * NT + t];
        }
#pragma unroll
        for (int k = 0; k < XD; ++k) fi[k] = 0.0;
        const int p1 = L.off[i + 1];
        for (int p = L.off[i]; p < p1; ++p) {
            const int j = L.nl[p];
            const double4 xj = pos[j];
            const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
            const double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            const bool owned = j < P.natms;
            double pf[XD];
#pragma unroll
            for (int k = 0; k < XD; ++k) pf[k] = 0.0;
            if (rsq < P.cut_sq_mod) {
                double aj[NT], cj[NT], dEj[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    aj[t] = alpha[(size_t)j * NT + t];
                    cj[t] = conc[(size_t)j * NT + t];
                    dEj[t] = dE[(size_t)j * NT + t];
                }
                const double r = sqrt(rsq);
                const double r_inv = 1.0 / r;
                // all ordered type pairs: rho {val, dr, dalpha}; phi for it >= jt
                double gr[NT][NT][2], fr[NPHI][2];
#pragma unroll
                for (int it = 0; it < NT; ++it) {
#pragma unroll
                    for (int jt = 0; jt < NT; ++jt) {
                        const bool want_phi = it >= jt;
                        const double a = (ai[it] * aj[jt]) / (ai[it] + aj[jt]);
                        double g[3] = {0, 0, 0}, f[3] = {0, 0, 0};
                        if (P.alpha_min < a && a < P.alpha_max) {
                            Q gq[3], fq[3];
                            const Q* tab = quad + (size_t)(it * NT + jt) * P.nr * kQuadRow;
                            bool hit;
                            if (want_phi)
                                hit = gh_average<Q, NQ, true, true>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, gq, fq);
                            else
                                hit = gh_average<Q, NQ, true, false>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, gq, fq);
                            if (hit) {
                                g[1] = (double)gq[1];
                                g[2] = (double)gq[2];
                                if (want_phi) {
                                    f[1] = (double)fq[1];
                                    f[2] = (double)fq[2];
                                }
                            }
                        } else if (P.alpha_max <= a && rsq < P.cut_sq) {
                            classical_pair(P, rsq, r_inv, rho7 + (size_t)type2rho(NT, it, jt) * P.nr * 7,
                                           phi7 + (size_t)type2phi(it, jt) * P.nr * 7, want_phi, g, f);
                        }
                        gr[it][jt][0] = g[1];
                        gr[it][jt][1] = g[2];
                        if (want_phi) {
                            fr[type2phi(it, jt)][0] = f[1];
                            fr[type2phi(it, jt)][1] = f[2];
                        }
                    }
                }
                // src/ff_dmd_I.cpp:374-410 (cache slot type2rho[jt][it] pairs with dE_i)
#pragma unroll
                for (int it = 0; it < NT; ++it) {
#pragma unroll
                    for (int jt = 0; jt < NT; ++jt) {
                        const int ph = type2phi(it, jt);
                        const double cc = ci[it] * cj[jt];
                        const double fpair =
                            -(gr[jt][it][0] * dEi[it] + gr[it][jt][0] * dEj[jt] + fr[ph][0]) * cc * r_inv;
                        const double s = ai[it] + aj[jt];
                        const double apair =
                            -(gr[jt][it][1] * dEi[it] + gr[it][jt][1] * dEj[jt] + fr[ph][1]) * cc / (s * s);
                        if (apair != 0.0 || fpair != 0.0) {
                            fi[0] += dx0 * fpair;
                            fi[1] += dx1 * fpair;
                            fi[2] += dx2 * fpair;
                            fi[3 + it] += apair * aj[jt] * aj[jt];
                            const double w = owned ? 1.0 : 0.5;
                            if (owned) {
                                pf[0] -= dx0 * fpair;
                                pf[1] -= dx1 * fpair;
                                pf[2] -= dx2 * fpair;
                                pf[3 + jt] += apair * ai[it] * ai[it];
                            }
                            const double wf = w * fpair;
                            vir[0] -= wf * dx0 * dx0;
                            vir[1] -= wf * dx1 * dx1;
                            vir[2] -= wf * dx2 * dx2;
                            vir[3] -= wf * dx1 * dx2;
                            vir[4] -= wf * dx2 * dx0;
                            vir[5] -= wf * dx0 * dx1;
                        }
                    }
                }
            }
            if (owned) {
#pragma unroll
                for (int k = 0; k < XD; ++k) pairF[(size_t)p * XD + k] = pf[k];
            }
        }
#pragma unroll
        for (int k = 0; k < XD; ++k) fown[(size_t)i * XD + k] = fi[k];
    }
    block_sum_store<6>(vir, part);
}

// ---------------------------------------------------------------- force gather
template <int NT>
__global__ __launch_bounds__(kBlock) void k_fgather(int natms, PairBufs L, const double* __restrict__ fown,
                                                    const double* __restrict__ pairF,
                                                    const double* __restrict__ fa, double* __restrict__ f) {
    constexpr int XD = 3 + NT;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= natms) return;
    double acc[XD];
#pragma unroll
    for (int k = 0; k < XD; ++k) acc[k] = fown[(size_t)i * XD + k];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[3 + t] += fa[(size_t)i * NT + t];
    const int q1 = L.roff[i + 1];
    for (int q = L.roff[i]; q < q1; ++q) {
        const size_t p = (size_t)L.rev[q];
#pragma unroll
        for (int k = 0; k < XD; ++k) acc[k] += pairF[p * XD + k];
    }
#pragma unroll
    for (int k = 0; k < XD; ++k) f[(size_t)i * XD + k] = acc[k];
}

// sums block partials in block order; out[k] for k < nv.  One block.
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ part, int nblocks, int nv,
                                                            int stride, int offset, double* __restrict__ out) {
    __shared__ double sh[kBlock];
    for (int k = 0; k < nv; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) s += part[(size_t)b * stride + offset + k];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = sh[0];
        __syncthreads();
    }
}

}  // namespace mappdmd

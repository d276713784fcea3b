// Concentration-rate kernels of the DMD force field, gather form:
//   create_2nd_neigh_lst   src/ff_dmd_I.cpp:2138-2217
//   c_d_calc  DMD_I / DMD_II / DMD   src/ff_dmd_I.cpp:2241-2398, src/ff_dmd_II.cpp:2186-2337, src/ff_dmd.cpp:2241-2398
//   calc_g    DMD_I / DMD_II / DMD   src/ff_dmd_I.cpp:2402-2674, src/ff_dmd_II.cpp:2341-2630, src/ff_dmd.cpp:2402-2671
// The reference sweeps the half list and scatters into both ends.  Each
// kernel here is one lane per owned site that adds (a) its own half-list
// rows and (b) the rows of lower-index sites that list it, found through the
// reverse CSR, recomputing the pair term with the same operands in the same
// roles, so every pair term is bitwise the reference's and only the order of
// the per-site sums differs.  The index quirk of the chemical-potential sweep
// (dE[jatm+jtype], dE[iatm+itype] instead of [atom*nt+type]) is kept.
//
// The pair values rho_bar / phi_bar come from a per-pair cache (as the
// reference's caches filled by force_calc): k_pair_values fills it for the
// current positions with the same quadrature as force_calc.
#pragma once
#include "force.hpp"

#pragma clang fp contract(fast)

namespace mappdmd {

struct RevBufs {
    const int* off;    // half list (or 2nd list) CSR
    const int* nl;
    const int* roff;   // reverse CSR over owned partners
    const int* rev;    // pair index
    const int* pair_i; // owner row of each pair
};

// pair_i[p] = i for p in the row of i
__global__ void k_pair_owner(const int* __restrict__ off, int natms, int* __restrict__ pair_i) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms) return;
    for (int p = off[i]; p < off[i + 1]; ++p) pair_i[p] = i;
}

// rho_bar for every ordered type pair and phi_bar for it >= jt, per half-list
// pair; zeros where the reference computes nothing (outside the cut, alpha out
// of range, or the rule beyond rc)
template <int NT, typename Q, int NQ>
__global__ __launch_bounds__(kBlock) void k_pair_values(FFParams<double> P, FFParams<Q> PQ,
                                                        const double4* __restrict__ pos,
                                                        const double* __restrict__ alpha, const int* __restrict__ off,
                                                        const int* __restrict__ nl, const Q* __restrict__ quad,
                                                        const double* __restrict__ rho7,
                                                        const double* __restrict__ phi7, double* __restrict__ rhoc,
                                                        double* __restrict__ phic) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
    const double4 xi = pos[i];
    double ai[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) ai[t] = alpha[(size_t)i * NT + t];
    for (int p = off[i]; p < off[i + 1]; ++p) {
        const int j = nl[p];
        const double4 xj = pos[j];
        const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
        const double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
        double rv[NT * NT], pv[NPHI];
#pragma unroll
        for (int k = 0; k < NT * NT; ++k) rv[k] = 0.0;
#pragma unroll
        for (int k = 0; k < NPHI; ++k) pv[k] = 0.0;
        if (rsq < P.cut_sq_mod) {
            const double r = sqrt(rsq), r_inv = 1.0 / r;
#pragma unroll
            for (int it = 0; it < NT; ++it)
#pragma unroll
                for (int jt = 0; jt < NT; ++jt) {
                    const bool want_phi = it >= jt;
                    const double aj = alpha[(size_t)j * NT + jt];
                    const double a = (ai[it] * aj) / (ai[it] + aj);
                    if (P.alpha_min < a && a < P.alpha_max) {
                        Q g[3], f[3];
                        const Q* tab = quad + (size_t)(it * NT + jt) * P.nr * kQuadRow;
                        bool hit = want_phi ? gh_average<Q, NQ, false, true>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, g, f)
                                            : gh_average<Q, NQ, false, false>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, g, f);
                        if (hit) {
                            rv[type2rho(NT, it, jt)] = (double)g[0];
                            if (want_phi) pv[type2phi(it, jt)] = (double)f[0];
                        }
                    } else if (P.alpha_max <= a && rsq < P.cut_sq) {
                        double g[3], f[3];
                        classical_pair(P, rsq, r_inv, rho7 + (size_t)type2rho(NT, it, jt) * P.nr * 7,
                                       phi7 + (size_t)type2phi(it, jt) * P.nr * 7, want_phi, g, f);
                        rv[type2rho(NT, it, jt)] = g[0];
                        if (want_phi) pv[type2phi(it, jt)] = f[0];
                    }
                }
        }
#pragma unroll
        for (int k = 0; k < NT * NT; ++k) rhoc[(size_t)p * NT * NT + k] = rv[k];
#pragma unroll
        for (int k = 0; k < NPHI; ++k) phic[(size_t)p * NPHI + k] = pv[k];
    }
}

// E = sum c rho_bar from the cache; n = E = F(E), dE = F', ddE = F''
// (first sweep of c_d_calc, src/ff_dmd_I.cpp:2269-2312, and calc_g, :2447-2493)
template <int NT>
__global__ __launch_bounds__(kBlock) void k_density_cache(FFParams<double> P, const double* __restrict__ conc,
                                                          RevBufs L, const double* __restrict__ rhoc,
                                                          const double* __restrict__ F9, double* __restrict__ E,
                                                          double* __restrict__ n, double* __restrict__ dE,
                                                          double* __restrict__ ddE) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
    double e[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) e[t] = 0.0;
    // rows of lower-index sites listing i: E[i][jt] += c[I][it] * rho[slot(it,jt)]
    for (int q = L.roff[i]; q < L.roff[i + 1]; ++q) {
        const int p = L.rev[q], I = L.pair_i[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt)
                e[jt] += conc[(size_t)I * NT + it] * rhoc[(size_t)p * NT * NT + type2rho(NT, it, jt)];
    }
    // own row: E[i][it] += c[j][jt] * rho[slot(jt,it)]
    for (int p = L.off[i]; p < L.off[i + 1]; ++p) {
        const int j = L.nl[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt)
                e[it] += conc[(size_t)j * NT + jt] * rhoc[(size_t)p * NT * NT + type2rho(NT, jt, it)];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double Fv, dF, ddF;
        embed_F(P, F9, t, e[t], &Fv, &dF, &ddF);
        E[(size_t)i * NT + t] = Fv;
        n[(size_t)i * NT + t] = Fv;
        dE[(size_t)i * NT + t] = dF;
        if (ddE) ddE[(size_t)i * NT + t] = ddF;
    }
}

// n += sum c (rho dE[flat jatm+jtype] + phi), then the closing terms:
// CLOSE 0: c_d_calc DMD / DMD_I   n = (n + c_0 + 1.5 kT log a) * beta
// CLOSE 1: c_d_calc DMD_II        n =  n + c_0 + 1.5 kT log a
// CLOSE 2: calc_g   DMD / DMD_I   n = (n + c_0) * beta + 1.5 log a
// CLOSE 3: calc_g   DMD_II        n = (n + c_0) + 1.5 kT log a
template <int NT, int CLOSE>
__global__ __launch_bounds__(kBlock) void k_chem_pot(FFParams<double> P, const double* __restrict__ alpha,
                                                     const double* __restrict__ conc, RevBufs L,
                                                     const double* __restrict__ rhoc,
                                                     const double* __restrict__ phic, const double* __restrict__ dE,
                                                     const int* __restrict__ qidx, double* __restrict__ n) {
    // qidx[j*NT + t] is the element that the reference's dE[jatm + jtype] reads for
    // site j (jatm + t within the site's own cell numbering: j + t for one cell)
    constexpr int NPHI = NT * (NT + 1) / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = n[(size_t)i * NT + t];
    for (int q = L.roff[i]; q < L.roff[i + 1]; ++q) {
        const int p = L.rev[q], I = L.pair_i[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt)
                acc[jt] += conc[(size_t)I * NT + it] * (rhoc[(size_t)p * NT * NT + type2rho(NT, jt, it)] *
                                                            dE[qidx[(size_t)I * NT + it]] +
                                                        phic[(size_t)p * NPHI + type2phi(it, jt)]);
    }
    for (int p = L.off[i]; p < L.off[i + 1]; ++p) {
        const int j = L.nl[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt)
                acc[it] += conc[(size_t)j * NT + jt] * (rhoc[(size_t)p * NT * NT + type2rho(NT, it, jt)] *
                                                            dE[qidx[(size_t)j * NT + jt]] +
                                                        phic[(size_t)p * NPHI + type2phi(it, jt)]);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double la = log(alpha[(size_t)i * NT + t]);
        double v = acc[t];
        if (CLOSE == 0) {
            v += P.c0[t] + 1.5 * P.kbT * la;
            v *= P.beta;
        } else if (CLOSE == 1) {
            v += P.c0[t] + 1.5 * P.kbT * la;
        } else if (CLOSE == 2) {
            v += P.c0[t];
            v *= P.beta;
            v += 1.5 * la;
        } else {
            v += P.c0[t];
            v += 1.5 * P.kbT * la;
        }
        n[(size_t)i * NT + t] = v;
    }
}

// exchange term of a 2nd-list pair (i, j) for type t, in the reference's roles
// VAR 0 = DMD, 1 = DMD_I, 2 = DMD_II.  Returns the c_d_calc s_ij.
template <int VAR>
__device__ __forceinline__ double rate_sij(const FFParams<double>& P, double ni, double nj, double ci, double cj) {
    if (VAR == 2) {
        const double mt = ni > nj ? ni + 0.5 : nj + 0.5;
        const double efi = exp(P.beta * (ni - mt));
        const double efj = exp(P.beta * (nj - mt));
        const double tmpi = ci * (1.0 - cj);
        const double tmpj = cj * (1.0 - ci);
        return tmpj * efj - tmpi * efi;
    }
    const double Mat = VAR == 0 ? 5.2 * P.beta : 0.4 * P.beta;
    double t0, t1;
    if (VAR == 1) {
        t0 = exp(0.5 * (ni - nj) + Mat);
        t1 = exp(0.5 * (nj - ni) + Mat);
    } else {
        t0 = exp(ni + Mat);
        t1 = exp(nj + Mat);
    }
    return ((1.0 - ci) * cj * t1 - (1.0 - cj) * ci * t0);
}

// c_d over the 2nd list
template <int NT, int VAR>
__global__ __launch_bounds__(kBlock) void k_c_d(FFParams<double> P, const double* __restrict__ conc, RevBufs L2,
                                                const double* __restrict__ n, double* __restrict__ c_d) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double ni = n[(size_t)i * NT + t], ci = conc[(size_t)i * NT + t];
        double acc = 0.0;
        for (int p = L2.off[i]; p < L2.off[i + 1]; ++p) {
            const int j = L2.nl[p];
            acc += rate_sij<VAR>(P, ni, n[(size_t)j * NT + t], ci, conc[(size_t)j * NT + t]);
        }
        for (int q = L2.roff[i]; q < L2.roff[i + 1]; ++q) {
            const int I = L2.pair_i[L2.rev[q]];
            acc -= rate_sij<VAR>(P, n[(size_t)I * NT + t], ni, conc[(size_t)I * NT + t], ci);
        }
        c_d[(size_t)i * NT + t] = acc;
    }
}

// calc_g residual s = 2(c + a) + rate terms (and c_d for DMD_II); partial sums of s^2
template <int NT, int VAR>
__global__ __launch_bounds__(kBlock) void k_calc_g_s(FFParams<double> P, double alpha_t,
                                                     const double* __restrict__ conc, const double* __restrict__ a,
                                                     RevBufs L2, const double* __restrict__ n, double* __restrict__ s,
                                                     double* __restrict__ c_d, double* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc2[1] = {0.0};
    if (i < P.natms) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const size_t k = (size_t)i * NT + t;
            const double ni = n[k], ci = conc[k];
            double sv = 2.0 * (ci + a[k]);
            double cd = 0.0;
            for (int p = L2.off[i]; p < L2.off[i + 1]; ++p) {
                const int j = L2.nl[p];
                const double sij = rate_sij<VAR>(P, ni, n[(size_t)j * NT + t], ci, conc[(size_t)j * NT + t]);
                if (VAR == 2) {
                    sv -= 2.0 * alpha_t * sij;
                    cd += sij;
                } else {
                    sv += -2.0 * alpha_t * sij;
                }
            }
            for (int q = L2.roff[i]; q < L2.roff[i + 1]; ++q) {
                const int I = L2.pair_i[L2.rev[q]];
                const double sij = rate_sij<VAR>(P, n[(size_t)I * NT + t], ni, conc[(size_t)I * NT + t], ci);
                if (VAR == 2) {
                    sv += 2.0 * alpha_t * sij;
                    cd -= sij;
                } else {
                    sv -= -2.0 * alpha_t * sij;
                }
            }
            s[k] = sv;
            if (VAR == 2) c_d[k] = cd;
            acc2[0] += sv * sv;
        }
    }
    block_sum_store<1>(acc2, part);
}

// pair terms of the t / g sweep for the pair (I, J) of type t in the reference's
// roles; returns contributions to (t_I, g_I) and (t_J, g_J)
template <int VAR>
__device__ __forceinline__ void tg_pair(const FFParams<double>& P, double alpha_t, double nI, double nJ, double cI,
                                        double cJ, double sI, double sJ, double* tI, double* gI, double* tJ,
                                        double* gJ) {
    if (VAR == 2) {
        const double s_ij = sI - sJ;
        const double mt = nI > nJ ? nI + 0.5 : nJ + 0.5;
        const double efi = exp(P.beta * (nI - mt));
        const double efj = exp(P.beta * (nJ - mt));
        const double tmpi = cI * (1.0 - cJ);
        const double tmpj = cJ * (1.0 - cI);
        const double dij = nI > nJ ? 1.0 : 0.0, dji = nJ > nI ? 1.0 : 0.0;
        double ti = tmpi * efi + dij * (tmpj * efj - tmpi * efi);
        ti *= s_ij;
        double tj = tmpj * efj + dji * (tmpi * efi - tmpj * efj);
        tj *= s_ij;
        *tI = ti;
        *gI = -(alpha_t * ((1.0 - cJ) * efi + cJ * efj) * s_ij);
        *tJ = -tj;
        *gJ = alpha_t * ((1.0 - cI) * efj + cI * efi) * s_ij;
        return;
    }
    const double Mat = VAR == 0 ? 5.2 * P.beta : 0.4 * P.beta;
    double t0, t1;
    if (VAR == 1) {
        t0 = exp(0.5 * (nI - nJ) + Mat);
        t1 = exp(0.5 * (nJ - nI) + Mat);
        t0 *= (sI - sJ);
        t1 *= (sI - sJ);
        const double t_ij = 0.5 * (cI * (1.0 - cJ) * t0 + cJ * (1.0 - cI) * t1);
        *tI = t_ij;
        *tJ = -t_ij;
    } else {
        t0 = (sI - sJ) * exp(Mat + nI);
        t1 = (sI - sJ) * exp(Mat + nJ);
        *tI = cI * t0 * (1.0 - cJ);
        *tJ = -(cJ * t1 * (1.0 - cI));
    }
    *gI = -(alpha_t * (t0 * (1.0 - cJ) + t1 * cJ));
    *gJ = alpha_t * (t1 * (1.0 - cI) + cI * t0);
}

template <int NT, int VAR>
__global__ __launch_bounds__(kBlock) void k_calc_g_t(FFParams<double> P, double alpha_t,
                                                     const double* __restrict__ conc, RevBufs L2,
                                                     const double* __restrict__ n, const double* __restrict__ s,
                                                     double* __restrict__ t, double* __restrict__ g) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        const size_t k = (size_t)i * NT + ty;
        const double ni = n[k], ci = conc[k], si = s[k];
        double tv = 0.0, gv = -si;
        for (int p = L2.off[i]; p < L2.off[i + 1]; ++p) {
            const size_t kj = (size_t)L2.nl[p] * NT + ty;
            double tI, gI, tJ, gJ;
            tg_pair<VAR>(P, alpha_t, ni, n[kj], ci, conc[kj], si, s[kj], &tI, &gI, &tJ, &gJ);
            tv += tI;
            gv += gI;
        }
        for (int q = L2.roff[i]; q < L2.roff[i + 1]; ++q) {
            const size_t kI = (size_t)L2.pair_i[L2.rev[q]] * NT + ty;
            double tI, gI, tJ, gJ;
            tg_pair<VAR>(P, alpha_t, n[kI], ni, conc[kI], ci, s[kI], si, &tI, &gI, &tJ, &gJ);
            tv += tJ;
            gv += gJ;
        }
        t[k] = tv;
        g[k] = gv;
    }
}

// v and the first g sweep over the half list (src/ff_dmd_I.cpp:2610-2642)
template <int NT>
__global__ __launch_bounds__(kBlock) void k_calc_g_v(FFParams<double> P, double alpha_t, RevBufs L,
                                                     const double* __restrict__ rhoc,
                                                     const double* __restrict__ phic, const double* __restrict__ dE,
                                                     const double* __restrict__ t, double* __restrict__ v,
                                                     double* __restrict__ g) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
    const double ba = P.beta * alpha_t;
    double vv[NT], gg[NT];
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        vv[ty] = 0.0;
        gg[ty] = g[(size_t)i * NT + ty];
    }
    for (int p = L.off[i]; p < L.off[i + 1]; ++p) {
        const int j = L.nl[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt) {
                const double rji = rhoc[(size_t)p * NT * NT + type2rho(NT, jt, it)];
                const double rij = rhoc[(size_t)p * NT * NT + type2rho(NT, it, jt)];
                double tmp0 = phic[(size_t)p * NPHI + type2phi(jt, it)] + rji * dE[(size_t)i * NT + it] +
                              rij * dE[(size_t)j * NT + jt];
                tmp0 *= ba;
                vv[it] += rji * t[(size_t)j * NT + jt];
                gg[it] += tmp0 * t[(size_t)j * NT + jt];
            }
    }
    for (int q = L.roff[i]; q < L.roff[i + 1]; ++q) {
        const int p = L.rev[q], I = L.pair_i[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt) {
                const double rji = rhoc[(size_t)p * NT * NT + type2rho(NT, jt, it)];
                const double rij = rhoc[(size_t)p * NT * NT + type2rho(NT, it, jt)];
                double tmp0 = phic[(size_t)p * NPHI + type2phi(jt, it)] + rji * dE[(size_t)I * NT + it] +
                              rij * dE[(size_t)i * NT + jt];
                tmp0 *= ba;
                vv[jt] += rij * t[(size_t)I * NT + it];
                gg[jt] += tmp0 * t[(size_t)I * NT + it];
            }
    }
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        v[(size_t)i * NT + ty] = vv[ty];
        g[(size_t)i * NT + ty] = gg[ty];
    }
}

// last g sweep: alpha beta rho ddE c v (src/ff_dmd_I.cpp:2647-2671)
template <int NT>
__global__ __launch_bounds__(kBlock) void k_calc_g_w(FFParams<double> P, double alpha_t,
                                                     const double* __restrict__ conc, RevBufs L,
                                                     const double* __restrict__ rhoc, const double* __restrict__ ddE,
                                                     const double* __restrict__ v, double* __restrict__ g) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.natms) return;
    const double ab = alpha_t * P.beta;
    double gg[NT];
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) gg[ty] = g[(size_t)i * NT + ty];
    for (int p = L.off[i]; p < L.off[i + 1]; ++p) {
        const int j = L.nl[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt) {
                const size_t kj = (size_t)j * NT + jt;
                gg[it] += ab * rhoc[(size_t)p * NT * NT + type2rho(NT, it, jt)] * ddE[kj] * conc[kj] * v[kj];
            }
    }
    for (int q = L.roff[i]; q < L.roff[i + 1]; ++q) {
        const int p = L.rev[q], I = L.pair_i[p];
#pragma unroll
        for (int it = 0; it < NT; ++it)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt) {
                const size_t kI = (size_t)I * NT + it;
                gg[jt] += ab * rhoc[(size_t)p * NT * NT + type2rho(NT, jt, it)] * ddE[kI] * conc[kI] * v[kI];
            }
    }
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) g[(size_t)i * NT + ty] = gg[ty];
}

}  // namespace mappdmd

#pragma clang fp contract(off)

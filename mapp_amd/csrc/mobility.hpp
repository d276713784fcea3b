// Concentration-rate kernels of the DMD force field, gather form:
//   create_2nd_neigh_lst   src/ff_dmd_I.cpp:2138-2217
//   c_d_calc  DMD_I / DMD_II / DMD   src/ff_dmd_I.cpp:2241-2398, src/ff_dmd_II.cpp:2186-2337, src/ff_dmd.cpp:2241-2398
//   calc_g    DMD_I / DMD_II / DMD   src/ff_dmd_I.cpp:2402-2674, src/ff_dmd_II.cpp:2341-2630, src/ff_dmd.cpp:2402-2671
// The reference sweeps the half list and scatters every pair term into both
// ends.  Here each site gathers the terms of all its pairs and never writes
// another site's row, so there are no atomics and every sum has a fixed
// order.
//
// Pair values.  rho_bar (every ordered type pair) and phi_bar depend on
// positions and widths only; the reference caches them in force_calc.  Here
// k_pair_values evaluates them per half-list pair with the same quadrature as
// force_calc, and k_table_fill turns them into the site table: for each
// owned site, the entries of its own half-list row, then the entries of the
// lower-index sites that list it (ascending owner), each holding the partner
// index and the pair's nt^2 + nt(nt+1)/2 values in the owner's orientation.
// Pairs whose values are all zero (outside the cut, or the width pair out of
// range) add nothing to any sum and are left out.  Seen from site i with
// type m and a partner k of type b, each sweep's term is the same expression
// whichever end owns the pair (slot(b, m) for the density, slot(m, b) for the
// chemical potential and the last g sweep), so one loop covers both.
//
// Sweeps run G = 16 lanes per site: the lanes stride through the site's
// entries (coalesced: consecutive lanes read consecutive entries), then a
// fixed DPP rotation tree sums the lane partials.  The 2nd list (rates) runs
// 4 lanes per site over its own table, where the entry's sign tells which end
// owns the pair, since the rate terms are antisymmetric in the two ends.
// The index quirk of the chemical-potential sweep (dE[jatm+jtype],
// dE[iatm+itype] instead of [atom*nt+type]) is kept.
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
// pair (one thread per pair, owner from pair_i); zeros where the reference
// computes nothing (outside the cut, alpha out of range, or the rule beyond rc)
template <int NT, typename Q, int NQ>
__global__ __launch_bounds__(kBlock) void k_pair_values(FFParams<double> P, FFParams<Q> PQ,
                                                        const double4* __restrict__ pos,
                                                        const double* __restrict__ alpha, const int* __restrict__ nl,
                                                        const int* __restrict__ pair_i, long npairs,
                                                        const Q* __restrict__ quad, const double* __restrict__ rho7,
                                                        const double* __restrict__ phi7, double* __restrict__ rhoc,
                                                        double* __restrict__ phic) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    const long p = (long)blockIdx.x * kBlock + threadIdx.x;
    if (p >= npairs) return;
    const int i = pair_i[p], j = nl[p];
    const double4 xi = pos[i], xj = pos[j];
    const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
    double rsq;
    {
#pragma clang fp contract(off)
        rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;  // as force_calc
    }
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
                const double ai = alpha[(size_t)i * NT + it], aj = alpha[(size_t)j * NT + jt];
                const double a = (ai * aj) / (ai + aj);
                if (P.alpha_min < a && a < P.alpha_max) {
                    Q g[3], f[3];
                    const Q* tab = quad + (size_t)(it * NT + jt) * kQuadRow;
                    bool hit = want_phi ? gh_average<Q, NQ, false, true>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, quad_stride(NT), g, f)
                                        : gh_average<Q, NQ, false, false>(PQ, (Q)r, (Q)r_inv, (Q)a, tab, quad_stride(NT), g, f);
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

// ---------------------------------------------------------------- site table
constexpr int kTableG = 16;  // lanes per site over the pair table
constexpr int kList2G = 4;   // lanes per site over the 2nd list

struct SiteTable {
    const int* off;     // [natms + 1] entries of site i: off[i] .. off[i+1]
    const int* idx;     // partner
    const double* rho;  // [entry][nt*nt] rho_bar, slot type2rho of the owner's (it, jt)
    const double* phi;  // [entry][nt(nt+1)/2] phi_bar
};

template <int NT>
__device__ __forceinline__ bool pair_contributes(const double* __restrict__ rhoc, const double* __restrict__ phic,
                                                 long p) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    bool nz = false;
#pragma unroll
    for (int k = 0; k < NT * NT; ++k) nz |= rhoc[p * NT * NT + k] != 0.0;
#pragma unroll
    for (int k = 0; k < NPHI; ++k) nz |= phic[p * NPHI + k] != 0.0;
    return nz;
}

// entries per site: contributing pairs of the own row and of the reverse CSR
// (16 lanes per site, strided over both ranges)
template <int NT>
__global__ __launch_bounds__(kBlock) void k_table_count(RevBufs L, int natms, const double* __restrict__ rhoc,
                                                        const double* __restrict__ phic, int* __restrict__ cnt) {
    const SiteGroup<kTableG> g;
    const int i = g.site;
    if (i >= natms) return;
    int c = 0;
    for (int p = L.off[i] + g.l; p < L.off[i + 1]; p += kTableG) c += pair_contributes<NT>(rhoc, phic, p) ? 1 : 0;
    for (int q = L.roff[i] + g.l; q < L.roff[i + 1]; q += kTableG) c += pair_contributes<NT>(rhoc, phic, L.rev[q]) ? 1 : 0;
    const double s = group_sum<kTableG>((double)c);
    if (g.l == 0) cnt[i] = (int)s;
}

// rank of this lane's flag among the set flags of lower lanes of its 16-lane group,
// and the number of set flags in the group
__device__ __forceinline__ int group_rank16(bool flag, int l, int* total) {
    const unsigned long long m = __builtin_amdgcn_ballot_w64(flag);
    const int gbase = (threadIdx.x & 63) & ~15;
    const unsigned bits = (unsigned)((m >> gbase) & 0xFFFFull);
    *total = __popc(bits);
    return __popc(bits & ((1u << l) - 1u));
}

// the entries, in list order, own row first; 16 lanes per site, each stride of
// 16 entries placed by ballot rank
template <int NT>
__global__ __launch_bounds__(kBlock) void k_table_fill(RevBufs L, int natms, const double* __restrict__ rhoc,
                                                       const double* __restrict__ phic, const int* __restrict__ toff,
                                                       int* __restrict__ tidx, double* __restrict__ trho,
                                                       double* __restrict__ tphi) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    const SiteGroup<kTableG> g;
    const int i = g.site;
    if (i >= natms) return;
    long e = toff[i];
    auto put = [&](long dst, long p, int partner) {
        tidx[dst] = partner;
#pragma unroll
        for (int k = 0; k < NT * NT; ++k) trho[dst * NT * NT + k] = rhoc[p * NT * NT + k];
#pragma unroll
        for (int k = 0; k < NPHI; ++k) tphi[dst * NPHI + k] = phic[p * NPHI + k];
    };
    const int o0 = L.off[i], o1 = L.off[i + 1];
    for (int b = o0; b < o1; b += kTableG) {
        const int p = b + g.l;
        const bool in = p < o1 && pair_contributes<NT>(rhoc, phic, p);
        int tot;
        const int rk = group_rank16(in, g.l, &tot);
        if (in) put(e + rk, p, L.nl[p]);
        e += tot;
    }
    const int q0 = L.roff[i], q1 = L.roff[i + 1];
    for (int b = q0; b < q1; b += kTableG) {
        const int q = b + g.l;
        const int p = q < q1 ? L.rev[q] : 0;
        const bool in = q < q1 && pair_contributes<NT>(rhoc, phic, p);
        int tot;
        const int rk = group_rank16(in, g.l, &tot);
        if (in) put(e + rk, p, L.pair_i[p]);
        e += tot;
    }
}

// 2nd list of both ends: own rows as partner j, reverse entries as ~owner
__global__ void k_list2_count(RevBufs L2, int natms, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms) return;
    cnt[i] = (L2.off[i + 1] - L2.off[i]) + (L2.roff[i + 1] - L2.roff[i]);
}
__global__ void k_list2_fill(RevBufs L2, int natms, const int* __restrict__ foff, int* __restrict__ fidx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms) return;
    int e = foff[i];
    for (int p = L2.off[i]; p < L2.off[i + 1]; ++p) fidx[e++] = L2.nl[p];
    for (int q = L2.roff[i]; q < L2.roff[i + 1]; ++q) fidx[e++] = ~L2.pair_i[L2.rev[q]];
}


// E = sum c rho_bar; n = E = F(E), dE = F', ddE = F''
// (first sweep of c_d_calc, src/ff_dmd_I.cpp:2269-2312, and calc_g, :2447-2493).
// Site i of type m, partner k of type b: E[i][m] += c[k][b] rho[slot(b, m)]
// (own row :2285-2286 and the partner end :2287, with the reference's slot).
template <int NT>
__global__ __launch_bounds__(kBlock) void k_density_table(FFParams<double> P, const double* __restrict__ conc,
                                                          SiteTable T, const double* __restrict__ F9,
                                                          double* __restrict__ E, double* __restrict__ n,
                                                          double* __restrict__ dE, double* __restrict__ ddE) {
    const SiteGroup<kTableG> g;
    const int i = g.site;
    if (i >= P.natms) return;
    double e[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) e[t] = 0.0;
    for (int q = T.off[i] + g.l; q < T.off[i + 1]; q += kTableG) {
        const int k = T.idx[q];
        double r[NT * NT], c[NT];
#pragma unroll
        for (int x = 0; x < NT * NT; ++x) r[x] = T.rho[(size_t)q * NT * NT + x];
#pragma unroll
        for (int b = 0; b < NT; ++b) c[b] = conc[(size_t)k * NT + b];
#pragma unroll
        for (int m = 0; m < NT; ++m)
#pragma unroll
            for (int b = 0; b < NT; ++b) e[m] += c[b] * r[type2rho(NT, b, m)];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) e[t] = group_sum<kTableG>(e[t]);
    if (g.l < NT) {
        double et = e[0];
#pragma unroll
        for (int t = 1; t < NT; ++t)
            if (g.l == t) et = e[t];
        double Fv, dF, ddF;
        embed_F(P, F9, g.l, et, &Fv, &dF, &ddF);
        const size_t k = (size_t)i * NT + g.l;
        E[k] = Fv;
        n[k] = Fv;
        dE[k] = dF;
        if (ddE) ddE[k] = ddF;
    }
}

// n += sum c (rho dE[flat jatm+jtype] + phi), then the closing terms:
// CLOSE 0: c_d_calc DMD / DMD_I   n = (n + c_0 + 1.5 kT log a) * beta
// CLOSE 1: c_d_calc DMD_II        n =  n + c_0 + 1.5 kT log a
// CLOSE 2: calc_g   DMD / DMD_I   n = (n + c_0) * beta + 1.5 log a
// CLOSE 3: calc_g   DMD_II        n = (n + c_0) + 1.5 kT log a
// Site i of type m, partner k of type b: c[k][b] (rho[slot(m, b)] dE[q(k, b)] + phi[(m, b)])
// (src/ff_dmd_I.cpp:2336-2339 at both ends).
template <int NT, int CLOSE>
__global__ __launch_bounds__(kBlock) void k_chem_pot(FFParams<double> P, const double* __restrict__ alpha,
                                                     const double* __restrict__ conc, SiteTable T,
                                                     const double* __restrict__ dE, const int* __restrict__ qidx,
                                                     double* __restrict__ n) {
    // qidx[j*NT + t] is the element that the reference's dE[jatm + jtype] reads for
    // site j (jatm + t within the site's own cell numbering: j + t for one cell)
    constexpr int NPHI = NT * (NT + 1) / 2;
    const SiteGroup<kTableG> g;
    const int i = g.site;
    if (i >= P.natms) return;
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = 0.0;
    for (int q = T.off[i] + g.l; q < T.off[i + 1]; q += kTableG) {
        const int k = T.idx[q];
        double r[NT * NT], f[NPHI], cd[NT];
#pragma unroll
        for (int x = 0; x < NT * NT; ++x) r[x] = T.rho[(size_t)q * NT * NT + x];
#pragma unroll
        for (int x = 0; x < NPHI; ++x) f[x] = T.phi[(size_t)q * NPHI + x];
#pragma unroll
        for (int b = 0; b < NT; ++b) cd[b] = dE[qidx[(size_t)k * NT + b]];
#pragma unroll
        for (int m = 0; m < NT; ++m)
#pragma unroll
            for (int b = 0; b < NT; ++b)
                acc[m] += conc[(size_t)k * NT + b] * (r[type2rho(NT, m, b)] * cd[b] + f[type2phi(m, b)]);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = group_sum<kTableG>(acc[t]);
    if (g.l < NT) {
        const int t = g.l;
        double v = acc[0];
#pragma unroll
        for (int x = 1; x < NT; ++x)
            if (t == x) v = acc[x];
        const size_t k = (size_t)i * NT + t;
        v += n[k];
        const double la = log(alpha[k]);
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
        n[k] = v;
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

// c_d over the 2nd list; an entry e >= 0 is the own-row partner j (pair (i, j)),
// e < 0 the owner ~e of the pair (~e, i)
template <int NT, int VAR>
__global__ __launch_bounds__(kBlock) void k_c_d(FFParams<double> P, const double* __restrict__ conc,
                                                const int* __restrict__ foff, const int* __restrict__ fidx,
                                                const double* __restrict__ n, double* __restrict__ c_d) {
    const SiteGroup<kList2G> g;
    const int i = g.site;
    if (i >= P.natms) return;
    double acc[NT], ni[NT], ci[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        acc[t] = 0.0;
        ni[t] = n[(size_t)i * NT + t];
        ci[t] = conc[(size_t)i * NT + t];
    }
    for (int q = foff[i] + g.l; q < foff[i + 1]; q += kList2G) {
        const int e = fidx[q];
        const bool own = e >= 0;
        const int k = own ? e : ~e;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double nk = n[(size_t)k * NT + t], ck = conc[(size_t)k * NT + t];
            const double sij = own ? rate_sij<VAR>(P, ni[t], nk, ci[t], ck) : rate_sij<VAR>(P, nk, ni[t], ck, ci[t]);
            acc[t] += own ? sij : -sij;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = group_sum<kList2G>(acc[t]);
    if (g.l == 0) {
#pragma unroll
        for (int t = 0; t < NT; ++t) c_d[(size_t)i * NT + t] = acc[t];
    }
}

// calc_g residual s = 2(c + a) - 2 alpha (rate sums) (and c_d for DMD_II); partial sums of s^2
template <int NT, int VAR>
__global__ __launch_bounds__(kBlock) void k_calc_g_s(FFParams<double> P, double alpha_t,
                                                     const double* __restrict__ conc, const double* __restrict__ a,
                                                     const int* __restrict__ foff, const int* __restrict__ fidx,
                                                     const double* __restrict__ n, double* __restrict__ s,
                                                     double* __restrict__ c_d, double* __restrict__ part) {
    const SiteGroup<kList2G> g;
    const int i = g.site;
    double sq = 0.0;
    if (i < P.natms) {
        double acc[NT], ni[NT], ci[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            acc[t] = 0.0;
            ni[t] = n[(size_t)i * NT + t];
            ci[t] = conc[(size_t)i * NT + t];
        }
        for (int q = foff[i] + g.l; q < foff[i + 1]; q += kList2G) {
            const int e = fidx[q];
            const bool own = e >= 0;
            const int k = own ? e : ~e;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double nk = n[(size_t)k * NT + t], ck = conc[(size_t)k * NT + t];
                const double sij =
                    own ? rate_sij<VAR>(P, ni[t], nk, ci[t], ck) : rate_sij<VAR>(P, nk, ni[t], ck, ci[t]);
                acc[t] += own ? sij : -sij;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = group_sum<kList2G>(acc[t]);
        if (g.l == 0) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const size_t k = (size_t)i * NT + t;
                const double sv = 2.0 * (ci[t] + a[k]) - 2.0 * alpha_t * acc[t];
                s[k] = sv;
                if (VAR == 2) c_d[k] = acc[t];
                sq += sv * sv;
            }
        }
    }
    double v[1] = {sq};
    block_sum_store<1>(v, part);
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

// t and the first g term over the 2nd list (src/ff_dmd_I.cpp:2576-2606)
template <int NT, int VAR>
__global__ __launch_bounds__(kBlock) void k_calc_g_t(FFParams<double> P, double alpha_t,
                                                     const double* __restrict__ conc, const int* __restrict__ foff,
                                                     const int* __restrict__ fidx, const double* __restrict__ n,
                                                     const double* __restrict__ s, double* __restrict__ t,
                                                     double* __restrict__ g) {
    const SiteGroup<kList2G> grp;
    const int i = grp.site;
    if (i >= P.natms) return;
    double tv[NT], gv[NT], ni[NT], ci[NT], si[NT];
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        const size_t k = (size_t)i * NT + ty;
        tv[ty] = 0.0;
        gv[ty] = 0.0;
        ni[ty] = n[k];
        ci[ty] = conc[k];
        si[ty] = s[k];
    }
    for (int q = foff[i] + grp.l; q < foff[i + 1]; q += kList2G) {
        const int e = fidx[q];
        const bool own = e >= 0;
        const int k = own ? e : ~e;
#pragma unroll
        for (int ty = 0; ty < NT; ++ty) {
            const size_t kk = (size_t)k * NT + ty;
            const double nk = n[kk], ck = conc[kk], sk = s[kk];
            double tI, gI, tJ, gJ;
            if (own) {
                tg_pair<VAR>(P, alpha_t, ni[ty], nk, ci[ty], ck, si[ty], sk, &tI, &gI, &tJ, &gJ);
                tv[ty] += tI;
                gv[ty] += gI;
            } else {
                tg_pair<VAR>(P, alpha_t, nk, ni[ty], ck, ci[ty], sk, si[ty], &tI, &gI, &tJ, &gJ);
                tv[ty] += tJ;
                gv[ty] += gJ;
            }
        }
    }
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        tv[ty] = group_sum<kList2G>(tv[ty]);
        gv[ty] = group_sum<kList2G>(gv[ty]);
    }
    if (grp.l == 0) {
#pragma unroll
        for (int ty = 0; ty < NT; ++ty) {
            const size_t k = (size_t)i * NT + ty;
            t[k] = tv[ty];
            g[k] = gv[ty] - si[ty];
        }
    }
}

// v and the next g term over the pair table (src/ff_dmd_I.cpp:2610-2642).
// Site i of type m, partner k of type b:
//   v[i][m] += rho[slot(b, m)] t[k][b]
//   g[i][m] += alpha beta (phi[(m, b)] + rho[slot(b, m)] dE[i][m] + rho[slot(m, b)] dE[k][b]) t[k][b]
template <int NT>
__global__ __launch_bounds__(kBlock) void k_calc_g_v(FFParams<double> P, double alpha_t, SiteTable T,
                                                     const double* __restrict__ dE, const double* __restrict__ t,
                                                     double* __restrict__ v, double* __restrict__ g) {
    constexpr int NPHI = NT * (NT + 1) / 2;
    const SiteGroup<kTableG> grp;
    const int i = grp.site;
    if (i >= P.natms) return;
    const double ba = P.beta * alpha_t;
    double vv[NT], gg[NT], dEi[NT];
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        vv[ty] = 0.0;
        gg[ty] = 0.0;
        dEi[ty] = dE[(size_t)i * NT + ty];
    }
    for (int q = T.off[i] + grp.l; q < T.off[i + 1]; q += kTableG) {
        const int k = T.idx[q];
        double r[NT * NT], f[NPHI], tk[NT], dk[NT];
#pragma unroll
        for (int x = 0; x < NT * NT; ++x) r[x] = T.rho[(size_t)q * NT * NT + x];
#pragma unroll
        for (int x = 0; x < NPHI; ++x) f[x] = T.phi[(size_t)q * NPHI + x];
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            tk[b] = t[(size_t)k * NT + b];
            dk[b] = dE[(size_t)k * NT + b];
        }
#pragma unroll
        for (int m = 0; m < NT; ++m)
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                const double rbm = r[type2rho(NT, b, m)];
                const double tmp0 = (f[type2phi(m, b)] + rbm * dEi[m] + r[type2rho(NT, m, b)] * dk[b]) * ba;
                vv[m] += rbm * tk[b];
                gg[m] += tmp0 * tk[b];
            }
    }
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) {
        vv[ty] = group_sum<kTableG>(vv[ty]);
        gg[ty] = group_sum<kTableG>(gg[ty]);
    }
    if (grp.l == 0) {
#pragma unroll
        for (int ty = 0; ty < NT; ++ty) {
            const size_t k = (size_t)i * NT + ty;
            v[k] = vv[ty];
            g[k] += gg[ty];
        }
    }
}

// last g term: alpha beta rho ddE c v (src/ff_dmd_I.cpp:2647-2671).
// Site i of type m, partner k of type b: alpha beta rho[slot(m, b)] ddE[k][b] c[k][b] v[k][b]
template <int NT>
__global__ __launch_bounds__(kBlock) void k_calc_g_w(FFParams<double> P, double alpha_t,
                                                     const double* __restrict__ conc, SiteTable T,
                                                     const double* __restrict__ ddE, const double* __restrict__ v,
                                                     double* __restrict__ g) {
    const SiteGroup<kTableG> grp;
    const int i = grp.site;
    if (i >= P.natms) return;
    const double ab = alpha_t * P.beta;
    double gg[NT];
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) gg[ty] = 0.0;
    for (int q = T.off[i] + grp.l; q < T.off[i + 1]; q += kTableG) {
        const int k = T.idx[q];
        double r[NT * NT], w[NT];
#pragma unroll
        for (int x = 0; x < NT * NT; ++x) r[x] = T.rho[(size_t)q * NT * NT + x];
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const size_t kk = (size_t)k * NT + b;
            w[b] = ddE[kk] * conc[kk] * v[kk];
        }
#pragma unroll
        for (int m = 0; m < NT; ++m)
#pragma unroll
            for (int b = 0; b < NT; ++b) gg[m] += ab * r[type2rho(NT, m, b)] * w[b];
    }
#pragma unroll
    for (int ty = 0; ty < NT; ++ty) gg[ty] = group_sum<kTableG>(gg[ty]);
    if (grp.l == 0) {
#pragma unroll
        for (int ty = 0; ty < NT; ++ty) g[(size_t)i * NT + ty] += gg[ty];
    }
}

}  // namespace mappdmd

#pragma clang fp contract(off)

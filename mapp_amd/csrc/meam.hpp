// MEAM (ForceField_meam, src/ff_meam.cpp) on the GPU.
//
// The reference sweeps each pair once from its lower index over the full
// list, adding the partner's shares to the partner (:1143-1400) and the pair
// forces to both ends (:1552-2140).  Here every owned site gathers over its own
// full list, both directions, so nothing is scattered and no atomics are used:
//   k_meam_density  one wavefront per site: the screening S_ij of every list
//                   entry over the site's list (staged in LDS), the site's 30
//                   density sums, its record of 38 values (rho0..3, the
//                   angular sums, t averages, Gamma, the three derivative
//                   factors, F') and its embedding energy
//   (ghosts)        records forwarded to the ghosts (atoms->update, :1545)
//   k_meam_force    one wavefront per site: the pair force of every entry
//                   (both records), the pair energy and virial at half weight
//                   per direction, and the site's own three-body screening
//                   forces (the derivative of S_ij by the site's position is
//                   the same in either role of the pair)
//   k_meam_kforce   one wavefront per site k: the forces on k as a screening
//                   atom, from every pair (o, p > o) of a local neighbour o
//                   whose list holds k itself -- the reference applies these
//                   only when k is local in the pair owner's list (:2076), and
//                   o local in k's list is exactly that.
// Sums run in another order than the reference's sequential loops, so the
// results agree to rounding.
#pragma once

#include "force.hpp"
#include "meam_setup.hpp"

namespace mappdmd {

constexpr int kMeamMaxNb = 192;  // list entries per site staged in LDS
constexpr int kMeamWaves = kBlock / 64;

enum { kR0 = 0, kR1 = 1, kR2 = 2, kR3 = 3, kA1 = 4, kA2 = 7, kA2b = 13, kA3 = 14, kA3b = 24, kT = 27, kTsq = 30,
       kGam = 33, kD1 = 34, kD2 = 35, kD3 = 36, kFp = 37 };
constexpr int kMeamSums = 30;  // slots 0..29 accumulate pair shares (rho0 at 0, slots 4..32 minus 1..3)

__device__ __forceinline__ double meam_fcut(double x) {
    if (x >= 1.0) return 1.0;
    if (x <= 0.0) return 0.0;
    double a = 1.0 - x;
    a *= a;
    a *= a;
    a = 1.0 - a;
    return a * a;
}
__device__ __forceinline__ double meam_dfcut(double x, double* dfc) {
    if (x > 1.0) {
        *dfc = 0.0;
        return 1.0;
    }
    if (x < 0.0) {
        *dfc = 0.0;
        return 0.0;
    }
    const double a = 1.0 - x;
    const double a3 = a * a * a, a4 = a * a * a * a;
    *dfc = 8 * (1.0 - a4) * a3;
    return (1.0 - a4) * (1.0 - a4);
}
// the reference's dCfunc2 with its parameter order (rij2, rjk2, rki2)
__device__ __forceinline__ void meam_dc2(double rij2, double p_jk, double p_ki, double* d1, double* d2) {
    const double rij4 = rij2 * rij2, rjk4 = p_jk * p_jk, rki4 = p_ki * p_ki;
    const double a = p_ki - p_jk;
    double den = rij4 - a * a;
    den *= den;
    *d1 = 4.0 * rij2 * (rij4 + rki4 + 2.0 * p_ki * p_jk - 3.0 * rjk4 - 2.0 * rij2 * a) / den;
    *d2 = 4.0 * rij2 * (rij4 - 3.0 * rki4 + 2.0 * p_ki * p_jk + rjk4 + 2.0 * rij2 * a) / den;
}
__device__ __forceinline__ double meam_G(const MeamDev& M, double gam, int ibar, double* dG) {
    const double gs = M.gsmooth;
    double G, d;
    if (ibar == 0 || ibar == 4) {
        const double sw = -gs / (gs + 1.0);
        G = gam < sw ? sqrt(pow(sw / gam, gs) / (gs + 1.0)) : sqrt(1.0 + gam);
        d = 0.5 / G;
    } else if (ibar == 1) {
        G = exp(gam * 0.5);
        d = 0.5 * G;
    } else if (ibar == 3) {
        G = 2.0 / (1.0 + exp(-gam));
        d = G * (1.0 - 0.5 * G);
    } else {  // -5 (the host admits no other value)
        if ((1.0 + gam) >= 0.0) {
            G = sqrt(1.0 + gam);
            d = 0.5 / G;
        } else {
            G = -sqrt(-1.0 - gam);
            d = -0.5 / G;
        }
    }
    *dG = d;
    return G;
}
__device__ __forceinline__ double meam_phi(const double* __restrict__ tab, const MeamDev& M, int a, int b, double r,
                                           double* dphi) {
    double pp = r * M.dr_inv;
    int kk = (int)pp;
    kk = kk < kMeamNr - 2 ? kk : kMeamNr - 2;
    pp -= (double)kk;
    pp = pp < 1.0 ? pp : 1.0;
    const int c = a < b ? (b + 1) * b / 2 + a : (a + 1) * a / 2 + b;
    const double* q = tab + ((size_t)c * kMeamNr + kk) * 7;
    if (dphi) *dphi = (q[6] * pp + q[5]) * pp + q[4];
    return ((q[3] * pp + q[2]) * pp + q[1]) * pp + q[0];
}

// per-wave LDS: the site's list entries {x, y, z, type}, indices and squared
// distances to the site, its record
struct MeamStage {
    double4* nb;
    int* idx;
    double* r2;
    double* rec;
    __device__ __forceinline__ explicit MeamStage(int w) {
        __shared__ double4 s_nb[kMeamWaves][kMeamMaxNb];
        __shared__ int s_idx[kMeamWaves][kMeamMaxNb];
        __shared__ double s_r2[kMeamWaves][kMeamMaxNb];
        __shared__ double s_rec[kMeamWaves][kMeamRec];
        nb = s_nb[w];
        idx = s_idx[w];
        r2 = s_r2[w];
        rec = s_rec[w];
    }
};

__device__ __forceinline__ void meam_stage(const MeamStage& S, const double4* __restrict__ pos,
                                           const double* __restrict__ type, const int* __restrict__ nl, int p0,
                                           int n, int lane, double4 xi) {
    for (int q = lane; q < n; q += 64) {
        const int j = nl[p0 + q];
        const double4 x = pos[j];
        S.nb[q] = make_double4(x.x, x.y, x.z, type[j]);
        S.idx[q] = j;
        // |x_k - x_i|^2 as the reference forms it for a screening atom k (:1198-1201)
        const double xki = x.x - xi.x, yki = x.y - xi.y, zki = x.z - xi.z;
        double r2;
        {
#pragma clang fp contract(off)
            r2 = xki * xki + yki * yki + zki * zki;
        }
        S.r2[q] = r2;
    }
    __builtin_amdgcn_wave_barrier();
}

// screening candidates of one lane per round of meam_screen (bytes in LDS,
// lane-interleaved: slot c of lane l at c * 64 + l)
constexpr int kMeamCand = 32;

// screening S_ij over the site's list, with the cutoff taper (:1171-1266).
// Each round first walks the candidates k with the two cheap ellipse tests and
// keeps the list positions of those inside both (up to kMeamCand, in list
// order, in this lane's column of cand), then runs the screening arithmetic
// on those only: the lanes of a wavefront no longer execute the divisions and
// the taper for every candidate that any one of them needs.  The candidates
// reach the product and the sum in the same order as in a single walk.
__device__ __forceinline__ void meam_screen(const MeamDev& M, const MeamStage& S, int n, int j, int it, int jt,
                                            double xi0, double xi1, double xi2, double4 xj, double rij2, double rij,
                                            unsigned char* __restrict__ cand, double* fc_o, double* s_o,
                                            double* ds_o) {
    if (!(rij < M.rc)) {
        *fc_o = *s_o = *ds_o = 0.0;
        return;
    }
    const double rnorm = (M.rc - rij) * M.delr_inv;
    double sij = 1.0, dsij = 0.0, dfcij;
    const double rbound = M.ebound[it][jt] * rij2, rij4 = rij2 * rij2;
    const double fcij = meam_dfcut(rnorm, &dfcij);
    dfcij *= M.delr_inv;
    for (int kn0 = 0; kn0 < n && sij != 0.0;) {
        int nc = 0;
        for (; kn0 < n && nc < kMeamCand; ++kn0) {
            if (S.idx[kn0] == j) continue;
            // both distances must lie inside the ellipse bound; the staged one first
            if (!(S.r2[kn0] < rbound)) continue;
            const double4 xk = S.nb[kn0];
            const double xjk = xj.x - xk.x, yjk = xj.y - xk.y, zjk = xj.z - xk.z;
            double rjk2;
            {
#pragma clang fp contract(off)
                rjk2 = xjk * xjk + yjk * yjk + zjk * zjk;
            }
            if (rjk2 < rbound) cand[(nc++) * 64] = (unsigned char)kn0;
        }
        for (int c = 0; c < nc && sij != 0.0; ++c) {
            const int kn = cand[c * 64];
            const double rki2 = S.r2[kn];
            const double4 xk = S.nb[kn];
            const int kt = (int)xk.w;
            const double xjk = xj.x - xk.x, yjk = xj.y - xk.y, zjk = xj.z - xk.z;
            double rjk2;
            {
#pragma clang fp contract(off)
                rjk2 = xjk * xjk + yjk * yjk + zjk * zjk;
            }
            const double x_ki = rki2 / rij2, x_jk = rjk2 / rij2;
            double a = 1.0 - (x_ki - x_jk) * (x_ki - x_jk);
            if (!(a > 0)) continue;
            double cikj = (2.0 * (x_ki + x_jk) + a - 2.0) / a;
            const double cmax = M.cmax[it][jt][kt], cmin = M.cmin[it][jt][kt];
            if (cikj < cmax && cikj > cmin) {
                cikj = (cikj - cmin) / (cmax - cmin);
                double dfikj;
                const double sikj = meam_dfcut(cikj, &dfikj);
                sij *= sikj;
                const double coef1 = dfikj / ((cmax - cmin) * sikj);
                a = rjk2 - rki2;
                const double b = rki2 + rjk2;
                double den = rij4 - a * a;
                den *= den;
                dsij += coef1 * (-4.0 * (-2.0 * rij2 * a * a + rij4 * b + a * a * b) / den);
            } else if (cikj < cmin) {
                sij = 0.0;
                dsij = 0.0;
            }
        }
    }
    const double coef1 = sij * fcij, coef2 = sij * dfcij / rij;
    dsij *= coef1;
    dsij -= coef2;
    if (!M.fix && (sij == 1.0 || sij == 0.0)) dsij = 0.0;
    if (sij * fcij == 1.0 || sij * fcij == 0.0) dsij = 0.0;
    *fc_o = fcij;
    *s_o = sij;
    *ds_o = dsij;
}

// pass 1: screening, densities, record and embedding; ENERGY adds the pair
// energy at half weight per direction (energy_calc, :2148-2589)
template <bool ENERGY>
__global__ __launch_bounds__(kBlock) void k_meam_density(MeamDev M, const double4* __restrict__ pos,
                                                         const double* __restrict__ type,
                                                         const int* __restrict__ off, const int* __restrict__ nl,
                                                         const double* __restrict__ tab, int natms,
                                                         double4* __restrict__ pairs, double* __restrict__ rec,
                                                         double* __restrict__ part) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kMeamWaves + w;
    const MeamStage S(w);
    __shared__ unsigned char s_cand[kMeamWaves][kMeamCand * 64];
    double en = 0.0;
    if (i < natms) {
        const int p0 = off[i], n = off[i + 1] - p0;
        const double4 xi = pos[i];
        meam_stage(S, pos, type, nl, p0, n, lane, xi);
        const int it = (int)type[i];
        double acc[kMeamSums];
#pragma unroll
        for (int q = 0; q < kMeamSums; ++q) acc[q] = 0.0;
        for (int jn = lane; jn < n; jn += 64) {
            const double4 xj = S.nb[jn];
            const int j = S.idx[jn], jt = (int)xj.w;
            const double d0 = xj.x - xi.x, d1 = xj.y - xi.y, d2 = xj.z - xi.z;
            const double xij = xi.x - xj.x, yij = xi.y - xj.y, zij = xi.z - xj.z;
            double rij2;
            {
#pragma clang fp contract(off)
                rij2 = xij * xij + yij * yij + zij * zij;
            }
            const double rij = sqrt(rij2);
            double fc, s, ds;
            meam_screen(M, S, n, j, it, jt, xi.x, xi.y, xi.z, xj, rij2, rij, &s_cand[w][lane], &fc, &s, &ds);
            pairs[p0 + jn] = make_double4(fc, s, ds, 0.0);
            if (s * fc != 0.0) {
                const double sij = s * fc;
                if (ENERGY) en += 0.5 * (meam_phi(tab, M, it, jt, rij, nullptr) * sij);
                // the partner's atomic densities at r, screened (:1270-1293)
                const double aj = rij / M.re[jt] - 1.0, r0 = M.rho0[jt];
                double ra[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) ra[q] = sij * r0 * exp(-M.beta[jt][q] * aj);
                if (M.ialloy == 1) {
#pragma unroll
                    for (int q = 1; q < 4; ++q) ra[q] *= M.t[jt][q - 1];
                }
                acc[0] += ra[0];
                if (M.ialloy != 2) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) acc[kT - 3 + q] += M.t[jt][q] * ra[0];
                }
                if (M.ialloy == 1) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) acc[kTsq - 3 + q] += M.t[jt][q] * M.t[jt][q] * ra[0];
                }
                acc[kA2b - 3] += ra[2];
                const double A1 = ra[1] / rij, B1 = ra[3] / rij, A2 = ra[2] / rij2, A3 = ra[3] / (rij2 * rij);
                const double d[3] = {d0, d1, d2};
                int n2 = kA2 - 3, n3 = kA3 - 3;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    acc[kA1 - 3 + a] += d[a] * A1;
                    acc[kA3b - 3 + a] += d[a] * B1;
#pragma unroll
                    for (int b = a; b < 3; ++b) {
                        acc[n2++] += d[a] * d[b] * A2;
#pragma unroll
                        for (int c = b; c < 3; ++c) acc[n3++] += d[a] * d[b] * d[c] * A3;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kMeamSums; ++q) acc[q] = wave_sum(acc[q]);
        en = wave_sum(en);
        if (lane == 0) {
            // per-site closing (:1406-1540)
            double R[kMeamRec];
            R[kR0] = acc[0];
#pragma unroll
            for (int q = 1; q < kMeamSums; ++q) R[q + 3] = acc[q];
            R[kR1] = 0.0;
            R[kR2] = -R[kA2b] * R[kA2b] / 3.0;
            R[kR3] = 0.0;
            const double v2[6] = {1, 2, 2, 1, 2, 1}, v3[10] = {1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
            for (int a = 0; a < 3; ++a) {
                R[kR1] += R[kA1 + a] * R[kA1 + a];
                R[kR3] -= 0.6 * R[kA3b + a] * R[kA3b + a];
            }
            for (int a = 0; a < 6; ++a) R[kR2] += v2[a] * R[kA2 + a] * R[kA2 + a];
            for (int a = 0; a < 10; ++a) R[kR3] += v3[a] * R[kA3 + a] * R[kA3 + a];
            if (R[kR0] > 0.0) {
                if (M.ialloy == 1) {
                    for (int a = 0; a < 3; ++a) R[kT + a] = R[kT + a] / R[kTsq + a];
                } else if (M.ialloy == 2) {
                    for (int a = 0; a < 3; ++a) R[kT + a] = M.t[it][a];
                } else {
                    for (int a = 0; a < 3; ++a) R[kT + a] = R[kT + a] / R[kR0];
                }
            }
            R[kGam] = R[kT] * R[kR1] + R[kT + 1] * R[kR2] + R[kT + 2] * R[kR3];
            if (R[kR0] > 0.0) R[kGam] = R[kGam] / (R[kR0] * R[kR0]);
            const double Z = M.Z[it];
            double dG;
            const double G = meam_G(M, R[kGam], M.ibar[it], &dG);
            const double rv = R[kR0] * G;
            const double bkgd = M.bkgd_dyn == 1 ? M.rho0[it] * Z : M.rho_ref[it];
            const double rhob = rv / bkgd, denom = 1.0 / bkgd;
            R[kD1] = (G - 2.0 * dG * R[kGam]) * denom;
            R[kD2] = R[kR0] != 0 ? (dG / R[kR0]) * denom : 0.0;
            R[kD3] = 0.0;
            const double B = M.A[it] * M.Ec[it];
            double F = 0.0;
            if (rhob != 0.0) {
                if (M.emb_lin_neg == 1 && rhob <= 0.0) {
                    R[kFp] = -B;
                    F = -(B * rhob);
                } else {
                    R[kFp] = B * (log(rhob) + 1.0);
                    F = B * rhob * log(rhob);
                }
            } else {
                R[kFp] = M.emb_lin_neg == 1 ? -B : B;
            }
            en += F;
            double* out = rec + (size_t)i * kMeamRec;
            for (int q = 0; q < kMeamRec; ++q) out[q] = R[q];
        }
    }
    double acc1[1] = {lane == 0 ? en : 0.0};
    block_store_waves<1>(acc1, part);
}

// pass 2: pair forces, energy and virial at half weight per direction, and
// the site's own three-body screening forces (:1552-2140); part[7 per block]
// = energy, virial
__global__ __launch_bounds__(kBlock) void k_meam_force(MeamDev M, const double4* __restrict__ pos,
                                                       const double* __restrict__ type,
                                                       const int* __restrict__ off, const int* __restrict__ nl,
                                                       const double* __restrict__ tab, int natms,
                                                       double4* __restrict__ pairs,
                                                       const double* __restrict__ rec, double* __restrict__ f,
                                                       int* __restrict__ nact, int* __restrict__ act,
                                                       double* __restrict__ part) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kMeamWaves + w;
    const MeamStage S(w);
    double en = 0.0, vir[6] = {0, 0, 0, 0, 0, 0}, fi[3] = {0, 0, 0};
    if (i < natms) {
        const int p0 = off[i], n = off[i + 1] - p0;
        const double4 xi = pos[i];
        meam_stage(S, pos, type, nl, p0, n, lane, xi);
        if (lane < kMeamRec) S.rec[lane] = rec[(size_t)i * kMeamRec + lane];
        __builtin_amdgcn_wave_barrier();
        const double* Ri = S.rec;
        const int it = (int)type[i];
        const double v2[6] = {1, 2, 2, 1, 2, 1}, v3[10] = {1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
        // per wave: the entries of the current chunk with three-body screening
        // forces, so that their (entry, k) candidates spread over the lanes
        __shared__ double4 s3x[kMeamWaves][64], s3p[kMeamWaves][64];
        __shared__ int s3j[kMeamWaves][64];
        double4* e3x = s3x[w];
        double4* e3p = s3p[w];
        int* e3j = s3j[w];
        // the entries this site owns (j > i) whose screening moves other atoms,
        // compacted in list order for k_meam_kforce
        int na = 0;
        for (int c = 0; c < n; c += 64) {
            const int jn = c + lane;
            const bool valid = jn < n;
            bool three = false;
            double4 pv = make_double4(0.0, 0.0, 0.0, 0.0), xj = make_double4(0.0, 0.0, 0.0, 0.0);
            double dUds_store = 0.0, rij2 = 1.0, sij3 = 0.0, dUds3 = 0.0;
            int j = -1;
            if (valid) {
            pv = pairs[p0 + jn];
            xj = S.nb[jn];
            j = S.idx[jn];
            const int jt = (int)xj.w;
            const double xij = xi.x - xj.x, yij = xi.y - xj.y, zij = xi.z - xj.z;
            const double d[3] = {-xij, -yij, -zij};
            {
#pragma clang fp contract(off)
                rij2 = xij * xij + yij * yij + zij * zij;
            }
            const double rij = sqrt(rij2);
            if (pv.y != 0.0 && rij < M.rc) {
                const double* Rj = rec + (size_t)j * kMeamRec;
                const double sij = pv.y * pv.x, dscr = pv.z;
                double phip;
                const double phi = meam_phi(tab, M, it, jt, rij, &phip);
                en += 0.5 * (phi * sij);
                const double recip = 1.0 / rij;
                double ra_i[4], dra_i[4], ra_j[4], dra_j[4];
                {
                    const double ai = rij * M.inv_re[it] - 1.0;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        ra_i[q] = M.rho0[it] * exp(-M.beta[it][q] * ai);
                        dra_i[q] = -M.beta[it][q] * M.inv_re[it] * ra_i[q];
                    }
                    if (it != jt) {
                        const double aj = rij * M.inv_re[jt] - 1.0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            ra_j[q] = M.rho0[jt] * exp(-M.beta[jt][q] * aj);
                            dra_j[q] = -M.beta[jt][q] * M.inv_re[jt] * ra_j[q];
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            ra_j[q] = ra_i[q];
                            dra_j[q] = dra_i[q];
                        }
                    }
                    if (M.ialloy == 1) {
#pragma unroll
                        for (int q = 1; q < 4; ++q) {
                            ra_j[q] *= M.t[jt][q - 1];
                            ra_i[q] *= M.t[it][q - 1];
                            dra_j[q] *= M.t[jt][q - 1];
                            dra_i[q] *= M.t[it][q - 1];
                        }
                    }
                }
                // contractions with the bond (the reference adds the rho2
                // components without their direction factor, :1692-1693)
                double a1i = 0, a1j = 0, a2i = 0, a2j = 0, a3i = 0, a3j = 0, b3i = 0, b3j = 0;
                {
                    int n2 = 0, n3 = 0;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
#pragma unroll
                        for (int b = a; b < 3; ++b) {
#pragma unroll
                            for (int c = b; c < 3; ++c) {
                                const double arg = d[a] * d[b] * d[c] * v3[n3];
                                a3i += Ri[kA3 + n3] * arg;
                                a3j -= Rj[kA3 + n3] * arg;
                                ++n3;
                            }
                            const double arg = d[a] * d[b] * v2[n2];
                            a2i += M.fix ? Ri[kA2 + n2] * arg : Ri[kA2 + n2];
                            a2j += M.fix ? Rj[kA2 + n2] * arg : Rj[kA2 + n2];
                            ++n2;
                        }
                        a1i += Ri[kA1 + a] * d[a];
                        a1j -= Rj[kA1 + a] * d[a];
                        b3i += Ri[kA3b + a] * d[a];
                        b3j -= Rj[kA3b + a] * d[a];
                    }
                }
                const double drho0dr1 = dra_j[0] * sij, drho0dr2 = dra_i[0] * sij;
                double k1 = 2.0 * sij / rij;
                const double drho1dr1 = k1 * (dra_j[1] - ra_j[1] / rij) * a1i;
                const double drho1dr2 = k1 * (dra_i[1] - ra_i[1] / rij) * a1j;
                double k2 = 2.0 * sij / rij2;
                const double drho2dr1 = k2 * (dra_j[2] - 2.0 * ra_j[2] / rij) * a2i - 2.0 / 3.0 * Ri[kA2b] * dra_j[2] * sij;
                const double drho2dr2 = k2 * (dra_i[2] - 2.0 * ra_i[2] / rij) * a2j - 2.0 / 3.0 * Rj[kA2b] * dra_i[2] * sij;
                k2 = 4.0 * sij / rij2;
                const double rij3 = rij * rij2;
                double k3 = 2.0 * sij / rij3, k3a = 1.2 * sij / rij;
                const double drho3dr1 = k3 * (dra_j[3] - 3.0 * ra_j[3] / rij) * a3i - k3a * (dra_j[3] - ra_j[3] / rij) * b3i;
                const double drho3dr2 = k3 * (dra_i[3] - 3.0 * ra_i[3] / rij) * a3j - k3a * (dra_i[3] - ra_i[3] / rij) * b3j;
                k3 = 6.0 * sij / rij3;
                k3a = 1.2 * sij / rij;
                const double t1i = Ri[kT], t2i = Ri[kT + 1], t3i = Ri[kT + 2];
                const double t1j = Rj[kT], t2j = Rj[kT + 1], t3j = Rj[kT + 2];
                double dUm[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double m1i = k1 * ra_j[1] * Ri[kA1 + a], m1j = -k1 * ra_i[1] * Rj[kA1 + a];
                    double s2i = 0.0, s2j = 0.0;
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const int q = a < b ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : (b == 0 ? a : (b == 1 ? 2 + a : 5));
                        s2i += Ri[kA2 + q] * d[b];
                        s2j -= Rj[kA2 + q] * d[b];
                    }
                    const double m2i = s2i * (k2 * ra_j[2]), m2j = s2j * (-k2 * ra_i[2]);
                    double s3i = 0.0, s3j = 0.0;
                    int q2 = 0;
#pragma unroll
                    for (int b = 0; b < 3; ++b)
#pragma unroll
                        for (int c = b; c < 3; ++c) {
                            // index of the sorted triple (a, b, c) among i <= j <= k
                            int x = a, y = b, z = c;
                            if (x > y) { const int u = x; x = y; y = u; }
                            if (y > z) { const int u = y; y = z; z = u; }
                            if (x > y) { const int u = x; x = y; y = u; }
                            const int q3 = x == 0 ? (y == 0 ? z : (y == 1 ? 2 + z : 5)) : (x == 1 ? (y == 1 ? 5 + z : 8) : 9);
                            const double arg = d[b] * d[c] * v2[q2];
                            s3i += Ri[kA3 + q3] * arg;
                            s3j += Rj[kA3 + q3] * arg;
                            ++q2;
                        }
                    const double m3i = (k3 * s3i - k3a * Ri[kA3b + a]) * ra_j[3];
                    const double m3j = -(k3 * s3j - k3a * Rj[kA3b + a]) * ra_i[3];
                    const double rmi = Ri[kD2] * (t1i * m1i + t2i * m2i + t3i * m3i);
                    const double rmj = Rj[kD2] * (t1j * m1j + t2j * m2j + t3j * m3j);
                    dUm[a] = Ri[kFp] * rmi + Rj[kFp] * rmj;
                }
                double dt1dr1, dt1dr2, dt2dr1, dt2dr2, dt3dr1, dt3dr2;
                if (M.ialloy == 1) {
                    const double q1i = Ri[kTsq] != 0.0 ? dra_j[0] * sij / Ri[kTsq] : 0.0;
                    const double q1j = Rj[kTsq] != 0.0 ? dra_i[0] * sij / Rj[kTsq] : 0.0;
                    const double q2i = Ri[kTsq + 1] != 0.0 ? dra_j[0] * sij / Ri[kTsq + 1] : 0.0;
                    const double q2j = Rj[kTsq + 1] != 0.0 ? dra_i[0] * sij / Rj[kTsq + 1] : 0.0;
                    const double q3i = Ri[kTsq + 2] != 0.0 ? dra_j[0] * sij / Ri[kTsq + 2] : 0.0;
                    const double q3j = Rj[kTsq + 2] != 0.0 ? dra_i[0] * sij / Rj[kTsq + 2] : 0.0;
                    dt1dr1 = q1i * (M.t[jt][0] - t1i * M.t[jt][0] * M.t[jt][0]);
                    dt1dr2 = q1j * (M.t[it][0] - t1j * M.t[it][0] * M.t[it][0]);
                    dt2dr1 = q2i * (M.t[jt][1] - t2i * M.t[jt][1] * M.t[jt][1]);
                    dt2dr2 = q2j * (M.t[it][1] - t2j * M.t[it][1] * M.t[it][1]);
                    dt3dr1 = q3i * (M.t[jt][2] - t3i * M.t[jt][2] * M.t[jt][2]);
                    dt3dr2 = q3j * (M.t[it][2] - t3j * M.t[it][2] * M.t[it][2]);
                } else if (M.ialloy == 2) {
                    dt1dr1 = dt1dr2 = dt2dr1 = dt2dr2 = dt3dr1 = dt3dr2 = 0.0;
                } else {
                    const double qi = Ri[kR0] != 0.0 ? dra_j[0] * sij / Ri[kR0] : 0.0;
                    const double qj = Rj[kR0] != 0.0 ? dra_i[0] * sij / Rj[kR0] : 0.0;
                    dt1dr1 = qi * (M.t[jt][0] - t1i);
                    dt1dr2 = qj * (M.t[it][0] - t1j);
                    dt2dr1 = qi * (M.t[jt][1] - t2i);
                    dt2dr2 = qj * (M.t[it][1] - t2j);
                    dt3dr1 = qi * (M.t[jt][2] - t3i);
                    dt3dr2 = qj * (M.t[it][2] - t3j);
                }
                const double* si = M.shape[it];
                const double* sj = M.shape[jt];
                const double drhodr1 = Ri[kD1] * drho0dr1 +
                                       Ri[kD2] * (dt1dr1 * Ri[kR1] + t1i * drho1dr1 + dt2dr1 * Ri[kR2] + t2i * drho2dr1 +
                                                  dt3dr1 * Ri[kR3] + t3i * drho3dr1) -
                                       Ri[kD3] * (si[0] * dt1dr1 + si[1] * dt2dr1 + si[2] * dt3dr1);
                const double drhodr2 = Rj[kD1] * drho0dr2 +
                                       Rj[kD2] * (dt1dr2 * Rj[kR1] + t1j * drho1dr2 + dt2dr2 * Rj[kR2] + t2j * drho2dr2 +
                                                  dt3dr2 * Rj[kR3] + t3j * drho3dr2) -
                                       Rj[kD3] * (sj[0] * dt1dr2 + sj[1] * dt2dr2 + sj[2] * dt3dr2);
                double dUdsij = 0.0;
                if (dscr != 0.0) {
                    const double drho0ds1 = ra_j[0], drho0ds2 = ra_i[0];
                    const double c1 = 2.0 / rij, c2 = 2.0 / rij2, c3 = 2.0 / rij3, c3a = 1.2 / rij;
                    const double drho1ds1 = c1 * ra_j[1] * a1i, drho1ds2 = c1 * ra_i[1] * a1j;
                    const double drho2ds1 = c2 * ra_j[2] * a2i - 2.0 / 3.0 * Ri[kA2b] * ra_j[2];
                    const double drho2ds2 = c2 * ra_i[2] * a2j - 2.0 / 3.0 * Rj[kA2b] * ra_i[2];
                    const double drho3ds1 = c3 * ra_j[3] * a3i - c3a * ra_j[3] * b3i;
                    const double drho3ds2 = c3 * ra_i[3] * a3j - c3a * ra_i[3] * b3j;
                    double dt1ds1, dt1ds2, dt2ds1, dt2ds2, dt3ds1, dt3ds2;
                    if (M.ialloy == 1) {
                        // the reference divides the r-derivative of the atomic density
                        // (:1896-1906) and squares t2, t3 as t2 t1, t3 t1 (:1910-1913)
                        const double wj = M.fix ? ra_j[0] : dra_j[0], wi = M.fix ? ra_i[0] : dra_i[0];
                        const double q1i = Ri[kTsq] != 0.0 ? wj / Ri[kTsq] : 0.0;
                        const double q1j = Rj[kTsq] != 0.0 ? wi / Rj[kTsq] : 0.0;
                        const double q2i = Ri[kTsq + 1] != 0.0 ? wj / Ri[kTsq + 1] : 0.0;
                        const double q2j = Rj[kTsq + 1] != 0.0 ? wi / Rj[kTsq + 1] : 0.0;
                        const double q3i = Ri[kTsq + 2] != 0.0 ? wj / Ri[kTsq + 2] : 0.0;
                        const double q3j = Rj[kTsq + 2] != 0.0 ? wi / Rj[kTsq + 2] : 0.0;
                        const double u2j = M.fix ? M.t[jt][1] : M.t[jt][0], u2i = M.fix ? M.t[it][1] : M.t[it][0];
                        const double u3j = M.fix ? M.t[jt][2] : M.t[jt][0], u3i = M.fix ? M.t[it][2] : M.t[it][0];
                        dt1ds1 = q1i * (M.t[jt][0] - t1i * M.t[jt][0] * M.t[jt][0]);
                        dt1ds2 = q1j * (M.t[it][0] - t1j * M.t[it][0] * M.t[it][0]);
                        dt2ds1 = q2i * (M.t[jt][1] - t2i * M.t[jt][1] * u2j);
                        dt2ds2 = q2j * (M.t[it][1] - t2j * M.t[it][1] * u2i);
                        dt3ds1 = q3i * (M.t[jt][2] - t3i * M.t[jt][2] * u3j);
                        dt3ds2 = q3j * (M.t[it][2] - t3j * M.t[it][2] * u3i);
                    } else if (M.ialloy == 2) {
                        dt1ds1 = dt1ds2 = dt2ds1 = dt2ds2 = dt3ds1 = dt3ds2 = 0.0;
                    } else {
                        const double qi = Ri[kR0] != 0.0 ? ra_j[0] / Ri[kR0] : 0.0;
                        const double qj = Rj[kR0] != 0.0 ? ra_i[0] / Rj[kR0] : 0.0;
                        dt1ds1 = qi * (M.t[jt][0] - t1i);
                        dt1ds2 = qj * (M.t[it][0] - t1j);
                        dt2ds1 = qi * (M.t[jt][1] - t2i);
                        dt2ds2 = qj * (M.t[it][1] - t2j);
                        dt3ds1 = qi * (M.t[jt][2] - t3i);
                        dt3ds2 = qj * (M.t[it][2] - t3j);
                    }
                    const double drhods1 = Ri[kD1] * drho0ds1 +
                                           Ri[kD2] * (dt1ds1 * Ri[kR1] + t1i * drho1ds1 + dt2ds1 * Ri[kR2] +
                                                      t2i * drho2ds1 + dt3ds1 * Ri[kR3] + t3i * drho3ds1) -
                                           Ri[kD3] * (si[0] * dt1ds1 + si[1] * dt2ds1 + si[2] * dt3ds1);
                    const double drhods2 = Rj[kD1] * drho0ds2 +
                                           Rj[kD2] * (dt1ds2 * Rj[kR1] + t1j * drho1ds2 + dt2ds2 * Rj[kR2] +
                                                      t2j * drho2ds2 + dt3ds2 * Rj[kR3] + t3j * drho3ds2) -
                                           Rj[kD3] * (sj[0] * dt1ds2 + sj[1] * dt2ds2 + sj[2] * dt3ds2);
                    dUdsij = phi + Ri[kFp] * drhods1 + Rj[kFp] * drhods2;
                }
                const double dUdrij = phip * sij + Ri[kFp] * drhodr1 + Rj[kFp] * drhodr2;
                const double force = dUdrij * recip + dUdsij * dscr;
                double fm[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    fm[a] = d[a] * force + dUm[a];
                    fi[a] += fm[a];
                }
                vir[0] += 0.5 * d[0] * fm[0];
                vir[1] += 0.5 * d[1] * fm[1];
                vir[2] += 0.5 * d[2] * fm[2];
                vir[3] += 0.25 * (d[1] * fm[2] + d[2] * fm[1]);
                vir[4] += 0.25 * (d[0] * fm[2] + d[2] * fm[0]);
                vir[5] += 0.25 * (d[0] * fm[1] + d[1] * fm[0]);
                // the site's own three-body screening forces (:2033-2072) follow
                // for the flagged entries, below
                three = sij != 0.0 && sij != 1.0 && dUdsij != 0.0;
                sij3 = sij;
                dUds3 = dUdsij;
                dUds_store = (sij != 0.0 && sij != 1.0) ? dUdsij : 0.0;
            }
            pairs[p0 + jn].w = dUds_store;
            }
            const bool own = valid && j > i && dUds_store != 0.0;
            int cnt;
            const int dest = pack_dest(own, &cnt);
            if (own) act[p0 + na + dest] = jn;
            na += cnt;
            // three-body screening forces on the site: (entry, k) over the lanes
            int n3;
            const int d3 = pack_dest(three, &n3);
            if (three) {
                e3x[d3] = xj;
                e3p[d3] = make_double4(rij2, sij3, dUds3, 0.0);
                e3j[d3] = j;
            }
            __builtin_amdgcn_wave_barrier();
            for (int tq = lane; tq < n3 * n; tq += 64) {
                const int e = tq / n, kn = tq - e * n;
                if (S.idx[kn] == e3j[e]) continue;
                const double4 xj3 = e3x[e];
                const double4 q3 = e3p[e];
                const int jt = (int)xj3.w;
                const double r2 = q3.x, sij = q3.y, dUdsij = q3.z;
                const double rbound = r2 * M.ebound[it][jt];
                const double rki2 = S.r2[kn];
                if (!(rki2 < rbound)) continue;
                const double4 xk = S.nb[kn];
                const int kt = (int)xk.w;
                const double xjk = xj3.x - xk.x, yjk = xj3.y - xk.y, zjk = xj3.z - xk.z;
                double rjk2;
                {
#pragma clang fp contract(off)
                    rjk2 = xjk * xjk + yjk * yjk + zjk * zjk;
                }
                if (!(rjk2 < rbound)) continue;
                const double xki = xk.x - xi.x, yki = xk.y - xi.y, zki = xk.z - xi.z;
                const double x_ki = rki2 / r2, x_jk = rjk2 / r2;
                const double a = 1.0 - (x_ki - x_jk) * (x_ki - x_jk);
                if (a == 0.0) continue;
                double cikj = (2.0 * (x_ki + x_jk) + a - 2.0) / a;
                const double cmax = M.cmax[it][jt][kt], cmin = M.cmin[it][jt][kt];
                if (!(cikj < cmax && cikj > cmin)) continue;
                cikj = (cikj - cmin) / (cmax - cmin);
                double dfc;
                const double sikj = meam_dfcut(cikj, &dfc);
                double dc1, dc2;
                if (M.fix) meam_dc2(r2, rjk2, rki2, &dc1, &dc2);
                else meam_dc2(r2, rki2, rjk2, &dc1, &dc2);
                const double aa = (sij / (cmax - cmin)) * (dfc / sikj);
                if (aa * dc1 == 0.0 && aa * dc2 == 0.0) continue;
                const double f1 = dUdsij * (aa * dc1), f2 = dUdsij * (aa * dc2);
                const double dik[3] = {xki, yki, zki}, djk[3] = {-xjk, -yjk, -zjk};
                double fa[3], fb[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    fa[q] = f1 * dik[q];
                    fb[q] = f2 * djk[q];
                    fi[q] += fa[q];
                }
                const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
                vir[0] += third * (dik[0] * fa[0] + djk[0] * fb[0]);
                vir[1] += third * (dik[1] * fa[1] + djk[1] * fb[1]);
                vir[2] += third * (dik[2] * fa[2] + djk[2] * fb[2]);
                vir[3] += sixth * (dik[1] * fa[2] + djk[1] * fb[2] + dik[2] * fa[1] + djk[2] * fb[1]);
                vir[4] += sixth * (dik[0] * fa[2] + djk[0] * fb[2] + dik[2] * fa[0] + djk[2] * fb[0]);
                vir[5] += sixth * (dik[0] * fa[1] + djk[0] * fb[1] + dik[1] * fa[0] + djk[1] * fb[0]);
            }
            __builtin_amdgcn_wave_barrier();
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) fi[a] = wave_sum(fi[a]);
        if (lane == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) f[(size_t)i * 3 + a] = fi[a];
            nact[i] = na;
        }
    }
    double acc[7] = {wave_sum(en)};
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[1 + q] = wave_sum(vir[q]);
    block_store_waves<7>(acc, part);
}

// pass 3: the forces on site k as a screening atom of the pairs (o, p > o) of
// its local neighbours o (:2033-2120 with k local in o's list); part[6 per
// block] = virial
__global__ __launch_bounds__(kBlock) void k_meam_kforce(MeamDev M, const double4* __restrict__ pos,
                                                        const double* __restrict__ type,
                                                        const int* __restrict__ off, const int* __restrict__ nl,
                                                        int natms, const double4* __restrict__ pairs,
                                                        const int* __restrict__ nact, const int* __restrict__ act,
                                                        double* __restrict__ f, double* __restrict__ part) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = blockIdx.x * kMeamWaves + w;
    double vir[6] = {0, 0, 0, 0, 0, 0}, fk[3] = {0, 0, 0};
    // per wave: the local neighbours o of k and the running count of their
    // compacted entries, so the (o, entry) candidates spread over the lanes
    __shared__ int s_o[kMeamWaves][kMeamMaxNb];
    __shared__ int s_pre[kMeamWaves][kMeamMaxNb + 1];
    if (k < natms) {
        const double4 xk = pos[k];
        const int kt = (int)type[k];
        const int q0 = off[k], nk = off[k + 1] - q0;
        int* so = s_o[w];
        int* pre = s_pre[w];
        int carry = 0;
        for (int c = 0; c < nk; c += 64) {
            const int e = c + lane;
            int o = -1, cnt = 0;
            if (e < nk) {
                o = nl[q0 + e];
                // k local in o's list <=> o local in k's list; only o's own pairs
                // (p > o) whose screening moves other atoms contribute
                cnt = o < natms ? nact[o] : 0;
                so[e] = o;
            }
            int incl = cnt;  // inclusive scan over the chunk
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(incl, d, 64);
                if (lane >= d) incl += v;
            }
            if (e < nk) pre[e + 1] = carry + incl;
            if (c == 0 && lane == 0) pre[0] = 0;
            carry += __shfl(incl, 63, 64);
        }
        __builtin_amdgcn_wave_barrier();
        const int total = carry;
        for (int t = lane; t < total; t += 64) {
            // the entry e with pre[e] <= t < pre[e + 1]
            int lo = 0, hi = nk;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (pre[mid] <= t) lo = mid;
                else hi = mid;
            }
            const int o = so[lo];
            const int p0 = off[o];
            const int pn = act[p0 + (t - pre[lo])];
            const double4 xo = pos[o];
            const int ot = (int)type[o];
            {
                const int p = nl[p0 + pn];
                if (p == k) continue;
                const double4 pv = pairs[p0 + pn];
                const double dUds = pv.w;
                if (dUds == 0.0) continue;
                const double sij = pv.y * pv.x;
                const double4 xp = pos[p];
                const int pt = (int)type[p];
                const double xij = xo.x - xp.x, yij = xo.y - xp.y, zij = xo.z - xp.z;
                double rij2, rjk2, rki2;
                {
#pragma clang fp contract(off)
                    rij2 = xij * xij + yij * yij + zij * zij;
                }
                const double rbound = rij2 * M.ebound[ot][pt];
                const double xjk = xp.x - xk.x, yjk = xp.y - xk.y, zjk = xp.z - xk.z;
                {
#pragma clang fp contract(off)
                    rjk2 = xjk * xjk + yjk * yjk + zjk * zjk;
                }
                if (!(rjk2 < rbound)) continue;
                const double xki = xk.x - xo.x, yki = xk.y - xo.y, zki = xk.z - xo.z;
                {
#pragma clang fp contract(off)
                    rki2 = xki * xki + yki * yki + zki * zki;
                }
                if (!(rki2 < rbound)) continue;
                const double x_ki = rki2 / rij2, x_jk = rjk2 / rij2;
                const double a = 1.0 - (x_ki - x_jk) * (x_ki - x_jk);
                if (a == 0.0) continue;
                double cikj = (2.0 * (x_ki + x_jk) + a - 2.0) / a;
                const double cmax = M.cmax[ot][pt][kt], cmin = M.cmin[ot][pt][kt];
                if (!(cikj < cmax && cikj > cmin)) continue;
                cikj = (cikj - cmin) / (cmax - cmin);
                double dfc;
                const double sikj = meam_dfcut(cikj, &dfc);
                double dc1, dc2;
                if (M.fix) meam_dc2(rij2, rjk2, rki2, &dc1, &dc2);
                else meam_dc2(rij2, rki2, rjk2, &dc1, &dc2);
                const double aa = (sij / (cmax - cmin)) * (dfc / sikj);
                if (aa * dc1 == 0.0 && aa * dc2 == 0.0) continue;
                const double f1 = dUds * (aa * dc1), f2 = dUds * (aa * dc2);
                const double dik[3] = {xki, yki, zki}, djk[3] = {-xjk, -yjk, -zjk};
                double fa[3], fb[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    fa[q] = f1 * dik[q];
                    fb[q] = f2 * djk[q];
                    fk[q] -= fa[q] + fb[q];
                }
                const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
                vir[0] += third * (dik[0] * fa[0] + djk[0] * fb[0]);
                vir[1] += third * (dik[1] * fa[1] + djk[1] * fb[1]);
                vir[2] += third * (dik[2] * fa[2] + djk[2] * fb[2]);
                vir[3] += sixth * (dik[1] * fa[2] + djk[1] * fb[2] + dik[2] * fa[1] + djk[2] * fb[1]);
                vir[4] += sixth * (dik[0] * fa[2] + djk[0] * fb[2] + dik[2] * fa[0] + djk[2] * fb[0]);
                vir[5] += sixth * (dik[0] * fa[1] + djk[0] * fb[1] + dik[1] * fa[0] + djk[1] * fb[0]);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) fk[a] = wave_sum(fk[a]);
        if (lane == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) f[(size_t)k * 3 + a] += fk[a];
        }
    }
    double acc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = wave_sum(vir[q]);
    block_store_waves<6>(acc, part);
}

// the longest list (the LDS stage holds kMeamMaxNb entries per site)
__global__ void k_meam_maxlist(const int* __restrict__ off, int natms, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < natms) atomicMax(out, off[i + 1] - off[i]);
}

}  // namespace mappdmd

// MD-mode potentials of the reference on the DMD context's geometry (same
// ghosts, half list and reverse CSR; a group of kEamG = 16 lanes per site, each
// lane a stride of the site's list, summed in a fixed tree; partner shares through
// pair buffers in reverse-CSR order, no floating-point atomics):
//   plain EAM   ForceField_eam::force_calc / energy_calc, src/ff_eam.cpp:44-356
//               (F the cubic rows of ForceField_eam::interpolate, 7 coefficients;
//               rho and r*phi the rows of the DMD tables; Horner as the reference)
//   LJ          ForceField_lj::force_calc / energy_calc, src/ff_lj.cpp:160-341
//   FS          ForceField_fs::force_calc / energy_calc, src/ff_fs.cpp:379-686
// Types are carried in the per-site concentration column.  FS keeps the
// reference's factors: rho_i stores -A/sqrt(rho) after the embedding and the
// pair records hold -(t1 + 1.5 t2 dr) dr / r, so the density force of a pair is
// the reference's drhoi_dr*rho[i] + drhoj_dr*rho[j].
#pragma once
#include "force.hpp"

#pragma clang fp contract(fast)

namespace mappdmd {

struct EamBufs {
    const double* rho7;  // [nt*nt][nr][7]
    const double* phi7;  // [nphi][nr][7]
    const double* F7;    // [nt][nrho+1][7]
    int nt;
};

// FSM is FS with the cutoff tests on squares (rsq < cut^2, src/ff_fsm.cpp:424,
// :463, :518); its one-sweep density force (:466-476) is the same sum as FS's
// two factors (-A/sqrt(rho) kept per site) in another rounding order
enum { kModelEam = 0, kModelLj = 1, kModelFs = 2, kModelFsm = 3 };

// 16 lanes (one DPP row) per site in every pass: each lane takes every 16th
// list entry, the group sums per-site values in a fixed tree
constexpr int kEamG = 16;

__device__ __forceinline__ void eam_row(const FFParams<double>& P, double r, int* m_out, double* p_out) {
    double p = r * P.dr_inv;
    int m = (int)p;
    m = m < P.nr - 2 ? m : P.nr - 2;
    p -= m;
    p = p < 1.0 ? p : 1.0;
    *m_out = m;
    *p_out = p;
}

__device__ __forceinline__ double h3(const double* c, double p) { return ((c[3] * p + c[2]) * p + c[1]) * p + c[0]; }
__device__ __forceinline__ double h2d(const double* c, double p) { return (c[6] * p + c[5]) * p + c[4]; }

// one pair (i of type it, j of type jt at distance^2 rsq) of a model:
// density shares rho_i (at i) and rho_j (at j), pair energy, pair force
// coefficient fpair (force on i = fpair * dx) and the density records
// {drho_i, drho_j} of pass 2; returns false when the pair contributes nothing
template <int MODEL, bool ENERGY_ONLY>
__device__ __forceinline__ bool md_pair(const FFParams<double>& P, const EamBufs& T, const MdParams& M, int it,
                                        int jt, double rsq, double* rho_i, double* rho_j, double* phi,
                                        double* fpair, double2* dr) {
    *rho_i = *rho_j = *phi = *fpair = 0.0;
    *dr = make_double2(0.0, 0.0);
    if constexpr (MODEL == kModelEam) {
        if (!(rsq < P.cut_sq)) return false;
        const int nt = T.nt;
        const double r = sqrt(rsq), r_inv = 1.0 / r;
        int m;
        double pp;
        eam_row(P, r, &m, &pp);
        const double* ci = T.rho7 + ((size_t)type2rho(nt, jt, it) * P.nr + m) * 7;
        const double* cj = T.rho7 + ((size_t)type2rho(nt, it, jt) * P.nr + m) * 7;
        const double* cp = T.phi7 + ((size_t)type2phi(it, jt) * P.nr + m) * 7;
        *rho_i = h3(ci, pp);
        *rho_j = h3(cj, pp);
        const double z2 = h3(cp, pp), z2p = h2d(cp, pp);
        *phi = ENERGY_ONLY ? z2 / r : z2 * r_inv;
        if (!ENERGY_ONLY) {
            const double dphi_dr = z2p * r_inv - *phi * r_inv;
            *fpair = -dphi_dr * r_inv;
            *dr = make_double2(-h2d(ci, pp) * r_inv, -h2d(cj, pp) * r_inv);
        }
        return true;
    } else if constexpr (MODEL == kModelLj) {
        const int c = md_comp(it, jt);
        if (!M.chk[c] || !(rsq < M.cut_sq[c])) return false;
        const double sig = M.sig[c], eps = M.eps[c];
        const double sig2 = sig * sig / rsq;
        const double sig6 = sig2 * sig2 * sig2;
        const double sig12 = sig6 * sig6;
        *phi = 4.0 * eps * (sig12 - sig6) + M.offset[c];
        if (!ENERGY_ONLY) *fpair = 24.0 * eps * (2.0 * sig12 - sig6) / rsq;
        return true;
    } else {
        const int c = md_comp(it, jt);
        if (!M.chk[c] || !(rsq < M.cut_sq[c])) return false;
        const double r = sqrt(rsq);
        const double r_inv = 1.0 / r;
        constexpr bool kSq = MODEL == kModelFsm;
        if (kSq ? rsq < M.cut_sq_rho[c] : r < M.cut_rho[c]) {
            const double d = r - M.cut_rho[c];
            *rho_i = d * d * (M.t1[jt][it] + M.t2[jt][it] * d);
            *rho_j = d * d * (M.t1[it][jt] + M.t2[it][jt] * d);
            if (!ENERGY_ONLY)
                *dr = make_double2(-r_inv * d * (M.t1[jt][it] + 1.5 * M.t2[jt][it] * d),
                                   -r_inv * d * (M.t1[it][jt] + 1.5 * M.t2[it][jt] * d));
        }
        if (kSq ? rsq < M.cut_sq_phi[c] : r < M.cut_phi[c]) {
            const double d = r - M.cut_phi[c];
            const double k1 = M.k1[c], k2 = M.k2[c], k3 = M.k3[c];
            *phi = d * d * (k1 + k2 * r + k3 * rsq);
            if (!ENERGY_ONLY) {
                double pc = 2.0 * d * (k1 + k2 * r + k3 * rsq) + d * d * (k2 + 2.0 * k3 * r);
                pc *= -r_inv;
                *fpair = pc;
            }
        }
        return true;
    }
}


// pass 1: own density and pair forces; partner shares {rho_j, -fpair dx} to pair
// buffers; per-pair density records for pass 2
template <int MODEL, bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_eam_pass1(FFParams<double> P, EamBufs T, MdParams M,
                                                      const double4* __restrict__ pos,
                                                      const double* __restrict__ type, PairBufs L,
                                                      const int* __restrict__ rpos, double* __restrict__ rho_own,
                                                      double* __restrict__ fown, double* __restrict__ pairE,
                                                      double* __restrict__ pairF, double2* __restrict__ drho,
                                                      double* __restrict__ part) {
    const SiteGroup<kEamG> g;
    const int i = g.site;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < P.natms) {
        const int it = (int)type[i];
        const double4 xi = pos[i];
        double rho = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        const int p1 = L.off[i + 1];
        for (int p = L.off[i] + g.l; p < p1; p += kEamG) {
            const int j = L.nl[p];
            const double4 xj = pos[j];
            const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
            double rsq;
            {
#pragma clang fp contract(off)
                rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            }
            const bool owned = j < P.natms;
            double rho_i, rho_j, phi, pf;
            double2 dr;
            if (md_pair<MODEL, ENERGY_ONLY>(P, T, M, it, (int)type[j], rsq, &rho_i, &rho_j, &phi, &pf, &dr)) {
                rho += rho_i;
                acc[0] += owned ? phi : 0.5 * phi;
                if (!ENERGY_ONLY) {
                    fx += pf * dx0;
                    fy += pf * dx1;
                    fz += pf * dx2;
                    const double w = owned ? 1.0 : 0.5;
                    acc[1] -= w * pf * dx0 * dx0;
                    acc[2] -= w * pf * dx1 * dx1;
                    acc[3] -= w * pf * dx2 * dx2;
                    acc[4] -= w * pf * dx1 * dx2;
                    acc[5] -= w * pf * dx2 * dx0;
                    acc[6] -= w * pf * dx0 * dx1;
                }
            }
            if (!ENERGY_ONLY && MODEL != kModelLj) drho[p] = dr;
            if (owned) {
                const size_t q = (size_t)rpos[p];
                if (MODEL != kModelLj) pairE[q] = rho_j;
                if (!ENERGY_ONLY) {
                    pairF[q * 3 + 0] = -pf * dx0;
                    pairF[q * 3 + 1] = -pf * dx1;
                    pairF[q * 3 + 2] = -pf * dx2;
                }
            }
        }
        rho = group_sum<kEamG>(rho);
        if (!ENERGY_ONLY) {
            fx = group_sum<kEamG>(fx);
            fy = group_sum<kEamG>(fy);
            fz = group_sum<kEamG>(fz);
        }
        if (g.l == 0) {
            rho_own[i] = rho;
            if (!ENERGY_ONLY) {
                fown[(size_t)i * 3 + 0] = fx;
                fown[(size_t)i * 3 + 1] = fy;
                fown[(size_t)i * 3 + 2] = fz;
            }
        }
    }
    block_sum_store<7>(acc, part);
}

// rho_i = own + partner shares; EAM: F(rho) (cubic rows, linear above
// rho_max), F' kept; FS: -A sqrt(rho), and -A/sqrt(rho) kept
// (src/ff_fs.cpp:524-526)
template <int MODEL, bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_eam_embed(FFParams<double> P, EamBufs T, MdParams M,
                                                      const double* __restrict__ type, PairBufs L,
                                                      const double* __restrict__ rho_own,
                                                      const double* __restrict__ pairE, double* __restrict__ rho,
                                                      double* __restrict__ dF, double* __restrict__ part) {
    const SiteGroup<kEamG> g;
    const int i = g.site;
    double acc[1] = {0.0};
    if (i < P.natms) {
        double part_r = 0.0;
        const int q1 = L.roff[i + 1];
        for (int q = L.roff[i] + g.l; q < q1; q += kEamG) part_r += pairE[q];
        double r = rho_own[i] + group_sum<kEamG>(part_r);
        const int it = (int)type[i];
        if (g.l != 0) {
            // one lane per site evaluates the embedding
        } else if constexpr (MODEL == kModelFs || MODEL == kModelFsm) {
            const double sq = sqrt(r);
            acc[0] = -(M.A[it] * sq);
            rho[i] = r;
            if (!ENERGY_ONLY) dF[i] = -M.A[it] / sq;
        } else {
            double pp = r * P.drho_inv;
            int m = (int)pp;
            m = m < P.nr - 2 ? m : P.nr - 2;
            m = m > 0 ? m : 0;
            pp -= m;
            pp = pp < 1.0 ? pp : 1.0;
            const double* c = T.F7 + ((size_t)it * (P.nrho + 1) + m) * 7;
            const double d = h2d(c, pp);
            double Fv = h3(c, pp);
            if (r > P.rho_max) Fv += d * (r - P.rho_max);
            acc[0] = Fv;
            rho[i] = r;
            if (!ENERGY_ONLY) dF[i] = d;
        }
    }
    block_sum_store<1>(acc, part);
}

// pass 2: embedding forces fpair = F'_i drho_i/dr + F'_j drho_j/dr
__global__ __launch_bounds__(kBlock) void k_eam_pass2(FFParams<double> P, const double4* __restrict__ pos, PairBufs L,
                                                      const int* __restrict__ rpos, const double2* __restrict__ drho,
                                                      const double* __restrict__ dF, double* __restrict__ fown,
                                                      double* __restrict__ pairF2, double* __restrict__ part) {
    const SiteGroup<kEamG> g;
    const int i = g.site;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < P.natms) {
        const double4 xi = pos[i];
        const double dFi = dF[i];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const int p1 = L.off[i + 1];
        for (int p = L.off[i] + g.l; p < p1; p += kEamG) {
            const int j = L.nl[p];
            const double2 dr = drho[p];
            const bool owned = j < P.natms;
            double fpair = 0.0, dx0 = 0.0, dx1 = 0.0, dx2 = 0.0;
            if (dr.x != 0.0 || dr.y != 0.0) {
                const double4 xj = pos[j];
                dx0 = xi.x - xj.x;
                dx1 = xi.y - xj.y;
                dx2 = xi.z - xj.z;
                fpair = dFi * dr.x + dF[j] * dr.y;
                fx += dx0 * fpair;
                fy += dx1 * fpair;
                fz += dx2 * fpair;
                const double w = owned ? 1.0 : 0.5;
                acc[0] -= w * fpair * dx0 * dx0;
                acc[1] -= w * fpair * dx1 * dx1;
                acc[2] -= w * fpair * dx2 * dx2;
                acc[3] -= w * fpair * dx1 * dx2;
                acc[4] -= w * fpair * dx2 * dx0;
                acc[5] -= w * fpair * dx0 * dx1;
            }
            if (owned) {
                const size_t q = (size_t)rpos[p];
                pairF2[q * 3 + 0] = -dx0 * fpair;
                pairF2[q * 3 + 1] = -dx1 * fpair;
                pairF2[q * 3 + 2] = -dx2 * fpair;
            }
        }
        fx = group_sum<kEamG>(fx);
        fy = group_sum<kEamG>(fy);
        fz = group_sum<kEamG>(fz);
        if (g.l == 0) {
            fown[(size_t)i * 3 + 0] += fx;
            fown[(size_t)i * 3 + 1] += fy;
            fown[(size_t)i * 3 + 2] += fz;
        }
    }
    block_sum_store<6>(acc, part);
}

__global__ __launch_bounds__(kBlock) void k_eam_gather(int natms, PairBufs L, const double* __restrict__ fown,
                                                       const double* __restrict__ pairF,
                                                       const double* __restrict__ pairF2, double* __restrict__ f) {
    const SiteGroup<kEamG> g;
    const int i = g.site;
    if (i >= natms) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int q1 = L.roff[i + 1];
    for (int q = L.roff[i] + g.l; q < q1; q += kEamG) {
        if (pairF2) {
            a0 += pairF[(size_t)q * 3] + pairF2[(size_t)q * 3];
            a1 += pairF[(size_t)q * 3 + 1] + pairF2[(size_t)q * 3 + 1];
            a2 += pairF[(size_t)q * 3 + 2] + pairF2[(size_t)q * 3 + 2];
        } else {
            a0 += pairF[(size_t)q * 3];
            a1 += pairF[(size_t)q * 3 + 1];
            a2 += pairF[(size_t)q * 3 + 2];
        }
    }
    a0 = group_sum<kEamG>(a0);
    a1 = group_sum<kEamG>(a1);
    a2 = group_sum<kEamG>(a2);
    if (g.l == 0) {
        f[(size_t)i * 3] = fown[(size_t)i * 3] + a0;
        f[(size_t)i * 3 + 1] = fown[(size_t)i * 3 + 1] + a1;
        f[(size_t)i * 3 + 2] = fown[(size_t)i * 3 + 2] + a2;
    }
}

}  // namespace mappdmd

#pragma clang fp contract(off)

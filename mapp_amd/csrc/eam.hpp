// Plain EAM of the reference (ForceField_eam::force_calc / energy_calc,
// src/ff_eam.cpp:44-356) on the DMD context's geometry: same ghosts, half list
// and reverse CSR, one lane per site, partner shares through pair buffers in
// reverse-CSR order (no floating-point atomics).  Types are carried in the
// per-site concentration column; F uses the cubic rows of
// ForceField_eam::interpolate (7 coefficients), rho and r*phi the same rows
// as the DMD tables.  Horner evaluation as in the reference.
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

// pass 1: own density and pair forces; partner shares {rho_j, -fpair dx} to pair
// buffers; per-pair drho_i/dr, drho_j/dr (already divided by -r) for pass 2
template <bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_eam_pass1(FFParams<double> P, EamBufs T,
                                                      const double4* __restrict__ pos,
                                                      const double* __restrict__ type, PairBufs L,
                                                      const int* __restrict__ rpos, double* __restrict__ rho_own,
                                                      double* __restrict__ fown, double* __restrict__ pairE,
                                                      double* __restrict__ pairF, double2* __restrict__ drho,
                                                      double* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < P.natms) {
        const int nt = T.nt;
        const int it = (int)type[i];
        const double4 xi = pos[i];
        double rho = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        for (int p = L.off[i]; p < L.off[i + 1]; ++p) {
            const int j = L.nl[p];
            const double4 xj = pos[j];
            const double dx0 = xi.x - xj.x, dx1 = xi.y - xj.y, dx2 = xi.z - xj.z;
            double rsq;
            {
#pragma clang fp contract(off)
                rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            }
            const bool owned = j < P.natms;
            double rj = 0.0, pf = 0.0;
            double2 dr = make_double2(0.0, 0.0);
            if (rsq < P.cut_sq) {
                const int jt = (int)type[j];
                const double r = sqrt(rsq), r_inv = 1.0 / r;
                int m;
                double pp;
                eam_row(P, r, &m, &pp);
                const double* ci = T.rho7 + ((size_t)type2rho(nt, jt, it) * P.nr + m) * 7;
                const double* cj = T.rho7 + ((size_t)type2rho(nt, it, jt) * P.nr + m) * 7;
                const double* cp = T.phi7 + ((size_t)type2phi(it, jt) * P.nr + m) * 7;
                const double rho_i = h3(ci, pp), rho_j = h3(cj, pp);
                const double z2 = h3(cp, pp), z2p = h2d(cp, pp);
                const double phi = ENERGY_ONLY ? z2 / r : z2 * r_inv;
                rho += rho_i;
                rj = rho_j;
                acc[0] += owned ? phi : 0.5 * phi;
                if (!ENERGY_ONLY) {
                    const double dphi_dr = z2p * r_inv - phi * r_inv;
                    const double fpair = -dphi_dr * r_inv;
                    fx += fpair * dx0;
                    fy += fpair * dx1;
                    fz += fpair * dx2;
                    pf = fpair;
                    const double w = owned ? 1.0 : 0.5;
                    acc[1] -= w * fpair * dx0 * dx0;
                    acc[2] -= w * fpair * dx1 * dx1;
                    acc[3] -= w * fpair * dx2 * dx2;
                    acc[4] -= w * fpair * dx1 * dx2;
                    acc[5] -= w * fpair * dx2 * dx0;
                    acc[6] -= w * fpair * dx0 * dx1;
                    dr = make_double2(-h2d(ci, pp) * r_inv, -h2d(cj, pp) * r_inv);
                }
            }
            if (!ENERGY_ONLY) drho[p] = dr;
            if (owned) {
                const size_t q = (size_t)rpos[p];
                pairE[q] = rj;
                if (!ENERGY_ONLY) {
                    pairF[q * 3 + 0] = -pf * dx0;
                    pairF[q * 3 + 1] = -pf * dx1;
                    pairF[q * 3 + 2] = -pf * dx2;
                }
            }
        }
        rho_own[i] = rho;
        if (!ENERGY_ONLY) {
            fown[(size_t)i * 3 + 0] = fx;
            fown[(size_t)i * 3 + 1] = fy;
            fown[(size_t)i * 3 + 2] = fz;
        }
    }
    block_sum_store<7>(acc, part);
}

// rho_i = own + partner shares; F(rho) (cubic rows, linear above rho_max); F' kept
template <bool ENERGY_ONLY>
__global__ __launch_bounds__(kBlock) void k_eam_embed(FFParams<double> P, EamBufs T,
                                                      const double* __restrict__ type, PairBufs L,
                                                      const double* __restrict__ rho_own,
                                                      const double* __restrict__ pairE, double* __restrict__ rho,
                                                      double* __restrict__ dF, double* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (i < P.natms) {
        double r = rho_own[i];
        for (int q = L.roff[i]; q < L.roff[i + 1]; ++q) r += pairE[q];
        const int it = (int)type[i];
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
    block_sum_store<1>(acc, part);
}

// pass 2: embedding forces fpair = F'_i drho_i/dr + F'_j drho_j/dr
__global__ __launch_bounds__(kBlock) void k_eam_pass2(FFParams<double> P, const double4* __restrict__ pos, PairBufs L,
                                                      const int* __restrict__ rpos, const double2* __restrict__ drho,
                                                      const double* __restrict__ dF, double* __restrict__ fown,
                                                      double* __restrict__ pairF2, double* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < P.natms) {
        const double4 xi = pos[i];
        const double dFi = dF[i];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int p = L.off[i]; p < L.off[i + 1]; ++p) {
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
        fown[(size_t)i * 3 + 0] += fx;
        fown[(size_t)i * 3 + 1] += fy;
        fown[(size_t)i * 3 + 2] += fz;
    }
    block_sum_store<6>(acc, part);
}

__global__ __launch_bounds__(kBlock) void k_eam_gather(int natms, PairBufs L, const double* __restrict__ fown,
                                                       const double* __restrict__ pairF,
                                                       const double* __restrict__ pairF2, double* __restrict__ f) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= natms) return;
    double a0 = fown[(size_t)i * 3], a1 = fown[(size_t)i * 3 + 1], a2 = fown[(size_t)i * 3 + 2];
    for (int q = L.roff[i]; q < L.roff[i + 1]; ++q) {
        a0 += pairF[(size_t)q * 3] + pairF2[(size_t)q * 3];
        a1 += pairF[(size_t)q * 3 + 1] + pairF2[(size_t)q * 3 + 1];
        a2 += pairF[(size_t)q * 3 + 2] + pairF2[(size_t)q * 3 + 2];
    }
    f[(size_t)i * 3] = a0;
    f[(size_t)i * 3 + 1] = a1;
    f[(size_t)i * 3 + 2] = a2;
}

}  // namespace mappdmd

#pragma clang fp contract(off)

/*
 * TEST INFRASTRUCTURE ONLY -- MEAM restatement (see meam_oracle.h).
 *
 * Follows, in the reference's operation order:
 *   G_gam / dG_gam              src/ff_meam.cpp:305-390
 *   fcut / dfcut                :395-437
 *   dCfunc / dCfunc2            :442-475
 *   get_shpfcn / get_Zij(2)     :477-600
 *   zbl / erose                 :603-650
 *   init                        :654-681
 *   force_calc                  :699-766 + :1138-2142 (the active loops; the
 *                               block :767-1138 is commented out in the source)
 *   energy_calc                 :2148-2589
 *   coef                        :2593-2627
 *   alloy_params                :3192-3249
 *   compute_pair_meam           :3252-3410
 *   get_sijk / get_dens_ref     :3413-3567
 *   get_tavref / phi_meam       :3571-3842
 *   compute_reference_density   :3844-3880
 *   compute_phi_dphi            :3899-3914
 * The neighbour list is the full one (pair_wise = 0, :675); every pair sweep
 * takes j > i from it, as the reference does.
 */
#include "meam_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NR 1000
#define RD 38
#define MT ORC_MEAM_MAXT

/* per-site record slots (src/ff_meam.h:46-100) */
enum { R_RHO0 = 0, R_RHO1 = 1, R_RHO2 = 2, R_RHO3 = 3, R_A1 = 4, R_A2 = 7, R_A2B = 13, R_A3 = 14,
       R_A3B = 24, R_T = 27, R_TSQ = 30, R_GAM = 33, R_D1 = 34, R_D2 = 35, R_D3 = 36, R_FP = 37 };

typedef struct {
    orc_meam_params p; /* after coef: t1 augmented, alloy rules applied */
    double rho_ref[MT];
    double dr, dr_inv;
    double* phir; /* [ncomp][NR][7] */
    double* rec;  /* [(natms + ghosts) * RD] */
    long rec_cap;
    double* rvec; /* rho_vec: rho0 G per owned site */
    double *scr, *dscr, *fcp;
    long pair_cap;
    int fix;
    int bad; /* an abort of the reference was reached */
} meam_t;

static int comp2(int i, int j) { return i < j ? (j + 1) * j / 2 + i : (i + 1) * i / 2 + j; }

/* multiplicities of the symmetric 2- and 3-index components and their indices
   (constructor, src/ff_meam.cpp:109-151) */
static const double V2[6] = {1, 2, 2, 1, 2, 1};
static const double V3[10] = {1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
static int vind2(int i, int j) {
    if (i > j) { int t = i; i = j; j = t; }
    static const int base[3] = {0, 3, 5};
    return base[i] + (j - i);
}
static int vind3(int i, int j, int k) {
    int a[3] = {i, j, k};
    for (int s = 0; s < 2; s++)
        for (int t = 0; t < 2 - s; t++)
            if (a[t] > a[t + 1]) { int u = a[t]; a[t] = a[t + 1]; a[t + 1] = u; }
    int n = 0;
    for (int x = 0; x < 3; x++)
        for (int y = x; y < 3; y++)
            for (int z = y; z < 3; z++) {
                if (x == a[0] && y == a[1] && z == a[2]) return n;
                n++;
            }
    return -1;
}

static void mfree(void* v) {
    meam_t* m = (meam_t*)v;
    if (!m) return;
    free(m->phir); free(m->rec); free(m->rvec); free(m->scr); free(m->dscr); free(m->fcp);
    free(m);
}
static meam_t* M(orc_t* o) { return (meam_t*)orc_ext(o); }

/* ---------------------------------------------------------------- scalar functions */
/* G(Gamma) of the background density; ibar 0/4 smooth below the switch point
   (the derivative there is 0.5/G, as in the reference) */
static double G_gam(meam_t* m, double gam, int ibar, double* dG) {
    const double gs = m->p.gsmooth;
    double G = 0.0, d = 0.0;
    if (ibar == 0 || ibar == 4) {
        const double sw = -gs / (gs + 1.0);
        if (gam < sw) {
            G = pow(sw / gam, gs) / (gs + 1.0);
            G = sqrt(G);
        } else {
            G = sqrt(1.0 + gam);
        }
        d = 0.5 / G;
    } else if (ibar == 1) {
        G = exp(gam * 0.5);
        d = 0.5 * G;
    } else if (ibar == 3) {
        G = 2.0 / (1.0 + exp(-gam));
        d = G * (1.0 - 0.5 * G);
    } else if (ibar == -5) {
        if ((1.0 + gam) >= 0.0) {
            G = sqrt(1.0 + gam);
            d = 0.5 / G;
        } else {
            G = -sqrt(-1.0 - gam);
            d = -0.5 / G;
        }
    } else {
        m->bad = 1;
    }
    if (dG) *dG = d;
    return G;
}

static double fcut(double xi) {
    if (xi >= 1.0) return 1.0;
    if (xi <= 0.0) return 0.0;
    double a = 1.0 - xi;
    a *= a;
    a *= a;
    a = 1.0 - a;
    return a * a;
}
static double dfcut(double xi, double* dfc) {
    if (xi > 1.0) { *dfc = 0.0; return 1.0; }
    if (xi < 0.0) { *dfc = 0.0; return 0.0; }
    const double a = 1.0 - xi;
    const double a3 = a * a * a, a4 = a * a * a * a;
    *dfc = 8 * (1.0 - a4) * a3;
    return (1.0 - a4) * (1.0 - a4);
}
/* dCfunc2 with the reference's parameter order (rij2, rjk2, rki2) */
static void dC2(double rij2, double p_jk, double p_ki, double* d1, double* d2) {
    const double rij4 = rij2 * rij2, rjk4 = p_jk * p_jk, rki4 = p_ki * p_ki;
    const double a = p_ki - p_jk;
    double den = rij4 - a * a;
    den *= den;
    *d1 = 4.0 * rij2 * (rij4 + rki4 + 2.0 * p_ki * p_jk - 3.0 * rjk4 - 2.0 * rij2 * a) / den;
    *d2 = 4.0 * rij2 * (rij4 - 3.0 * rki4 + 2.0 * p_ki * p_jk + rjk4 + 2.0 * rij2 * a) / den;
}

static int shape(int latt, double* s) {
    s[0] = s[1] = s[2] = 0.0;
    switch (latt) {
        case ORC_FCC: case ORC_BCC: case ORC_B1: case ORC_B2: return 0;
        case ORC_HCP: s[2] = 1.0 / 3.0; return 0;
        case ORC_DIA: s[2] = 32.0 / 9.0; return 0;
        case ORC_DIM: s[0] = 1.0; s[1] = 2.0 / 3.0; s[2] = 0.4; return 0;
        default: return -1;
    }
}
static double zij(int latt) {
    switch (latt) {
        case ORC_FCC: return 12.0;
        case ORC_BCC: return 8.0;
        case ORC_HCP: return 12.0;
        case ORC_DIM: return 1.0;
        case ORC_DIA: return 4.0;
        case ORC_B1: return 6.0;
        case ORC_C11: return 10.0;
        case ORC_L12: return 12.0;
        case ORC_B2: return 8.0;
        default: return 0.0;
    }
}
/* second-neighbour count Z2, distance ratio and screening of the reference lattice */
static int zij2(int latt, double cmin, double cmax, double* Z2, double* arat, double* S) {
    int num;
    switch (latt) {
        case ORC_FCC: *Z2 = 6.0; *arat = sqrt(2.0); num = 4; break;
        case ORC_BCC: *Z2 = 6.0; *arat = 2.0 / sqrt(3.0); num = 4; break;
        case ORC_HCP: *Z2 = 6.0; *arat = sqrt(2.0); num = 4; break;
        case ORC_DIM: *Z2 = 0.0; *arat = 1.0; *S = 0.0; return 0;
        case ORC_DIA: *Z2 = 0.0; *arat = sqrt(8.0 / 3.0); num = 4; break;
        case ORC_B1: *Z2 = 12.0; *arat = sqrt(2.0); num = 2; break;
        case ORC_L12: *Z2 = 6.0; *arat = sqrt(2.0); num = 4; break;
        case ORC_B2: *Z2 = 6.0; *arat = 2.0 / sqrt(3.0); num = 4; break;
        default: return -1;
    }
    const double C = 4.0 / (*arat * *arat) - 1.0;
    const double x = (C - cmin) / (cmax - cmin);
    *S = pow(fcut(x), num);
    return 0;
}
static double zbl(double r, double z1, double z2) {
    const double a = 0.4685 / (pow(z1, 0.23) + pow(z2, 0.23));
    const double x = r / a;
    double v = 0.028171 * exp(-0.20162 * x) + 0.28022 * exp(-0.40290 * x) + 0.50986 * exp(-0.94229 * x) +
               0.18175 * exp(-3.1998 * x);
    if (r > 0.0) v *= z1 * z2 * 14.3997 / r;
    return v;
}
static double erose(double r, double re, double alpha, double Ec, double repuls, double attrac, int form) {
    if (!(r > 0.0)) return 0.0;
    const double astar = alpha * (r / re - 1.0);
    const double a3 = astar >= 0.0 ? attrac : repuls;
    if (form == 1) return -Ec * (1.0 + astar + (repuls / r - attrac) * astar * astar * astar) * exp(-astar);
    if (form == 2) return -Ec * (1.0 + astar + a3 * astar * astar * astar) * exp(-astar);
    return -Ec * (1.0 + astar + a3 * (astar * astar * astar) / (r / re)) * exp(-astar);
}

/* ---------------------------------------------------------------- reference structure */
static double sijk(meam_t* m, double C, int i, int j, int k) {
    return fcut((C - m->p.cmin[i][j][k]) / (m->p.cmax[i][j][k] - m->p.cmin[i][j][k]));
}

/* partial densities of the a-b reference structure at r (get_dens_ref) */
static void dens_ref(meam_t* m, double r, int a, int b, double* r0a, double* r1a, double* r2a, double* r3a,
                     double* r0b, double* r1b, double* r2b, double* r3b) {
    const orc_meam_params* p = &m->p;
    double a1 = r / p->re[a][a] - 1.0, a2 = r / p->re[b][b] - 1.0;
    const double ra01 = p->rho0[a] * exp(-p->beta0[a] * a1), ra11 = p->rho0[a] * exp(-p->beta1[a] * a1);
    const double ra21 = p->rho0[a] * exp(-p->beta2[a] * a1), ra31 = p->rho0[a] * exp(-p->beta3[a] * a1);
    const double ra02 = p->rho0[b] * exp(-p->beta0[b] * a2), ra12 = p->rho0[b] * exp(-p->beta1[b] * a2);
    const double ra22 = p->rho0[b] * exp(-p->beta2[b] * a2), ra32 = p->rho0[b] * exp(-p->beta3[b] * a2);
    const int lat = p->lattice[a][b];
    *r1a = *r2a = *r3a = *r1b = *r2b = *r3b = 0.0;
    switch (lat) {
        case ORC_FCC: *r0a = 12.0 * ra02; *r0b = 12.0 * ra01; break;
        case ORC_BCC: *r0a = 8.0 * ra02; *r0b = 8.0 * ra01; break;
        case ORC_B1: *r0a = 6.0 * ra02; *r0b = 6.0 * ra01; break;
        case ORC_DIA:
            *r0a = 4.0 * ra02; *r0b = 4.0 * ra01;
            *r3a = 32.0 / 9.0 * ra32 * ra32; *r3b = 32.0 / 9.0 * ra31 * ra31;
            break;
        case ORC_HCP:
            *r0a = 12.0 * ra02; *r0b = 12.0 * ra01;
            *r3a = 1.0 / 3.0 * ra32 * ra32; *r3b = 1.0 / 3.0 * ra31 * ra31;
            break;
        case ORC_DIM: {
            double s[3];
            shape(lat, s);
            *r0a = ra02; *r0b = ra01;
            *r1a = s[0] * ra12 * ra12; *r1b = s[0] * ra11 * ra11;
            *r2a = s[1] * ra22 * ra22; *r2b = s[1] * ra21 * ra21;
            *r3a = s[2] * ra32 * ra32; *r3b = s[2] * ra31 * ra31;
            break;
        }
        case ORC_C11:
            *r0a = ra01; *r0b = ra02; *r1a = ra11; *r1b = ra12;
            *r2a = ra21; *r2b = ra22; *r3a = ra31; *r3b = ra32;
            break;
        case ORC_L12:
            *r0a = 8.0 * ra01 + 4.0 * ra02;
            *r0b = 12.0 * ra01;
            if (p->ialloy == 1) {
                *r2a = (8.0 / 3.0) * (ra21 * p->t2[a] - ra22 * p->t2[b]) * (ra21 * p->t2[a] - ra22 * p->t2[b]);
                const double den = 8.0 * ra01 * p->t2[a] * p->t2[a] + 4.0 * ra02 * p->t2[b] * p->t2[b];
                if (den > 0.0) *r2a *= *r0a / den;
            } else {
                *r2a = 8.0 / 3.0 * (ra21 - ra22) * (ra21 - ra22);
            }
            break;
        case ORC_B2: *r0a = 8.0 * ra02; *r0b = 8.0 * ra01; break;
        default: m->bad = 1; *r0a = *r0b = 0.0; return;
    }
    if (p->nn2[a][b] == 1) {
        double Z2, arat, scrn;
        if (zij2(lat, p->cmin[a][a][b], p->cmax[a][a][b], &Z2, &arat, &scrn)) { m->bad = 1; return; }
        a1 = arat * r / p->re[a][a] - 1.0;
        a2 = arat * r / p->re[b][b] - 1.0;
        const double ra01nn = p->rho0[a] * exp(-p->beta0[a] * a1);
        const double ra02nn = p->rho0[b] * exp(-p->beta0[b] * a2);
        if (lat == ORC_L12) {
            const double s111 = sijk(m, 1.0, a, a, a), s112 = sijk(m, 1.0, a, a, b), s221 = sijk(m, 1.0, b, b, a);
            const double S11 = s111 * s111 * s112 * s112, S22 = s221 * s221 * s221 * s221;
            *r0a += 6.0 * S11 * ra01nn;
            *r0b += 6.0 * S22 * ra02nn;
        } else {
            *r0a += Z2 * scrn * ra01nn;
            if (zij2(lat, p->cmin[b][b][a], p->cmax[b][b][a], &Z2, &arat, &scrn)) { m->bad = 1; return; }
            *r0b += Z2 * scrn * ra02nn;
        }
    }
}

/* averaged t of the reference structure (get_tavref); out[0..5] = t11 t21 t31 t12 t22 t32 */
static void tav_ref(meam_t* m, double r, int a, int b, int latt, double* out) {
    const orc_meam_params* p = &m->p;
    const double t1a = p->t1[a], t2a = p->t2[a], t3a = p->t3[a], t1b = p->t1[b], t2b = p->t2[b], t3b = p->t3[b];
    if (p->ialloy == 2) {
        out[0] = t1a; out[1] = t2a; out[2] = t3a; out[3] = t1b; out[4] = t2b; out[5] = t3b;
        return;
    }
    switch (latt) {
        case ORC_FCC: case ORC_BCC: case ORC_DIA: case ORC_HCP: case ORC_B1: case ORC_DIM: case ORC_B2:
            out[0] = t1b; out[1] = t2b; out[2] = t3b; out[3] = t1a; out[4] = t2a; out[5] = t3a;
            return;
        case ORC_L12: {
            const double a1 = r / p->re[a][a] - 1.0, a2 = r / p->re[b][b] - 1.0;
            const double ra01 = p->rho0[a] * exp(-p->beta0[a] * a1), ra02 = p->rho0[b] * exp(-p->beta0[b] * a2);
            const double r01 = 8.0 * ra01 + 4.0 * ra02;
            out[0] = (8.0 * t1a * ra01 + 4.0 * t1b * ra02) / r01;
            out[3] = t1a;
            out[1] = (8.0 * t2a * ra01 + 4.0 * t2b * ra02) / r01;
            out[4] = t2a;
            out[2] = (8.0 * t3a * ra01 + 4.0 * t3b * ra02) / r01;
            out[5] = t3a;
            return;
        }
        default: m->bad = 1; return;
    }
}

static double emb(const orc_meam_params* p, int a, double rb) {
    if (rb == 0.0) return 0.0;
    if (p->emb_lin_neg == 1 && rb <= 0.0) return -p->A[a] * p->Ec[a][a] * rb;
    return p->A[a] * p->Ec[a][a] * rb * log(rb);
}

/* pair potential of the a-b reference structure (phi_meam) */
static double phi_meam(meam_t* m, double r, int a, int b) {
    const orc_meam_params* p = &m->p;
    const int lat = p->lattice[a][b];
    const double Z12 = zij(lat);
    double r01, r11, r21, r31, r02, r12, r22, r32;
    dens_ref(m, r, a, b, &r01, &r11, &r21, &r31, &r02, &r12, &r22, &r32);
    if (r01 <= 1.0e-14 && r02 <= 1.0e-14) return 0.0;
    double t[6] = {0, 0, 0, 0, 0, 0};
    if (lat == ORC_C11) {
        if (p->ialloy == 2) {
            t[0] = p->t1[a]; t[1] = p->t2[a]; t[2] = p->t3[a]; t[3] = p->t1[b]; t[4] = p->t2[b]; t[5] = p->t3[b];
        } else {
            const double sc = 1.0 / (r01 + r02);
            t[0] = sc * (p->t1[a] * r01 + p->t1[b] * r02); t[3] = t[0];
            t[1] = sc * (p->t2[a] * r01 + p->t2[b] * r02); t[4] = t[1];
            t[2] = sc * (p->t3[a] * r01 + p->t3[b] * r02); t[5] = t[2];
        }
    } else {
        tav_ref(m, r, a, b, lat, t);
    }
    double rb1, rb2;
    if (lat == ORC_C11) {
        if (p->lattice[a][a] == ORC_DIA) {
            rb1 = 0.25 * Z12 * Z12 * (r02 + r01) * (r02 + r01) + t[0] * (r12 - r11) * (r12 - r11) +
                  t[1] / 6.0 * (r22 + r21) * (r22 + r21) + 3.025 * t[2] * (r32 - r31) * (r32 - r31);
            rb1 = sqrt(rb1);
            rb2 = Z12 * r01 * Z12 * r01 + 2.0 / 3.0 * t[1] * r21 * r21;
            rb2 = sqrt(rb2);
        } else {
            rb2 = 0.25 * Z12 * Z12 * (r01 + r02) * (r01 + r02) + t[3] * (r12 - r11) * (r12 - r11) +
                  t[4] / 6.0 * (r22 + r21) * (r22 + r21) + 3.025 * t[5] * (r32 - r31) * (r32 - r31);
            rb2 = sqrt(rb2);
            rb1 = Z12 * r02 * Z12 * r02 + 2.0 / 3.0 * t[4] * r22 * r22;
            rb1 = sqrt(rb1);
        }
    } else {
        double Gam1 = t[0] * r11 + t[1] * r21 + t[2] * r31;
        Gam1 = Gam1 / (r01 * r01);
        double Gam2 = t[3] * r12 + t[4] * r22 + t[5] * r32;
        Gam2 = Gam2 / (r02 * r02);
        const double G1 = G_gam(m, Gam1, p->ibar[a], NULL), G2 = G_gam(m, Gam2, p->ibar[b], NULL);
        double bk1, bk2;
        if (p->bkgd_dyn == 1) {
            bk1 = p->rho0[a] * p->Z[a];
            bk2 = p->rho0[b] * p->Z[b];
        } else {
            bk1 = m->rho_ref[a];
            bk2 = m->rho_ref[b];
        }
        rb1 = r01 / bk1 * G1;
        rb2 = r02 / bk2 * G2;
    }
    const double F1 = emb(p, a, rb1), F2 = emb(p, b, rb2);
    const double Eu = erose(r, p->re[a][b], p->alpha[a][b], p->Ec[a][b], p->repuls[a][b], p->attrac[a][b],
                            p->erose_form);
    double phi;
    if (lat == ORC_C11) {
        if (p->lattice[a][a] == ORC_DIA)
            phi = (3.0 * Eu - F1 - 2.0 * F2 - 5.0 * phi_meam(m, r, a, a)) / Z12;
        else
            phi = (3.0 * Eu - F1 - 2.0 * F2 - 5.0 * phi_meam(m, r, b, b)) / Z12;
    } else if (lat == ORC_L12) {
        double phiaa = phi_meam(m, r, a, a);
        double Z2nn, arat, scrn;
        if (zij2(p->lattice[a][a], p->cmin[a][a][a], p->cmax[a][a][a], &Z2nn, &arat, &scrn)) m->bad = 1;
        const double Z1nn = zij(p->lattice[a][a]);
        if (scrn > 0.0) {
            double t0 = 1.0, t1 = r;
            const double t2 = -Z2nn * scrn / Z1nn, t3 = arat;
            for (int n = 0; n < 10; n++) {
                t0 *= t2;
                t1 *= t3;
                phiaa += t2 * phi_meam(m, t1, a, a); /* the ratio, not its power (:3827) */
            }
        }
        phi = Eu / 3.0 - F1 / 4.0 - F2 / 12.0 - phiaa;
    } else {
        phi = (2.0 * Eu - F1 - F2) / Z12;
    }
    if (r == 0.0) phi = 0.0;
    return phi;
}

static void reference_density(meam_t* m) {
    const orc_meam_params* p = &m->p;
    for (int t = 0; t < p->nt; t++) {
        const double Z = p->Z[t];
        double Gbar = 1.0;
        if (p->ibar[t] > 0) {
            double s[3];
            if (shape(p->lattice[t][t], s)) m->bad = 1;
            const double gam = (p->t1[t] * s[0] + p->t2[t] * s[1] + p->t3[t] * s[2]) / (Z * Z);
            Gbar = G_gam(m, gam, p->ibar[t], NULL);
        }
        double rho0 = p->rho0[t] * Z;
        if (p->nn2[t][t] == 1) {
            double Z2, arat, scrn;
            if (zij2(p->lattice[t][t], p->cmin[t][t][t], p->cmax[t][t][t], &Z2, &arat, &scrn)) m->bad = 1;
            rho0 += Z2 * (p->rho0[t] * exp(-p->beta0[t] * (arat - 1.0))) * scrn;
        }
        m->rho_ref[t] = rho0 * Gbar;
    }
}

static void alloy_params(meam_t* m) {
    orc_meam_params* p = &m->p;
    const int nt = p->nt;
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt; j++) {
            if (p->Ec[i][j] == 0.0) {
                if (p->lattice[i][j] == ORC_L12)
                    p->Ec[i][j] = 0.25 * (3.0 * p->Ec[i][i] + p->Ec[j][j]) - p->delta[i][j];
                else if (p->lattice[i][j] == ORC_C11)
                    p->Ec[i][j] = p->lattice[i][i] == ORC_DIA
                                      ? (2.0 * p->Ec[i][i] + p->Ec[j][j]) / 3.0 - p->delta[i][j]
                                      : (p->Ec[i][i] + 2.0 * p->Ec[j][j]) / 3.0 - p->delta[i][j];
                else
                    p->Ec[i][j] = 0.5 * (p->Ec[i][i] + p->Ec[j][j]);
                p->Ec[j][i] = p->Ec[i][j];
            }
            if (p->alpha[i][j] == 0.0) {
                p->alpha[i][j] = 0.5 * (p->alpha[i][i] + p->alpha[j][j]);
                p->alpha[j][i] = p->alpha[i][j];
            }
            if (p->re[i][j] == 0.0) {
                p->re[i][j] = 0.5 * (p->re[i][i] + p->re[j][j]);
                p->re[j][i] = p->re[i][j];
            }
        }
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt; j++)
            for (int k = 0; k < nt; k++) {
                const double c = p->cmax[i][j][k];
                const double e = 0.25 * (c * c) / (c - 1.0);
                p->ebound[i][j] = p->ebound[i][j] > e ? p->ebound[i][j] : e;
            }
}

/* the phi table per type pair: values at r = n dr (with the second-neighbour
   series and the ZBL blend), 5-point derivatives, Hermite cubic rows */
static void pair_table(meam_t* m) {
    const orc_meam_params* p = &m->p;
    const int nt = p->nt;
    const double dr = m->dr, dr_inv = m->dr_inv;
    for (int a = 0; a < nt; a++)
        for (int b = a; b < nt; b++) {
            double* T = m->phir + (size_t)comp2(a, b) * NR * 7;
            const int lat = p->lattice[a][b];
            for (int n = 0; n < NR; n++) {
                const double r = (double)n * dr;
                double v = phi_meam(m, r, a, b);
                if (p->nn2[a][b] == 1) {
                    double Z1 = zij(lat), Z2, arat, scrn, scrn2;
                    if (zij2(lat, p->cmin[a][a][b], p->cmax[a][a][b], &Z2, &arat, &scrn)) m->bad = 1;
                    if (lat == ORC_B1 || lat == ORC_B2 || lat == ORC_L12) {
                        const double rarat = r * arat;
                        double phiaa = phi_meam(m, rarat, a, a);
                        Z1 = zij(p->lattice[a][a]);
                        if (zij2(p->lattice[a][a], p->cmin[a][a][a], p->cmax[a][a][a], &Z2, &arat, &scrn)) m->bad = 1;
                        if (scrn > 0.0) {
                            double t0 = 1.0, t1 = rarat;
                            const double t2 = -Z2 * scrn / Z1, t3 = arat;
                            for (int q = 0; q < 10; q++) {
                                t0 *= t2;
                                t1 *= t3;
                                phiaa += t0 * phi_meam(m, t1, a, a);
                            }
                        }
                        double phibb = phi_meam(m, rarat, b, b);
                        Z1 = zij(p->lattice[b][b]);
                        if (zij2(p->lattice[b][b], p->cmin[b][b][b], p->cmax[b][b][b], &Z2, &arat, &scrn)) m->bad = 1;
                        if (scrn > 0.0) {
                            double t0 = 1.0, t1 = rarat;
                            const double t2 = -Z2 * scrn / Z1, t3 = arat;
                            for (int q = 0; q < 10; q++) {
                                t0 *= t2;
                                t1 *= t3;
                                phibb += t0 * phi_meam(m, t1, b, b);
                            }
                        }
                        if (lat == ORC_B1 || lat == ORC_B2) {
                            Z1 = zij(lat);
                            zij2(lat, p->cmin[a][a][b], p->cmax[a][a][b], &Z2, &arat, &scrn);
                            v -= 0.5 * Z2 * scrn / Z1 * phiaa;
                            zij2(lat, p->cmin[b][b][a], p->cmax[b][b][a], &Z2, &arat, &scrn2);
                            v -= 0.5 * Z2 * scrn2 / Z1 * phibb;
                        } else {
                            const double s111 = sijk(m, 1.0, a, a, a), s112 = sijk(m, 1.0, a, a, b);
                            const double s221 = sijk(m, 1.0, b, b, a);
                            const double S11 = s111 * s111 * s112 * s112, S22 = s221 * s221 * s221 * s221;
                            v += -0.75 * S11 * phiaa - 0.25 * S22 * phibb;
                        }
                    } else {
                        double t0 = 1.0, t1 = r;
                        const double t2 = -Z2 * scrn / Z1, t3 = arat;
                        for (int q = 0; q < 10; q++) {
                            t0 *= t2;
                            t1 *= t3;
                            v += t0 * phi_meam(m, t1, a, b);
                        }
                    }
                }
                if (p->zbl[a][b] == 1) {
                    const double astar = p->alpha[a][b] * (r / p->re[a][b] - 1.0);
                    if (astar <= -3.0) {
                        v = zbl(r, (double)p->ielt[a], (double)p->ielt[b]);
                    } else if (astar > -3.0 && astar < -1.0) {
                        const double frac = fcut(1.0 + 0.5 * (astar + 1.0));
                        const double pz = zbl(r, (double)p->ielt[a], (double)p->ielt[b]);
                        v = frac * v + (1.0 - frac) * pz;
                    }
                }
                T[n * 7] = v;
            }
#define V(n) T[(n) * 7]
#define D(n) T[(n) * 7 + 1]
            D(0) = V(1) - V(0);
            D(1) = 0.5 * (V(2) - V(0));
            D(NR - 2) = 0.5 * (V(NR - 1) - V(NR - 3));
            D(NR - 1) = 0.0;
            for (int n = 2; n < NR - 2; n++) D(n) = ((V(n - 2) - V(n + 2)) + 8.0 * (V(n + 1) - V(n - 1))) / 12.0;
            T[(NR - 1) * 7 + 2] = 0.0;
            T[(NR - 1) * 7 + 3] = 0.0;
            for (int n = 0; n < NR - 1; n++) {
                T[n * 7 + 2] = 3.0 * (V(n + 1) - V(n)) - 2.0 * D(n) - D(n + 1);
                T[n * 7 + 3] = D(n) + D(n + 1) - 2.0 * (V(n + 1) - V(n));
            }
            for (int n = 0; n < NR; n++) {
                T[n * 7 + 4] = D(n) * dr_inv;
                T[n * 7 + 5] = 2.0 * T[n * 7 + 2] * dr_inv;
                T[n * 7 + 6] = 3.0 * T[n * 7 + 3] * dr_inv;
            }
#undef V
#undef D
        }
}

int orc_meam_coef(orc_t* o, const orc_meam_params* p) {
    if (!p || p->nt < 1 || p->nt > MT || p->mix_ref_t == 1) return -1;
    meam_t* m = (meam_t*)calloc(1, sizeof(meam_t));
    m->p = *p;
    const int nt = p->nt;
    for (int a = 0; a < nt; a++)
        for (int b = 0; b < nt; b++)
            if (p->lattice[a][b] < ORC_FCC || p->lattice[a][b] > ORC_B2) { free(m); return -1; }
    for (int t = 0; t < nt; t++) m->p.t1[t] += (double)p->augt1 * 0.6 * p->t3[t];
    alloy_params(m);
    reference_density(m);
    m->dr = 1.1 * p->rc / (double)NR;
    m->dr_inv = 1.0 / m->dr;
    m->phir = (double*)calloc((size_t)(nt * (nt + 1) / 2) * NR * 7, sizeof(double));
    pair_table(m);
    if (m->bad) { mfree(m); return -1; }
    orc_set_ext(o, m, mfree);
    orc_mark_params(o, p->rc);
    return 0;
}

void orc_meam_set_fix(orc_t* o, int fix) {
    if (M(o)) M(o)->fix = fix;
}
const double* orc_meam_table(orc_t* o, int* nr, double* dr) {
    meam_t* m = M(o);
    if (!m) return NULL;
    *nr = NR;
    *dr = m->dr;
    return m->phir;
}
const double* orc_meam_records(orc_t* o) { return M(o) ? M(o)->rec : NULL; }
const double* orc_meam_rho_ref(orc_t* o) { return M(o) ? M(o)->rho_ref : NULL; }
double orc_meam_phi(orc_t* o, double r, int a, int b) { return M(o) ? phi_meam(M(o), r, a, b) : 0.0; }

/* ---------------------------------------------------------------- the sweeps */
/* reset (:2628-2656): records and rho_vec zeroed, pair arrays sized to the
   half-pair count */
static void reset(meam_t* m, orc_t* o) {
    const long ntot = (long)orc_natms(o) + orc_nghost(o);
    if (ntot * RD > m->rec_cap) {
        free(m->rec);
        free(m->rvec);
        m->rec = (double*)calloc((size_t)ntot * RD, sizeof(double));
        m->rvec = (double*)calloc((size_t)ntot, sizeof(double));
        m->rec_cap = ntot * RD;
    }
    memset(m->rec, 0, (size_t)ntot * RD * sizeof(double));
    memset(m->rvec, 0, (size_t)ntot * sizeof(double));
    const long np = orc_npairs(o) + 1;
    if (np > m->pair_cap) {
        free(m->scr); free(m->dscr); free(m->fcp);
        m->scr = (double*)calloc((size_t)np, sizeof(double));
        m->dscr = (double*)calloc((size_t)np, sizeof(double));
        m->fcp = (double*)calloc((size_t)np, sizeof(double));
        m->pair_cap = np;
    }
}

/* screening of pair (i, j) by the atoms k of i's list, with the cutoff taper:
   out fc, s and the combined derivative ds (:1171-1266) */
static void screen(meam_t* m, const double* x, const int* ty, const int* lst, int n, int i, int j, double rij2,
                   double rij, double* fc_out, double* s_out, double* ds_out) {
    const orc_meam_params* p = &m->p;
    const int it = ty[i], jt = ty[j];
    if (!(rij < p->rc)) {
        *fc_out = *s_out = *ds_out = 0.0;
        return;
    }
    const double rnorm = (p->rc - rij) * p->delr_inv;
    double sij = 1.0, dsij = 0.0, dfcij;
    const double rbound = p->ebound[it][jt] * rij2, rij4 = rij2 * rij2;
    const double fcij = dfcut(rnorm, &dfcij);
    dfcij *= p->delr_inv;
    for (int kn = 0; kn < n && sij != 0.0; kn++) {
        const int k = lst[kn];
        if (k == j) continue;
        const int kt = ty[k];
        const double xjk = x[3 * j] - x[3 * k], yjk = x[3 * j + 1] - x[3 * k + 1], zjk = x[3 * j + 2] - x[3 * k + 2];
        const double rjk2 = xjk * xjk + yjk * yjk + zjk * zjk;
        if (!(rjk2 < rbound)) continue;
        const double xki = x[3 * k] - x[3 * i], yki = x[3 * k + 1] - x[3 * i + 1], zki = x[3 * k + 2] - x[3 * i + 2];
        const double rki2 = xki * xki + yki * yki + zki * zki;
        if (!(rki2 < rbound)) continue;
        const double x_ki = rki2 / rij2, x_jk = rjk2 / rij2;
        double a = 1.0 - (x_ki - x_jk) * (x_ki - x_jk);
        if (!(a > 0)) continue;
        double cikj = (2.0 * (x_ki + x_jk) + a - 2.0) / a;
        const double cmax = p->cmax[it][jt][kt], cmin = p->cmin[it][jt][kt];
        if (cikj < cmax && cikj > cmin) {
            cikj = (cikj - cmin) / (cmax - cmin);
            double dfikj;
            const double sikj = dfcut(cikj, &dfikj);
            sij *= sikj;
            const double coef1 = dfikj / ((cmax - cmin) * sikj);
            a = rjk2 - rki2;
            const double b = rki2 + rjk2;
            double den = rij4 - a * a;
            den *= den;
            const double dcikj = -4.0 * (-2.0 * rij2 * a * a + rij4 * b + a * a * b) / den;
            dsij += coef1 * dcikj;
        } else if (cikj < cmin) {
            sij = 0.0;
            dsij = 0.0;
        }
    }
    const double coef1 = sij * fcij, coef2 = sij * dfcij / rij;
    dsij *= coef1;
    dsij -= coef2;
    if (!m->fix && (sij == 1.0 || sij == 0.0)) dsij = 0.0;
    if (sij * fcij == 1.0 || sij * fcij == 0.0) dsij = 0.0;
    *fc_out = fcij;
    *s_out = sij;
    *ds_out = dsij;
}

/* the four atomic densities of type t at r times s (and t1..t3 for ialloy = 1) */
static void atomic(const orc_meam_params* p, int t, double r, double s, double* ra) {
    const double ai = r / p->re[t][t] - 1.0, r0 = p->rho0[t];
    ra[0] = s * r0 * exp(-p->beta0[t] * ai);
    ra[1] = s * r0 * exp(-p->beta1[t] * ai);
    ra[2] = s * r0 * exp(-p->beta2[t] * ai);
    ra[3] = s * r0 * exp(-p->beta3[t] * ai);
    if (p->ialloy == 1) {
        ra[1] *= p->t1[t];
        ra[2] *= p->t2[t];
        ra[3] *= p->t3[t];
    }
}

/* a partner's share of type tp with atomic densities ra along d (d = x_partner
   - x_site for the site i; the partner j adds with sgn = -1 on odd ranks) */
static void add_share(const orc_meam_params* p, double* R, int tp, const double* ra, const double* d, double r,
                      double r2, double sgn) {
    R[R_RHO0] += ra[0];
    if (p->ialloy != 2) {
        R[R_T] += p->t1[tp] * ra[0];
        R[R_T + 1] += p->t2[tp] * ra[0];
        R[R_T + 2] += p->t3[tp] * ra[0];
    }
    if (p->ialloy == 1) {
        R[R_TSQ] += p->t1[tp] * p->t1[tp] * ra[0];
        R[R_TSQ + 1] += p->t2[tp] * p->t2[tp] * ra[0];
        R[R_TSQ + 2] += p->t3[tp] * p->t3[tp] * ra[0];
    }
    R[R_A2B] += ra[2];
    const double A1 = ra[1] / r, B1 = ra[3] / r, A2 = ra[2] / r2, A3 = ra[3] / (r2 * r);
    int n2 = R_A2, n3 = R_A3;
    for (int a = 0; a < 3; a++) {
        if (sgn > 0) {
            R[R_A1 + a] += d[a] * A1;
            R[R_A3B + a] += d[a] * B1;
        } else {
            R[R_A1 + a] -= d[a] * A1;
            R[R_A3B + a] -= d[a] * B1;
        }
        for (int b = a; b < 3; b++) {
            R[n2++] += d[a] * d[b] * A2;
            for (int c = b; c < 3; c++) {
                if (sgn > 0) R[n3++] += d[a] * d[b] * d[c] * A3;
                else R[n3++] -= d[a] * d[b] * d[c] * A3;
            }
        }
    }
}

/* per-site closing of the density sweep (:1406-1540): rho1..3, t averages,
   Gamma, G, background, derivative factors and the embedding; returns F */
static double finish_site(meam_t* m, double* R, int it, double* rvec) {
    const orc_meam_params* p = &m->p;
    R[R_RHO1] = 0.0;
    R[R_RHO2] = -R[R_A2B] * R[R_A2B] / 3.0;
    R[R_RHO3] = 0.0;
    for (int a = 0; a < 3; a++) {
        R[R_RHO1] += R[R_A1 + a] * R[R_A1 + a];
        R[R_RHO3] -= 0.6 * R[R_A3B + a] * R[R_A3B + a];
    }
    for (int a = 0; a < 6; a++) R[R_RHO2] += V2[a] * R[R_A2 + a] * R[R_A2 + a];
    for (int a = 0; a < 10; a++) R[R_RHO3] += V3[a] * R[R_A3 + a] * R[R_A3 + a];
    if (R[R_RHO0] > 0.0) {
        if (p->ialloy == 1) {
            for (int a = 0; a < 3; a++) R[R_T + a] = R[R_T + a] / R[R_TSQ + a];
        } else if (p->ialloy == 2) {
            R[R_T] = p->t1[it];
            R[R_T + 1] = p->t2[it];
            R[R_T + 2] = p->t3[it];
        } else {
            for (int a = 0; a < 3; a++) R[R_T + a] = R[R_T + a] / R[R_RHO0];
        }
    }
    R[R_GAM] = R[R_T] * R[R_RHO1] + R[R_T + 1] * R[R_RHO2] + R[R_T + 2] * R[R_RHO3];
    if (R[R_RHO0] > 0.0) R[R_GAM] = R[R_GAM] / (R[R_RHO0] * R[R_RHO0]);
    const double Z = p->Z[it];
    double G = G_gam(m, R[R_GAM], p->ibar[it], NULL);
    if (p->ibar[it] > 0) { /* Gbar of the element's lattice: evaluated, used only with mix_ref_t */
        double s[3];
        if (shape(p->lattice[it][it], s)) m->bad = 1;
        (void)G_gam(m, (p->t1[it] * s[0] + p->t2[it] * s[1] + p->t3[it] * s[2]) / (Z * Z), p->ibar[it], NULL);
    }
    *rvec = R[R_RHO0] * G;
    const double bkgd = p->bkgd_dyn == 1 ? p->rho0[it] * Z : m->rho_ref[it];
    const double rhob = *rvec / bkgd, denom = 1.0 / bkgd;
    double dG;
    G = G_gam(m, R[R_GAM], p->ibar[it], &dG);
    R[R_D1] = (G - 2.0 * dG * R[R_GAM]) * denom;
    R[R_D2] = R[R_RHO0] != 0 ? (dG / R[R_RHO0]) * denom : 0.0;
    R[R_D3] = 0.0;
    const double B = p->A[it] * p->Ec[it][it];
    double F = 0.0;
    if (rhob != 0.0) {
        if (p->emb_lin_neg == 1 && rhob <= 0.0) {
            R[R_FP] = -B;
            F = -(B * rhob);
        } else {
            R[R_FP] = B * (log(rhob) + 1.0);
            F = B * rhob * log(rhob);
        }
    } else {
        R[R_FP] = p->emb_lin_neg == 1 ? -B : B;
    }
    return F;
}

static double table_phi(const meam_t* m, int it, int jt, double r, double* dphi) {
    double pp = r * m->dr_inv;
    int kk = (int)pp;
    kk = kk < NR - 2 ? kk : NR - 2;
    pp -= (double)kk;
    pp = pp < 1.0 ? pp : 1.0;
    const double* c = m->phir + ((size_t)comp2(it, jt) * NR + kk) * 7;
    if (dphi) *dphi = (c[6] * pp + c[5]) * pp + c[4];
    return ((c[3] * pp + c[2]) * pp + c[1]) * pp + c[0];
}

/* density sweep shared by force_calc (pair screening kept, energy = embedding)
   and energy_calc (pair energy added in the same sweep) */
static double density_sweep(meam_t* m, orc_t* o, int with_pair_energy) {
    const orc_meam_params* p = &m->p;
    const int natms = orc_natms(o);
    const double* x = orc_x(o);
    const double* c = orc_c(o);
    const int* off = orc_list_offsets(o);
    const int* nl = orc_list(o);
    const int ntot = natms + orc_nghost(o);
    int* ty = (int*)malloc((size_t)ntot * sizeof(int));
    for (int i = 0; i < ntot; i++) ty[i] = (int)c[i];
    double en = 0.0;
    long istart = 0;
    for (int i = 0; i < natms; i++) {
        const int it = ty[i];
        double* Ri = m->rec + (size_t)i * RD;
        const int* lst = nl + off[i];
        const int n = off[i + 1] - off[i];
        for (int jn = 0; jn < n; jn++) {
            const int j = lst[jn];
            if (j <= i) continue;
            const int jt = ty[j];
            double d[3];
            for (int a = 0; a < 3; a++) d[a] = -(x[3 * i + a] - x[3 * j + a]);
            const double xij = x[3 * i] - x[3 * j], yij = x[3 * i + 1] - x[3 * j + 1], zij = x[3 * i + 2] - x[3 * j + 2];
            const double rij2 = xij * xij + yij * yij + zij * zij, rij = sqrt(rij2);
            double fc, s, ds;
            screen(m, x, ty, lst, n, i, j, rij2, rij, &fc, &s, &ds);
            m->fcp[istart] = fc;
            m->scr[istart] = s;
            m->dscr[istart] = ds;
            if (s * fc != 0.0) {
                const double sij = s * fc;
                if (with_pair_energy) {
                    const double phi = table_phi(m, it, jt, rij, NULL);
                    en += j < natms ? phi * sij : 0.5 * phi * sij;
                }
                double rai[4], raj[4];
                atomic(p, it, rij, sij, rai);
                atomic(p, jt, rij, sij, raj);
                add_share(p, Ri, jt, raj, d, rij, rij2, 1.0);
                if (j < natms) add_share(p, m->rec + (size_t)j * RD, it, rai, d, rij, rij2, -1.0);
            }
            istart++;
        }
        en += finish_site(m, Ri, it, m->rvec + i);
    }
    free(ty);
    return en;
}

/* energy and virial running sums: plain (the reference's) or Neumaier
   compensated, as the DMD oracle's compensated mode selects */
typedef struct {
    double s, c;
} macc_t;
static inline void macc_add(macc_t* a, double x, int comp) {
    if (!comp) {
        a->s += x;
        return;
    }
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}

double meam_energy_calc(orc_t* o) {
    meam_t* m = M(o);
    reset(m, o);
    return density_sweep(m, o, 1);
}

void meam_force_calc(orc_t* o, int st_clc, double* en_st) {
    meam_t* m = M(o);
    const orc_meam_params* p = &m->p;
    macc_t ns[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int kahan = orc_compensated(o);
    reset(m, o);
    ns[0].s = density_sweep(m, o, 0);
    orc_update_vec(o, m->rec, RD);

    const int natms = orc_natms(o), ntot = natms + orc_nghost(o);
    const double* x = orc_x(o);
    const double* c = orc_c(o);
    const int* off = orc_list_offsets(o);
    const int* nl = orc_list(o);
    double* f = orc_vec(o, "f");
    for (long q = 0; q < (long)natms * 3; q++) f[q] = 0.0;
    int* ty = (int*)malloc((size_t)ntot * sizeof(int));
    for (int i = 0; i < ntot; i++) ty[i] = (int)c[i];
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    long istart = 0;
    for (int i = 0; i < natms; i++) {
        const int it = ty[i];
        const int* lst = nl + off[i];
        const int n = off[i + 1] - off[i];
        const double* Ri = m->rec + (size_t)i * RD;
        for (int jn = 0; jn < n; jn++) {
            const int j = lst[jn];
            if (j <= i) continue;
            const long ip = istart++;
            if (m->scr[ip] == 0.0) continue;
            const int jt = ty[j];
            const double* Rj = m->rec + (size_t)j * RD;
            const double xij = x[3 * i] - x[3 * j], yij = x[3 * i + 1] - x[3 * j + 1], zij = x[3 * i + 2] - x[3 * j + 2];
            const double d[3] = {-xij, -yij, -zij};
            const double rij2 = xij * xij + yij * yij + zij * zij, rij = sqrt(rij2);
            if (!(rij < p->rc)) continue;
            const double sij = m->scr[ip] * m->fcp[ip];
            double phip;
            const double phi = table_phi(m, it, jt, rij, &phip);
            macc_add(&ns[0], j < natms ? phi * sij : 0.5 * phi * sij, kahan);

            /* atomic densities and their r-derivatives */
            const double recip = 1.0 / rij;
            double ra_i[4], dra_i[4], ra_j[4], dra_j[4];
            const double rinv_i = 1.0 / p->re[it][it];
            const double ai = rij * rinv_i - 1.0;
            const double b_i[4] = {p->beta0[it], p->beta1[it], p->beta2[it], p->beta3[it]};
            for (int q = 0; q < 4; q++) {
                ra_i[q] = p->rho0[it] * exp(-b_i[q] * ai);
                dra_i[q] = -b_i[q] * rinv_i * ra_i[q];
            }
            if (it != jt) {
                const double rinv_j = 1.0 / p->re[jt][jt];
                const double aj = rij * rinv_j - 1.0;
                const double b_j[4] = {p->beta0[jt], p->beta1[jt], p->beta2[jt], p->beta3[jt]};
                for (int q = 0; q < 4; q++) {
                    ra_j[q] = p->rho0[jt] * exp(-b_j[q] * aj);
                    dra_j[q] = -b_j[q] * rinv_j * ra_j[q];
                }
            } else {
                for (int q = 0; q < 4; q++) {
                    ra_j[q] = ra_i[q];
                    dra_j[q] = dra_i[q];
                }
            }
            if (p->ialloy == 1) {
                const double tj[3] = {p->t1[jt], p->t2[jt], p->t3[jt]}, ti[3] = {p->t1[it], p->t2[it], p->t3[it]};
                for (int q = 1; q < 4; q++) {
                    ra_j[q] *= tj[q - 1];
                    ra_i[q] *= ti[q - 1];
                    dra_j[q] *= tj[q - 1];
                    dra_i[q] *= ti[q - 1];
                }
            }

            /* contractions of the angular sums with the bond (rho2 without its
               direction factor in the reference) */
            double a1i = 0, a1j = 0, a2i = 0, a2j = 0, a3i = 0, a3j = 0, b3i = 0, b3j = 0;
            int n2 = 0, n3 = 0;
            for (int a = 0; a < 3; a++) {
                for (int b = a; b < 3; b++) {
                    for (int cc = b; cc < 3; cc++) {
                        const double arg = d[a] * d[b] * d[cc] * V3[n3];
                        a3i += Ri[R_A3 + n3] * arg;
                        a3j -= Rj[R_A3 + n3] * arg;
                        n3++;
                    }
                    const double arg = d[a] * d[b] * V2[n2];
                    if (m->fix) {
                        a2i += Ri[R_A2 + n2] * arg;
                        a2j += Rj[R_A2 + n2] * arg;
                    } else {
                        a2i += Ri[R_A2 + n2];
                        a2j += Rj[R_A2 + n2];
                    }
                    n2++;
                }
                a1i += Ri[R_A1 + a] * d[a];
                a1j -= Rj[R_A1 + a] * d[a];
                b3i += Ri[R_A3B + a] * d[a];
                b3j -= Rj[R_A3B + a] * d[a];
            }

            /* radial derivatives of rho0..rho3 at i (1) and j (2) */
            const double drho0dr1 = dra_j[0] * sij, drho0dr2 = dra_i[0] * sij;
            double k1 = 2.0 * sij / rij;
            const double drho1dr1 = k1 * (dra_j[1] - ra_j[1] / rij) * a1i;
            const double drho1dr2 = k1 * (dra_i[1] - ra_i[1] / rij) * a1j;
            double drho1m1[3], drho1m2[3], drho2m1[3], drho2m2[3], drho3m1[3], drho3m2[3];
            for (int a = 0; a < 3; a++) {
                drho1m1[a] = k1 * ra_j[1] * Ri[R_A1 + a];
                drho1m2[a] = -k1 * ra_i[1] * Rj[R_A1 + a];
            }
            double k2 = 2.0 * sij / rij2;
            const double drho2dr1 = k2 * (dra_j[2] - 2.0 * ra_j[2] / rij) * a2i - 2.0 / 3.0 * Ri[R_A2B] * dra_j[2] * sij;
            const double drho2dr2 = k2 * (dra_i[2] - 2.0 * ra_i[2] / rij) * a2j - 2.0 / 3.0 * Rj[R_A2B] * dra_i[2] * sij;
            k2 = 4.0 * sij / rij2;
            for (int a = 0; a < 3; a++) {
                double s1 = 0.0, s2 = 0.0;
                for (int b = 0; b < 3; b++) {
                    s1 += Ri[R_A2 + vind2(a, b)] * d[b];
                    s2 -= Rj[R_A2 + vind2(a, b)] * d[b];
                }
                drho2m1[a] = s1 * (k2 * ra_j[2]);
                drho2m2[a] = s2 * (-k2 * ra_i[2]);
            }
            const double rij3 = rij * rij2;
            double k3 = 2.0 * sij / rij3, k3a = 1.2 * sij / rij;
            const double drho3dr1 = k3 * (dra_j[3] - 3.0 * ra_j[3] / rij) * a3i - k3a * (dra_j[3] - ra_j[3] / rij) * b3i;
            const double drho3dr2 = k3 * (dra_i[3] - 3.0 * ra_i[3] / rij) * a3j - k3a * (dra_i[3] - ra_i[3] / rij) * b3j;
            k3 = 6.0 * sij / rij3;
            k3a = 1.2 * sij / rij;
            for (int a = 0; a < 3; a++) {
                double s1 = 0.0, s2 = 0.0;
                int q2 = 0;
                for (int b = 0; b < 3; b++)
                    for (int cc = b; cc < 3; cc++) {
                        const double arg = d[b] * d[cc] * V2[q2];
                        s1 += Ri[R_A3 + vind3(a, b, cc)] * arg;
                        s2 += Rj[R_A3 + vind3(a, b, cc)] * arg;
                        q2++;
                    }
                drho3m1[a] = (k3 * s1 - k3a * Ri[R_A3B + a]) * ra_j[3];
                drho3m2[a] = -(k3 * s2 - k3a * Rj[R_A3B + a]) * ra_i[3];
            }

            /* derivatives of the t averages */
            const double t1i = Ri[R_T], t2i = Ri[R_T + 1], t3i = Ri[R_T + 2];
            const double t1j = Rj[R_T], t2j = Rj[R_T + 1], t3j = Rj[R_T + 2];
            double dt1dr1, dt1dr2, dt2dr1, dt2dr2, dt3dr1, dt3dr2;
            if (p->ialloy == 1) {
                const double q1i = Ri[R_TSQ] != 0.0 ? dra_j[0] * sij / Ri[R_TSQ] : 0.0;
                const double q1j = Rj[R_TSQ] != 0.0 ? dra_i[0] * sij / Rj[R_TSQ] : 0.0;
                const double q2i = Ri[R_TSQ + 1] != 0.0 ? dra_j[0] * sij / Ri[R_TSQ + 1] : 0.0;
                const double q2j = Rj[R_TSQ + 1] != 0.0 ? dra_i[0] * sij / Rj[R_TSQ + 1] : 0.0;
                const double q3i = Ri[R_TSQ + 2] != 0.0 ? dra_j[0] * sij / Ri[R_TSQ + 2] : 0.0;
                const double q3j = Rj[R_TSQ + 2] != 0.0 ? dra_i[0] * sij / Rj[R_TSQ + 2] : 0.0;
                dt1dr1 = q1i * (p->t1[jt] - t1i * p->t1[jt] * p->t1[jt]);
                dt1dr2 = q1j * (p->t1[it] - t1j * p->t1[it] * p->t1[it]);
                dt2dr1 = q2i * (p->t2[jt] - t2i * p->t2[jt] * p->t2[jt]);
                dt2dr2 = q2j * (p->t2[it] - t2j * p->t2[it] * p->t2[it]);
                dt3dr1 = q3i * (p->t3[jt] - t3i * p->t3[jt] * p->t3[jt]);
                dt3dr2 = q3j * (p->t3[it] - t3j * p->t3[it] * p->t3[it]);
            } else if (p->ialloy == 2) {
                dt1dr1 = dt1dr2 = dt2dr1 = dt2dr2 = dt3dr1 = dt3dr2 = 0.0;
            } else {
                const double qi = Ri[R_RHO0] != 0.0 ? dra_j[0] * sij / Ri[R_RHO0] : 0.0;
                const double qj = Rj[R_RHO0] != 0.0 ? dra_i[0] * sij / Rj[R_RHO0] : 0.0;
                dt1dr1 = qi * (p->t1[jt] - t1i);
                dt1dr2 = qj * (p->t1[it] - t1j);
                dt2dr1 = qi * (p->t2[jt] - t2i);
                dt2dr2 = qj * (p->t2[it] - t2j);
                dt3dr1 = qi * (p->t3[jt] - t3i);
                dt3dr2 = qj * (p->t3[it] - t3j);
            }
            double si[3], sj[3];
            shape(p->lattice[it][it], si);
            shape(p->lattice[jt][jt], sj);
            const double drhodr1 = Ri[R_D1] * drho0dr1 +
                                   Ri[R_D2] * (dt1dr1 * Ri[R_RHO1] + t1i * drho1dr1 + dt2dr1 * Ri[R_RHO2] +
                                               t2i * drho2dr1 + dt3dr1 * Ri[R_RHO3] + t3i * drho3dr1) -
                                   Ri[R_D3] * (si[0] * dt1dr1 + si[1] * dt2dr1 + si[2] * dt3dr1);
            const double drhodr2 = Rj[R_D1] * drho0dr2 +
                                   Rj[R_D2] * (dt1dr2 * Rj[R_RHO1] + t1j * drho1dr2 + dt2dr2 * Rj[R_RHO2] +
                                               t2j * drho2dr2 + dt3dr2 * Rj[R_RHO3] + t3j * drho3dr2) -
                                   Rj[R_D3] * (sj[0] * dt1dr2 + sj[1] * dt2dr2 + sj[2] * dt3dr2);
            double drhodrm1[3], drhodrm2[3];
            for (int a = 0; a < 3; a++) {
                drhodrm1[a] = Ri[R_D2] * (t1i * drho1m1[a] + t2i * drho2m1[a] + t3i * drho3m1[a]);
                drhodrm2[a] = Rj[R_D2] * (t1j * drho1m2[a] + t2j * drho2m2[a] + t3j * drho3m2[a]);
            }

            /* derivatives with respect to the screening */
            double drhods1 = 0.0, drhods2 = 0.0;
            const double dscr = m->dscr[ip];
            if (dscr != 0.0) {
                const double drho0ds1 = ra_j[0], drho0ds2 = ra_i[0];
                k1 = 2.0 / rij;
                const double drho1ds1 = k1 * ra_j[1] * a1i, drho1ds2 = k1 * ra_i[1] * a1j;
                k2 = 2.0 / rij2;
                const double drho2ds1 = k2 * ra_j[2] * a2i - 2.0 / 3.0 * Ri[R_A2B] * ra_j[2];
                const double drho2ds2 = k2 * ra_i[2] * a2j - 2.0 / 3.0 * Rj[R_A2B] * ra_i[2];
                k3 = 2.0 / rij3;
                k3a = 1.2 / rij;
                const double drho3ds1 = k3 * ra_j[3] * a3i - k3a * ra_j[3] * b3i;
                const double drho3ds2 = k3 * ra_i[3] * a3j - k3a * ra_i[3] * b3j;
                double dt1ds1, dt1ds2, dt2ds1, dt2ds2, dt3ds1, dt3ds2;
                if (p->ialloy == 1) {
                    /* the reference divides the r-derivative of the atomic
                       density (:1896-1906) where its value belongs */
                    const double wj = m->fix ? ra_j[0] : dra_j[0], wi = m->fix ? ra_i[0] : dra_i[0];
                    const double q1i = Ri[R_TSQ] != 0.0 ? wj / Ri[R_TSQ] : 0.0;
                    const double q1j = Rj[R_TSQ] != 0.0 ? wi / Rj[R_TSQ] : 0.0;
                    const double q2i = Ri[R_TSQ + 1] != 0.0 ? wj / Ri[R_TSQ + 1] : 0.0;
                    const double q2j = Rj[R_TSQ + 1] != 0.0 ? wi / Rj[R_TSQ + 1] : 0.0;
                    const double q3i = Ri[R_TSQ + 2] != 0.0 ? wj / Ri[R_TSQ + 2] : 0.0;
                    const double q3j = Rj[R_TSQ + 2] != 0.0 ? wi / Rj[R_TSQ + 2] : 0.0;
                    /* the reference squares t2, t3 as t2 t1, t3 t1 here */
                    const double u2j = m->fix ? p->t2[jt] : p->t1[jt], u2i = m->fix ? p->t2[it] : p->t1[it];
                    const double u3j = m->fix ? p->t3[jt] : p->t1[jt], u3i = m->fix ? p->t3[it] : p->t1[it];
                    dt1ds1 = q1i * (p->t1[jt] - t1i * p->t1[jt] * p->t1[jt]);
                    dt1ds2 = q1j * (p->t1[it] - t1j * p->t1[it] * p->t1[it]);
                    dt2ds1 = q2i * (p->t2[jt] - t2i * p->t2[jt] * u2j);
                    dt2ds2 = q2j * (p->t2[it] - t2j * p->t2[it] * u2i);
                    dt3ds1 = q3i * (p->t3[jt] - t3i * p->t3[jt] * u3j);
                    dt3ds2 = q3j * (p->t3[it] - t3j * p->t3[it] * u3i);
                } else if (p->ialloy == 2) {
                    dt1ds1 = dt1ds2 = dt2ds1 = dt2ds2 = dt3ds1 = dt3ds2 = 0.0;
                } else {
                    const double qi = Ri[R_RHO0] != 0.0 ? ra_j[0] / Ri[R_RHO0] : 0.0;
                    const double qj = Rj[R_RHO0] != 0.0 ? ra_i[0] / Rj[R_RHO0] : 0.0;
                    dt1ds1 = qi * (p->t1[jt] - t1i);
                    dt1ds2 = qj * (p->t1[it] - t1j);
                    dt2ds1 = qi * (p->t2[jt] - t2i);
                    dt2ds2 = qj * (p->t2[it] - t2j);
                    dt3ds1 = qi * (p->t3[jt] - t3i);
                    dt3ds2 = qj * (p->t3[it] - t3j);
                }
                drhods1 = Ri[R_D1] * drho0ds1 +
                          Ri[R_D2] * (dt1ds1 * Ri[R_RHO1] + t1i * drho1ds1 + dt2ds1 * Ri[R_RHO2] + t2i * drho2ds1 +
                                      dt3ds1 * Ri[R_RHO3] + t3i * drho3ds1) -
                          Ri[R_D3] * (si[0] * dt1ds1 + si[1] * dt2ds1 + si[2] * dt3ds1);
                drhods2 = Rj[R_D1] * drho0ds2 +
                          Rj[R_D2] * (dt1ds2 * Rj[R_RHO1] + t1j * drho1ds2 + dt2ds2 * Rj[R_RHO2] + t2j * drho2ds2 +
                                      dt3ds2 * Rj[R_RHO3] + t3j * drho3ds2) -
                          Rj[R_D3] * (sj[0] * dt1ds2 + sj[1] * dt2ds2 + sj[2] * dt3ds2);
            }

            const double dUdrij = phip * sij + Ri[R_FP] * drhodr1 + Rj[R_FP] * drhodr2;
            const double dUdsij = dscr != 0.0 ? phi + Ri[R_FP] * drhods1 + Rj[R_FP] * drhods2 : 0.0;
            double dUm[3];
            for (int a = 0; a < 3; a++) dUm[a] = Ri[R_FP] * drhodrm1[a] + Rj[R_FP] * drhodrm2[a];
            const double force = dUdrij * recip + dUdsij * dscr;
            double fi[3];
            for (int a = 0; a < 3; a++) {
                const double fm = d[a] * force + dUm[a];
                f[3 * i + a] += fm;
                if (j < natms) f[3 * j + a] -= fm;
                fi[a] = d[a] * force + dUm[a];
            }
            if (st_clc) {
                const double w = j < natms ? 1.0 : 0.5, h = j < natms ? 0.5 : 0.25;
                if (j < natms) {
                    macc_add(&ns[1], d[0] * fi[0], kahan);
                    macc_add(&ns[2], d[1] * fi[1], kahan);
                    macc_add(&ns[3], d[2] * fi[2], kahan);
                } else {
                    macc_add(&ns[1], w * d[0] * fi[0], kahan);
                    macc_add(&ns[2], w * d[1] * fi[1], kahan);
                    macc_add(&ns[3], w * d[2] * fi[2], kahan);
                }
                macc_add(&ns[4], h * (d[1] * fi[2] + d[2] * fi[1]), kahan);
                macc_add(&ns[5], h * (d[0] * fi[2] + d[2] * fi[0]), kahan);
                macc_add(&ns[6], h * (d[0] * fi[1] + d[1] * fi[0]), kahan);
            }

            /* three-body forces of the screening (:2033-2120) */
            if (sij != 0.0 && sij != 1.0) {
                const double rbound = rij2 * p->ebound[it][jt];
                for (int kn = 0; kn < n; kn++) {
                    const int k = lst[kn];
                    if (k == j) continue;
                    double ds1 = 0.0, ds2 = 0.0;
                    const int kt = ty[k];
                    const double xjk = x[3 * j] - x[3 * k], yjk = x[3 * j + 1] - x[3 * k + 1], zjk = x[3 * j + 2] - x[3 * k + 2];
                    const double rjk2 = xjk * xjk + yjk * yjk + zjk * zjk;
                    if (!(rjk2 < rbound)) continue;
                    const double xki = x[3 * k] - x[3 * i], yki = x[3 * k + 1] - x[3 * i + 1], zki = x[3 * k + 2] - x[3 * i + 2];
                    const double rki2 = xki * xki + yki * yki + zki * zki;
                    if (rki2 < rbound) {
                        const double x_ki = rki2 / rij2, x_jk = rjk2 / rij2;
                        const double a = 1.0 - (x_ki - x_jk) * (x_ki - x_jk);
                        if (a != 0.0) {
                            double cikj = (2.0 * (x_ki + x_jk) + a - 2.0) / a;
                            const double cmax = p->cmax[it][jt][kt], cmin = p->cmin[it][jt][kt];
                            if (cikj < cmax && cikj > cmin) {
                                cikj = (cikj - cmin) / (cmax - cmin);
                                double dfc;
                                const double sikj = dfcut(cikj, &dfc);
                                double dc1, dc2;
                                /* the reference passes (rki2, rjk2) to a (rjk2, rki2)
                                   signature, so its dsij1 is the derivative by
                                   rjk2 applied along i-k (and dsij2 the other) */
                                if (m->fix) dC2(rij2, rjk2, rki2, &dc1, &dc2);
                                else dC2(rij2, rki2, rjk2, &dc1, &dc2);
                                const double aa = (sij / (cmax - cmin)) * (dfc / sikj);
                                ds1 = aa * dc1;
                                ds2 = aa * dc2;
                            }
                        }
                    }
                    if (ds1 != 0.0 || ds2 != 0.0) {
                        const double f1 = dUdsij * ds1, f2 = dUdsij * ds2;
                        const double dik[3] = {xki, yki, zki}, djk[3] = {-xjk, -yjk, -zjk};
                        for (int a = 0; a < 3; a++) f[3 * i + a] += f1 * dik[a];
                        if (j < natms)
                            for (int a = 0; a < 3; a++) f[3 * j + a] += f2 * djk[a];
                        if (k < natms)
                            for (int a = 0; a < 3; a++) f[3 * k + a] -= f1 * dik[a] + f2 * djk[a];
                        if (st_clc) {
                            double fa[3], fb[3], v[6];
                            for (int a = 0; a < 3; a++) {
                                fa[a] = f1 * dik[a];
                                fb[a] = f2 * djk[a];
                            }
                            v[0] = -third * (dik[0] * fa[0] + djk[0] * fb[0]);
                            v[1] = -third * (dik[1] * fa[1] + djk[1] * fb[1]);
                            v[2] = -third * (dik[2] * fa[2] + djk[2] * fb[2]);
                            v[3] = -sixth * (dik[1] * fa[2] + djk[1] * fb[2] + dik[2] * fa[1] + djk[2] * fb[1]);
                            v[4] = -sixth * (dik[0] * fa[2] + djk[0] * fb[2] + dik[2] * fa[0] + djk[2] * fb[0]);
                            v[5] = -sixth * (dik[0] * fa[1] + djk[0] * fb[1] + dik[1] * fa[0] + djk[1] * fb[0]);
                            for (int a = 0; a < 6; a++) macc_add(&ns[a + 1], -(v[a]), kahan);
                            if (j < natms)
                                for (int a = 0; a < 6; a++) macc_add(&ns[a + 1], -(v[a]), kahan);
                            if (k < natms)
                                for (int a = 0; a < 6; a++) macc_add(&ns[a + 1], -(v[a]), kahan);
                        }
                    }
                }
            }
        }
    }
    free(ty);
    en_st[0] = ns[0].s + ns[0].c;
    if (st_clc)
        for (int a = 1; a < 7; a++) en_st[a] = ns[a].s + ns[a].c;
}

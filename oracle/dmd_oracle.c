/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see dmd_oracle.h).
 *
 * Restates, in the reference's operation order:
 *   tables      ForceField_DMD_II::set_fs         src/ff_dmd_II.cpp:1253-1480
 *               ForceField_DMD_I::allocate        src/ff_dmd_I.cpp:1654-1725
 *               ForceField_DMD_I::interpolate     src/ff_dmd_I.cpp:1760-1790
 *               ForceField_DMD_I::interpolate_m   src/ff_dmd_I.cpp:1794-1836
 *               ForceField_DMD_I::set_weight_abs  src/ff_dmd_I.cpp:1841-2134
 *               ForceField_DMD_I::coef            src/ff_dmd_I.cpp:735-796
 *   geometry    M3INV                             src/init.h:72-85
 *               Atoms::set_ph / ph_setup          src/atoms.cpp:249-372
 *               single-rank ph_xchng              src/atoms.cpp:464-497
 *               Atoms::update (single rank)       src/atoms.cpp:1088-1220
 *               Atoms::update_0                   src/atoms.cpp:1780-1865
 *               Atoms::x2s / s2x                  src/atoms.cpp:1869-1941
 *   neighbors   Neighbor::create_list (pairwise)  src/neighbor.cpp:137-238
 *               create_bin_list / find_bin_no     src/neighbor.cpp:296-392
 *               bin_atoms_s / s2bin               src/neighbor.cpp:467-525
 *   DMD         force_calc                        src/ff_dmd_I.cpp:88-450
 *               energy_calc                       src/ff_dmd_I.cpp:454-668
 *               init                              src/ff_dmd_I.cpp:672-699
 *               create_2nd_neigh_lst              src/ff_dmd_I.cpp:2138-2217
 *               Mat, calc_ent                     src/ff_dmd_I.cpp:2221-2237, src/ff_dmd.cpp:2221-2226
 *               c_d_calc  (I / II / DMD)          src/ff_dmd_I.cpp:2241-2398, src/ff_dmd_II.cpp:2186-2337, src/ff_dmd.cpp:2241-2398
 *               calc_g    (I / II / DMD)          src/ff_dmd_I.cpp:2402-2674, src/ff_dmd_II.cpp:2341-2630, src/ff_dmd.cpp:2402-2671
 *               mat, dmat (II)                    src/ff_dmd_II.cpp:2151-2171
 *   EAM         ForceField_eam::force_calc        src/ff_eam.cpp:44-264
 *               ForceField_eam::energy_calc       src/ff_eam.cpp:268-356
 *               ForceField_eam::init / coef       src/ff_eam.cpp:360-429
 *
 * Deliberate deviations (documented, all outside the reference's defined
 * behaviour): a density below zero would make the reference index its F
 * table at a negative row (undefined); the index is clamped to row 0 here.
 */
#include "dmd_oracle.h"
#include "gh_table.h"
#include "meam_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_IN_SQ 0.564189583547756286948079451561
#define MINV(a, b) ((a) < (b) ? (a) : (b))
#define MAXV(a, b) ((a) > (b) ? (a) : (b))

struct orc {
    int nt, variant, xdim;
    /* tables */
    int nr, nrho;
    double dr, drho, dr_inv, drho_inv, rho_max;
    double* F_arr;   /* [nt][nrho+1][ncF] */
    double* rho_arr; /* [nt*nt][nr][7] */
    double* phi_arr; /* [nphi][nr][7] */
    int ncF;
    int have_tables;
    /* parameters */
    int no_i;
    double xi[14], wi_0[14], wi_1[14], wi_2[14];
    double alpha_min, alpha_max, kbT, beta, rsq_crd;
    double rc, cut_sq, mod_rc, cut_sq_mod;
    double c_0[16];
    int have_params;
    /* box + atoms */
    double H[3][3], B[3][3];
    double skin, tot_cut_ph, cut_sk_sq;
    double cut_ph_s[3], s_ph_lo[3], s_ph_hi[3];
    int comm_need[3];
    int natms, nghost, cap;
    double* x;  /* cap*xdim */
    double* c;  /* cap*nt (EAM: type as double) */
    double* x0; /* natms*3 */
    int nswap;
    int swap_off[64];
    int swap_dim[64], swap_dir[64];
    int* ghost_src; /* cap */
    /* bins */
    int tot_bin_grid[3], tot_bin, bin_denom[3];
    double bin_size[3];
    /* half list */
    int* nl_off;
    int* nl;
    long no_pairs;
    /* per-atom vectors (with ghosts) */
    double *E, *dE, *ddE, *n, *s, *t, *v;
    int* crd;
    double *f, *c_d;
    /* pair caches */
    long max_pairs;
    double *rho, *drho_dr, *drho_dalpha, *phi, *dphi_dr, *dphi_dalpha;
    int fresh;
    int fix_dE_index; /* 1: dE[atom*nt+type] in the chemical-potential sweep */
    int compensated;  /* 1: energy and virial sums of force_calc / energy_calc compensated */
    /* MD-mode potentials: per type pair (COMP order) list cutoffs and coefficients */
    double md_cut_sq[36], md_cut_sk_sq[36];
    int md_chk[36], md_set[36];
    double lj_eps[36], lj_sig[36], lj_cut[36], lj_offset[36];
    double fs_A[8], fs_t1[8][8], fs_t2[8][8];
    double fs_cut_rho[36], fs_cut_phi[36], fs_k1[36], fs_k2[36], fs_k3[36];
    double fs_cut_sq_rho[36], fs_cut_sq_phi[36];  /* FSM keeps the squares */
    int energy_full;  /* 1: consistent energy: energy_calc over all type pairs, cross-type
                         pair terms weighted as in the forces (force_calc's energy too) */
    /* MEAM state (meam_oracle.c) */
    void* ext;
    void (*free_ext)(void*);
    /* 2nd list */
    int* nl2_off;
    int* nl2;
    long no_pairs2;
};

static int type2rho(int nt, int i, int j) { return j * nt + i; }
/* weight of a cross-type pair term in the energy: 1 in the reference
   (0.5 (c c + c c) phi for itype > jtype, src/ff_dmd_I.cpp:244-255), 2 in the
   consistent energy mode, where the energy is the one whose gradient the
   forces are (the force loop weights each ordered type pair by c_i c_j, :375-380) */
#define XW(o, it, jt) (((o)->energy_full && (it) != (jt)) ? 2.0 : 1.0)
static int comp(int i, int j) { return i < j ? (j + 1) * j / 2 + i : (i + 1) * i / 2 + j; }
/* MD layout: x rows [3], c holds the type (plain EAM, LJ, FS) */
#define ORC_MD(o) ((o)->variant >= ORC_EAM)
/* list cutoff per type pair (LJ, FS, FSM) */
#define ORC_PAIRWISE(o) ((o)->variant == ORC_LJ || (o)->variant == ORC_FS || (o)->variant == ORC_FSM)

orc_t* orc_create(int nt, int variant) {
    orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
    o->nt = nt;
    o->variant = variant;
    o->xdim = variant >= ORC_EAM ? 3 : 3 + nt;
    o->fresh = 0;
    return o;
}

static void free_atoms(orc_t* o) {
    free(o->x); free(o->c); free(o->x0); free(o->ghost_src);
    free(o->E); free(o->dE); free(o->ddE); free(o->n); free(o->s); free(o->t); free(o->v);
    free(o->crd); free(o->f); free(o->c_d);
    free(o->nl_off); free(o->nl); free(o->nl2_off); free(o->nl2);
    o->x = o->c = o->x0 = NULL; o->ghost_src = NULL;
    o->E = o->dE = o->ddE = o->n = o->s = o->t = o->v = NULL;
    o->crd = NULL; o->f = o->c_d = NULL;
    o->nl_off = o->nl = o->nl2_off = o->nl2 = NULL;
}

static void free_caches(orc_t* o) {
    free(o->rho); free(o->drho_dr); free(o->drho_dalpha);
    free(o->phi); free(o->dphi_dr); free(o->dphi_dalpha);
    o->rho = o->drho_dr = o->drho_dalpha = o->phi = o->dphi_dr = o->dphi_dalpha = NULL;
    o->max_pairs = 0;
}

void orc_destroy(orc_t* o) {
    if (!o) return;
    free_atoms(o);
    free_caches(o);
    free(o->F_arr); free(o->rho_arr); free(o->phi_arr);
    if (o->ext && o->free_ext) o->free_ext(o->ext);
    free(o);
}

void* orc_ext(orc_t* o) { return o->ext; }
void orc_set_ext(orc_t* o, void* ext, void (*free_ext)(void*)) {
    if (o->ext && o->free_ext) o->free_ext(o->ext);
    o->ext = ext;
    o->free_ext = free_ext;
}
/* MEAM: coefficients are in; rc is its one cutoff */
void orc_mark_params(orc_t* o, double rc) {
    o->rc = rc;
    o->cut_sq = rc * rc;
    o->have_params = 1;
}

void orc_set_fresh_caches(orc_t* o, int fresh) { o->fresh = fresh; }
void orc_set_fix_dE_index(orc_t* o, int fix) { o->fix_dE_index = fix; }
void orc_set_compensated(orc_t* o, int on) { o->compensated = on; }
int orc_compensated(const orc_t* o) { return o->compensated; }

/* A running sum of the DMD energy and virial.  comp = 0: the reference's plain
   double running sum (src/ff_dmd_I.cpp:250-257, :343-348, :403-430); comp = 1:
   Neumaier's compensated sum of the same terms in the same order (the
   rounding error of every addition is carried in c), whose error does not grow
   with the number of terms -- the checker for sums over ~1e8 terms. */
typedef struct {
    double s, c;
} acc_t;
static inline void acc_add(acc_t* a, double x, int comp) {
    if (!comp) {
        a->s += x;
        return;
    }
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}
static inline double acc_value(const acc_t* a) { return a->s + a->c; }
void orc_set_energy_full(orc_t* o, int full) { o->energy_full = full; }

/* ---------------------------------------------------------------- tables */

/* ForceField_DMD_I::interpolate, src/ff_dmd_I.cpp:1760-1790 (stride 7) and
   interpolate_m, :1794-1836 (stride 9, quartic term); ForceField_eam::interpolate
   is the stride-7 form */
static void interpolate(int n, double delta, double* sp, int stride, int quartic) {
#define S(i, k) sp[(long)(i) * stride + (k)]
    S(0, 1) = S(1, 0) - S(0, 0);
    S(1, 1) = 0.5 * (S(2, 0) - S(0, 0));
    S(n - 2, 1) = 0.5 * (S(n - 1, 0) - S(n - 3, 0));
    S(n - 1, 1) = S(n - 1, 0) - S(n - 2, 0);
    for (int i = 2; i < n - 2; i++)
        S(i, 1) = ((S(i - 2, 0) - S(i + 2, 0)) + 8.0 * (S(i + 1, 0) - S(i - 1, 0))) / 12.0;
    for (int i = 0; i < n - 1; i++) {
        S(i, 2) = 3.0 * (S(i + 1, 0) - S(i, 0)) - 2.0 * S(i, 1) - S(i + 1, 1);
        S(i, 3) = S(i, 1) + S(i + 1, 1) - 2.0 * (S(i + 1, 0) - S(i, 0));
    }
    S(n - 1, 2) = 0.0;
    S(n - 1, 3) = 0.0;
    if (!quartic) {
        for (int i = 0; i < n; i++) {
            S(i, 4) = S(i, 1) / delta;
            S(i, 5) = 2.0 * S(i, 2) / delta;
            S(i, 6) = 3.0 * S(i, 3) / delta;
        }
    } else {
        for (int i = 0; i < n - 1; i++)
            S(i, 4) = (S(i + 1, 2) - S(i, 2)) / 6.0 - 0.5 * S(i, 3);
        S(n - 1, 4) = 0.0;
        for (int i = 0; i < n; i++) {
            S(i, 5) = S(i, 1) / delta;
            S(i, 6) = 2.0 * S(i, 2) / delta;
            S(i, 7) = 3.0 * S(i, 3) / delta;
            S(i, 8) = 4.0 * S(i, 4) / delta;
        }
    }
#undef S
}

/* set_fs (src/ff_dmd_II.cpp:1253-1480) with the file's element order equal to
   the type order, followed by set_arrays (src/ff_dmd_I.cpp:1729-1757) */
int orc_set_fs(orc_t* o, int nrho, double drho, int nr, double dr,
               const double* F, const double* rho, const double* rphi) {
    int nt = o->nt;
    if (nr < 5 || nrho < 5) return -1;
    o->nrho = nrho; o->nr = nr; o->drho = drho; o->dr = dr;
    o->dr_inv = 1.0 / dr;
    o->drho_inv = 1.0 / drho;
    int eam = o->variant == ORC_EAM;
    o->ncF = eam ? 7 : 9;
    int nphi = nt * (nt + 1) / 2;
    free(o->F_arr); free(o->rho_arr); free(o->phi_arr);
    o->F_arr = (double*)calloc((size_t)nt * (nrho + 1) * o->ncF, sizeof(double));
    o->rho_arr = (double*)calloc((size_t)nt * nt * nr * 7, sizeof(double));
    o->phi_arr = (double*)calloc((size_t)nphi * nr * 7, sizeof(double));
    /* element a: F_a, then rho_{a,b} for every b -> slot type2rho[a][b] */
    for (int a = 0; a < nt; a++) {
        for (int i = 0; i < nrho; i++) o->F_arr[((long)a * (nrho + 1) + i) * o->ncF] = F[(long)a * nrho + i];
        for (int b = 0; b < nt; b++) {
            int slot = type2rho(nt, a, b);
            const double* src = rho + ((long)a * nt + b) * nr;
            for (int i = 0; i < nr; i++) o->rho_arr[((long)slot * nr + i) * 7] = src[i];
        }
    }
    /* r*phi for i, j<=i -> slot type2phi[i][j] */
    long k = 0;
    for (int i = 0; i < nt; i++)
        for (int j = 0; j <= i; j++, k++) {
            int slot = comp(i, j);
            for (int r = 0; r < nr; r++) o->phi_arr[((long)slot * nr + r) * 7] = rphi[k * nr + r];
        }
    for (int a = 0; a < nt; a++)
        interpolate(nrho, drho, o->F_arr + (long)a * (nrho + 1) * o->ncF, o->ncF, !eam);
    for (int a = 0; a < nt; a++)
        for (int b = 0; b < nt; b++)
            interpolate(nr, dr, o->rho_arr + (long)type2rho(nt, a, b) * nr * 7, 7, 0);
    for (int i = 0; i < nt; i++)
        for (int j = 0; j < i + 1; j++)
            interpolate(nr, dr, o->phi_arr + (long)comp(i, j) * nr * 7, 7, 0);
    o->have_tables = 1;
    return 0;
}

/* coef (src/ff_dmd_I.cpp:735-796), set_weight_abs (:1841-2134);
   EAM: coef (src/ff_eam.cpp:400-429) */
int orc_set_params(orc_t* o, double r_crd, int no_i, double alpha_min,
                   double alpha_max, double T, const double* mass) {
    if (!o->have_tables) return -1;
    if (o->variant == ORC_EAM) {
        o->cut_sq = (o->nr - 1.0) * (o->nr - 1.0) * o->dr * o->dr;
        o->rho_max = (o->nrho - 1.0) * o->drho;
        o->have_params = 1;
        return 0;
    }
    if (no_i < 1 || no_i > 14) return -2;
    if (alpha_min == 0.0) return -3;
    o->rsq_crd = r_crd;
    o->rsq_crd *= o->rsq_crd;
    o->no_i = no_i;
    for (int i = 0; i < no_i; i++) {
        o->xi[i] = GH_NODES[no_i - 1][i];
        o->wi_0[i] = GH_WEIGHTS[no_i - 1][i];
    }
    for (int i = 0; i < no_i; i++) {
        o->wi_1[i] = o->wi_0[i] * o->xi[i];
        o->wi_2[i] = o->wi_0[i] * o->xi[i] * o->xi[i];
    }
    o->alpha_min = alpha_min;
    o->alpha_max = alpha_max;
    double kb = 8.617332478e-5, hbar = 6.5821192815e-16;
    for (int i = 0; i < o->nt; i++) {
        double m = mass[i];
        m *= 1.0364269184093291236e-28;
        double deb_l = hbar * hbar * 2.0 / (m * kb * T);
        o->c_0[i] = 1.5 * kb * T * (log(deb_l) - 1.0);
    }
    o->kbT = kb * T;
    o->beta = 1.0 / o->kbT;
    o->rc = ((double)o->nr - 1.0) * o->dr;
    o->rho_max = ((double)o->nrho - 1.0) * o->drho;
    o->cut_sq = o->rc * o->rc;
    o->mod_rc = o->rc + o->xi[no_i - 1] / sqrt(alpha_min);
    o->cut_sq_mod = o->mod_rc * o->mod_rc;
    o->have_params = 1;
    return 0;
}

/* --------------------------------------------------------------- geometry */

/* M3INV, src/init.h:72-85 */
void orc_m3inv(const double* Hf, double* Bf) {
    double A[3][3], Bm[3][3], det;
    memcpy(A, Hf, sizeof(A));
    Bm[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    Bm[1][1] = A[2][2] * A[0][0] - A[2][0] * A[0][2];
    Bm[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    Bm[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    Bm[2][1] = A[2][0] * A[0][1] - A[2][1] * A[0][0];
    Bm[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    Bm[2][0] = A[1][0] * A[2][1] - A[2][0] * A[1][1];
    Bm[0][1] = A[2][1] * A[0][2] - A[0][1] * A[2][2];
    Bm[1][2] = A[0][2] * A[1][0] - A[1][2] * A[0][0];
    det = A[0][0] * Bm[0][0] + A[0][1] * Bm[1][0] + A[0][2] * Bm[2][0];
    Bm[0][0] /= det; Bm[1][1] /= det; Bm[2][2] /= det;
    Bm[1][0] /= det; Bm[2][1] /= det; Bm[0][2] /= det;
    Bm[2][0] /= det; Bm[0][1] /= det; Bm[1][2] /= det;
    memcpy(Bf, Bm, sizeof(Bm));
}

/* Atoms::x2s, src/atoms.cpp:1869-1897 */
void orc_x2s(const double* Bf, int n, int xdim, double* x) {
    const double(*B)[3] = (const double(*)[3])Bf;
    for (int i = 0; i < n; i++) {
        double* p = x + (long)i * xdim;
        for (int j = 0; j < 3; j++) {
            p[j] = p[j] * B[j][j];
            for (int k = j + 1; k < 3; k++) p[j] += B[k][j] * p[k];
        }
        for (int j = 0; j < 3; j++) {
            while (p[j] < 0.0) p[j]++;
            while (p[j] >= 1.0) p[j]--;
        }
    }
}

/* Atoms::s2x, src/atoms.cpp:1924-1941 */
void orc_s2x(const double* Hf, int n, int xdim, double* x) {
    const double(*H)[3] = (const double(*)[3])Hf;
    for (int i = 0; i < n; i++) {
        double* p = x + (long)i * xdim;
        for (int j = 0; j < 3; j++) {
            p[j] = p[j] * H[j][j];
            for (int k = j + 1; k < 3; k++) p[j] += H[k][j] * p[k];
        }
    }
}

int orc_set_system(orc_t* o, const double* Hf, int natms, const double* x, const double* c) {
    free_atoms(o);
    free_caches(o);
    memcpy(o->H, Hf, sizeof(o->H));
    orc_m3inv(Hf, &o->B[0][0]);
    o->natms = natms;
    o->nghost = 0;
    o->cap = natms > 16 ? natms : 16;
    o->x = (double*)malloc((size_t)o->cap * o->xdim * sizeof(double));
    o->c = (double*)malloc((size_t)o->cap * o->nt * sizeof(double));
    int cd = ORC_MD(o) ? 1 : o->nt;
    memcpy(o->x, x, (size_t)natms * o->xdim * sizeof(double));
    memcpy(o->c, c, (size_t)natms * cd * sizeof(double));
    o->x0 = (double*)calloc((size_t)natms * 3, sizeof(double));
    o->ghost_src = NULL;
    return 0;
}

static void grow_atoms(orc_t* o, int need) {
    if (need <= o->cap) return;
    int nc = o->cap;
    while (nc < need) nc = nc + nc / 2 + 16;
    o->x = (double*)realloc(o->x, (size_t)nc * o->xdim * sizeof(double));
    o->c = (double*)realloc(o->c, (size_t)nc * o->nt * sizeof(double));
    o->cap = nc;
}

/* Atoms::set_ph(TYPE0), src/atoms.cpp:249-280 (single rank: s_lo=0, s_hi=1) */
static void set_ph(orc_t* o, double cut) {
    o->tot_cut_ph = cut + o->skin;
    for (int i = 0; i < 3; i++) {
        double tmp = 0.0;
        for (int j = i; j < 3; j++) tmp += o->B[j][i] * o->B[j][i];
        o->cut_ph_s[i] = o->tot_cut_ph * sqrt(tmp);
    }
    for (int i = 0; i < 3; i++) {
        o->s_ph_hi[i] = 1.0 - o->cut_ph_s[i];
        o->s_ph_lo[i] = 0.0 + o->cut_ph_s[i];
    }
    for (int i = 0; i < 3; i++) o->comm_need[i] = (int)(o->cut_ph_s[i] / (1.0 - 0.0)) + 1;
}

/* Atoms::ph_setup, src/atoms.cpp:317-372, with the single-rank copy of
   ph_xchng, :464-497.  x holds reduced coordinates on entry. */
static void ph_setup(orc_t* o) {
    int natms = o->natms, xd = o->xdim;
    int cd = ORC_MD(o) ? 1 : o->nt;
    o->nghost = 0;
    o->nswap = 0;
    free(o->ghost_src);
    int gcap = natms / 2 + 64;
    o->ghost_src = (int*)malloc((size_t)gcap * sizeof(int));
    int* cand = (int*)malloc((size_t)(natms + 1) * sizeof(int));
    int cand_cap = natms + 1;
    int lo_dim = 0, hi_dim = natms;
    o->swap_off[0] = 0;
    for (int idim = 0; idim < 3; idim++) {
        for (int dir = 0; dir < 2; dir++) {
            int lo = lo_dim, hi = hi_dim;
            for (int k = 0; k < o->comm_need[idim]; k++) {
                int tmp_r = natms + o->nghost;
                int nc = 0;
                if (hi - lo > cand_cap) {
                    cand_cap = hi - lo;
                    cand = (int*)realloc(cand, (size_t)cand_cap * sizeof(int));
                }
                for (int i = lo; i < hi; i++) {
                    double sv = o->x[(long)i * xd + idim];
                    if (dir == 0 ? (sv <= o->s_ph_lo[idim]) : (o->s_ph_hi[idim] <= sv)) cand[nc++] = i;
                }
                grow_atoms(o, natms + o->nghost + nc);
                if (o->nghost + nc > gcap) {
                    while (gcap < o->nghost + nc) gcap = gcap * 2;
                    o->ghost_src = (int*)realloc(o->ghost_src, (size_t)gcap * sizeof(int));
                }
                int base = natms + o->nghost;
                for (int q = 0; q < nc; q++) {
                    memcpy(o->x + (long)(base + q) * xd, o->x + (long)cand[q] * xd, xd * sizeof(double));
                    memcpy(o->c + (long)(base + q) * cd, o->c + (long)cand[q] * cd, cd * sizeof(double));
                    o->ghost_src[o->nghost + q] = cand[q];
                }
                for (int q = 0; q < nc; q++) {
                    if (dir == 0) o->x[(long)(base + q) * xd + idim]++;
                    else o->x[(long)(base + q) * xd + idim]--;
                }
                o->nghost += nc;
                o->swap_dim[o->nswap] = idim;
                o->swap_dir[o->nswap] = dir;
                o->nswap++;
                o->swap_off[o->nswap] = o->nghost;
                hi = natms + o->nghost;
                lo = tmp_r;
            }
        }
        lo_dim = 0;
        hi_dim = natms + o->nghost;
    }
    free(cand);
}

/* Atoms::update for a per-atom double vector of dimension d (single rank,
   src/atoms.cpp:1181-1216): owner values copied to ghosts in swap order */
static void update_vec(orc_t* o, double* v, int d) {
    for (int g = 0; g < o->nghost; g++)
        memcpy(v + (long)(o->natms + g) * d, v + (long)o->ghost_src[g] * d, d * sizeof(double));
}
void orc_update_vec(orc_t* o, double* v, int d) { update_vec(o, v, d); }
static void update_int(orc_t* o, int* v) {
    for (int g = 0; g < o->nghost; g++) v[o->natms + g] = v[o->ghost_src[g]];
}
/* Atoms::update of x: copy then add/subtract the box row H[idim][0..idim] */
static void update_x(orc_t* o) {
    int xd = o->xdim;
    for (int sw = 0; sw < o->nswap; sw++) {
        int idim = o->swap_dim[sw], dir = o->swap_dir[sw];
        for (int g = o->swap_off[sw]; g < o->swap_off[sw + 1]; g++) {
            double* dst = o->x + (long)(o->natms + g) * xd;
            memcpy(dst, o->x + (long)o->ghost_src[g] * xd, xd * sizeof(double));
            for (int j = 0; j <= idim; j++) {
                if (dir == 0) dst[j] += o->H[idim][j];
                else dst[j] -= o->H[idim][j];
            }
        }
    }
}

/* Neighbor::create_bin_list + find_bin_no, src/neighbor.cpp:296-392 */
static void create_bin_list(orc_t* o) {
    o->tot_bin = 1;
    for (int i = 0; i < 3; i++) {
        o->bin_size[i] = o->cut_ph_s[i];
        o->tot_bin_grid[i] = (int)((2.0 * o->cut_ph_s[i] + 1.0 - 0.0) / o->bin_size[i]) + 1;
        o->tot_bin *= o->tot_bin_grid[i];
    }
    for (int i = 0; i < 3; i++) {
        int no = 1;
        for (int j = 0; j < i; j++) no *= o->tot_bin_grid[j];
        o->bin_denom[i] = no;
    }
}

static int s2bin(const orc_t* o, const double* s) {
    int no = 0;
    for (int i = 0; i < 3; i++)
        no += (int)((s[i] + o->cut_ph_s[i] - 0.0) / o->bin_size[i]) * o->bin_denom[i];
    return no;
}

/* bin_atoms_s (src/neighbor.cpp:467-503), s2x of every atom, then the pairwise
   branch of create_list (src/neighbor.cpp:191-238), or for MEAM
   (pair_wise = 0, src/ff_meam.cpp:675) the full branch (:240-284): every
   j != i in the same stencil order, ghosts included */
static void create_list(orc_t* o) {
    int ntot = o->natms + o->nghost, xd = o->xdim;
    int* first = (int*)malloc((size_t)o->tot_bin * sizeof(int));
    int* next = (int*)malloc((size_t)ntot * sizeof(int));
    int* atm_bin = (int*)malloc((size_t)(o->natms + 1) * sizeof(int));
    for (int b = 0; b < o->tot_bin; b++) first[b] = -1;
    for (int i = ntot - 1; i > -1; i--) {
        int b = s2bin(o, o->x + (long)i * xd);
        if (i < o->natms) atm_bin[i] = b;
        next[i] = first[b];
        first[b] = i;
    }
    orc_s2x(&o->H[0][0], ntot, xd, o->x);

    free(o->nl_off); free(o->nl);
    o->nl_off = (int*)malloc((size_t)(o->natms + 1) * sizeof(int));
    long cap = (long)o->natms * 48 + 64, np = 0;
    o->nl = (int*)malloc((size_t)cap * sizeof(int));
    double cut = o->cut_sk_sq;
    for (int i = 0; i < o->natms; i++) {
        o->nl_off[i] = (int)np;
        int ibin = atm_bin[i];
        int loc[3], tmp = ibin;
        for (int d = 2; d > -1; d--) {
            loc[d] = tmp / o->bin_denom[d];
            tmp -= loc[d] * o->bin_denom[d];
        }
        const double* xi = o->x + (long)i * xd;
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
                for (int cc = -1; cc <= 1; cc++) {
                    int dd[3] = {a, b, cc}, no = 0, bad = 0;
                    for (int d = 0; d < 3; d++) {
                        int t = loc[d] + dd[d];
                        if (t < 0 || t > o->tot_bin_grid[d] - 1) bad = 1;
                        no += t * o->bin_denom[d];
                    }
                    if (bad) continue;
                    for (int j = first[no]; j != -1; j = next[j]) {
                        if (o->variant == ORC_MEAM ? j == i : j <= i) continue;
                        const double* xj = o->x + (long)j * xd;
                        double rsq = 0.0;
                        for (int d = 0; d < 3; d++) rsq += (xi[d] - xj[d]) * (xi[d] - xj[d]);
                        if (ORC_PAIRWISE(o)) cut = o->md_cut_sk_sq[comp((int)o->c[i], (int)o->c[j])];
                        if (rsq < cut) {
                            if (np == cap) {
                                cap = cap * 3 / 2;
                                o->nl = (int*)realloc(o->nl, (size_t)cap * sizeof(int));
                            }
                            o->nl[np++] = j;
                        }
                    }
                }
    }
    o->nl_off[o->natms] = (int)np;
    o->no_pairs = np;
    if (o->variant == ORC_MEAM) { /* no_pairs counts j > i only (:272) */
        long nh = 0;
        for (int i = 0; i < o->natms; i++)
            for (int q = o->nl_off[i]; q < o->nl_off[i + 1]; q++) nh += o->nl[q] > i;
        o->no_pairs = nh;
    }
    free(first); free(next); free(atm_bin);
}

static void alloc_vectors(orc_t* o) {
    int ntot = o->natms + o->nghost, nt = o->nt;
    free(o->E); free(o->dE); free(o->ddE); free(o->n); free(o->s); free(o->t); free(o->v);
    free(o->crd); free(o->f); free(o->c_d);
    o->E = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->dE = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->ddE = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->n = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->s = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->t = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->v = (double*)calloc((size_t)ntot * nt, sizeof(double));
    o->crd = (int*)calloc((size_t)ntot, sizeof(int));
    o->f = (double*)calloc((size_t)o->natms * o->xdim, sizeof(double));
    o->c_d = (double*)calloc((size_t)o->natms * nt, sizeof(double));
}

static void rebuild(orc_t* o) {
    ph_setup(o);
    create_bin_list(o);
    create_list(o);
    alloc_vectors(o);
    for (int i = 0; i < o->natms; i++)
        for (int d = 0; d < 3; d++) o->x0[(long)i * 3 + d] = o->x[(long)i * o->xdim + d];
}

/* ForceField_DMD_I::init (src/ff_dmd_I.cpp:672-699) / ForceField_eam::init
   (src/ff_eam.cpp:360-381), Atoms::ph_setup, Neighbor::create_list(0,1),
   Atoms::store_0 -- the order of Min_LBFGS::init (src/min_l-bfgs.cpp:235-239) */
int orc_init(orc_t* o, double skin) {
    if (!o->have_params || !o->x) return -1;
    o->skin = skin;
    if (o->variant == ORC_MEAM) {
        /* ForceField_meam::init (src/ff_meam.cpp:654-681): cutoff+skin of rc_meam
           for every type pair, ghosts for rc_meam */
        o->cut_sk_sq = o->cut_sq + skin * skin + 2 * sqrt(o->cut_sq) * skin;
        set_ph(o, o->rc);
        rebuild(o);
        return 0;
    }
    if (ORC_PAIRWISE(o)) {
        /* ForceField_lj::init (src/ff_lj.cpp:129-153), ForceField_fs::init
           (src/ff_fs.cpp:339-360): cutoff+skin of every type pair with
           coefficients, ghosts for the largest cutoff */
        double ph_cut = 0.0;
        int ncomp = o->nt * (o->nt + 1) / 2;
        for (int k = 0; k < ncomp; k++)
            if (o->md_chk[k]) {
                o->md_cut_sk_sq[k] = o->md_cut_sq[k] + skin * skin + 2 * sqrt(o->md_cut_sq[k]) * skin;
                ph_cut = MAXV(ph_cut, sqrt(o->md_cut_sq[k]));
            } else if (o->md_set[k] && o->variant == ORC_LJ) {
                o->md_cut_sk_sq[k] = (o->lj_cut[k] + skin) * (o->lj_cut[k] + skin);  /* coef time value */
            }
        o->cut_sk_sq = 0.0;
        set_ph(o, ph_cut);
        rebuild(o);
        return 0;
    }
    double base = o->variant == ORC_EAM ? o->cut_sq : o->cut_sq_mod;
    o->cut_sk_sq = base + skin * skin + 2.0 * sqrt(base) * skin;
    set_ph(o, sqrt(base));
    rebuild(o);
    return 0;
}

/* ForceField_lj::coef (src/ff_lj.cpp:72-125); the shift of the constructor is
   off (:35), so the offset is zero.  A type pair with a zero cutoff, sigma or
   epsilon has no coefficients but keeps the list cutoff (cut+skin)^2 of the
   coef call (taken here with the skin of init); a type pair never given
   coefficients, whose cutoff the reference leaves uninitialised, lists nothing. */
int orc_lj_coef(orc_t* o, int itype, int jtype, double eps, double sig, double cut) {
    if (o->variant != ORC_LJ || itype < 0 || jtype < 0 || itype >= o->nt || jtype >= o->nt) return -1;
    if (eps < 0.0 || sig < 0.0 || cut < 0.0) return -2;
    int k = comp(itype, jtype);
    o->lj_eps[k] = eps;
    o->lj_sig[k] = sig;
    o->lj_cut[k] = cut;
    o->lj_offset[k] = 0.0;
    o->md_cut_sq[k] = cut * cut;
    o->md_chk[k] = (cut == 0.0 || sig == 0.0 || eps == 0.0) ? 0 : 1;
    o->md_set[k] = 1;
    o->have_params = 1;
    return 0;
}

/* ForceField_fs::coef (src/ff_fs.cpp:113-335) with the file's values in type
   order; cut_sq = max(cut_phi^2, cut_rho^2) at init (src/ff_fs.cpp:346).
   ForceField_fsm::coef (src/ff_fsm.cpp:111-333) reads the same file and keeps
   cut_rho^2, cut_phi^2 (:281-282); cut_sq = max of the two at init (:345) */
int orc_fs_coef(orc_t* o, const double* A, const double* t1, const double* t2, const double* cut_rho,
                const double* cut_phi, const double* k1, const double* k2, const double* k3) {
    if (o->variant != ORC_FS && o->variant != ORC_FSM) return -1;
    int nt = o->nt, ncomp = nt * (nt + 1) / 2;
    for (int i = 0; i < nt; i++) {
        o->fs_A[i] = A[i];
        for (int j = 0; j < nt; j++) {
            o->fs_t1[i][j] = t1[i * nt + j];
            o->fs_t2[i][j] = t2[i * nt + j];
        }
    }
    for (int k = 0; k < ncomp; k++) {
        o->fs_cut_rho[k] = cut_rho[k];
        o->fs_cut_phi[k] = cut_phi[k];
        o->fs_k1[k] = k1[k];
        o->fs_k2[k] = k2[k];
        o->fs_k3[k] = k3[k];
        o->fs_cut_sq_rho[k] = cut_rho[k] * cut_rho[k];
        o->fs_cut_sq_phi[k] = cut_phi[k] * cut_phi[k];
        o->md_cut_sq[k] = MAXV(o->fs_cut_sq_phi[k], o->fs_cut_sq_rho[k]);
        o->md_chk[k] = 1;
        o->md_set[k] = 1;
    }
    o->have_params = 1;
    return 0;
}

/* Atoms::update_0 (src/atoms.cpp:1780-1865), single rank */
int orc_set_x_update(orc_t* o, const double* x_owned, int* rebuilt) {
    int xd = o->xdim;
    memcpy(o->x, x_owned, (size_t)o->natms * xd * sizeof(double));
    double sq_skin = o->skin * o->skin;
    int check = 0;
    for (int i = 0; i < o->natms && !check; i++) {
        double sq = 0.0;
        for (int d = 0; d < 3; d++) {
            double dx = o->x[(long)i * xd + d] - o->x0[(long)i * 3 + d];
            sq += dx * dx;
        }
        if (sq > sq_skin) check = 1;
    }
    if (check) {
        orc_x2s(&o->B[0][0], o->natms, xd, o->x);
        rebuild(o);
    } else {
        update_x(o);
    }
    if (rebuilt) *rebuilt = check;
    return 0;
}

/* new owned c, forwarded to the ghosts (Atoms::update of c) */
void orc_set_c(orc_t* o, const double* c_owned) {
    int cd = ORC_MD(o) ? 1 : o->nt;
    memcpy(o->c, c_owned, (size_t)o->natms * cd * sizeof(double));
    update_vec(o, o->c, cd);
}

/* ------------------------------------------------------------------- DMD */

static void ensure_caches(orc_t* o) {
    int nt = o->nt, nphi = nt * (nt + 1) / 2;
    if (o->max_pairs < o->no_pairs) {
        free_caches(o);
        o->max_pairs = o->no_pairs;
        long n0 = o->max_pairs * nt * nt, n1 = o->max_pairs * nphi;
        o->rho = (double*)calloc((size_t)n0 + 1, sizeof(double));
        o->drho_dr = (double*)calloc((size_t)n0 + 1, sizeof(double));
        o->drho_dalpha = (double*)calloc((size_t)n0 + 1, sizeof(double));
        o->phi = (double*)calloc((size_t)n1 + 1, sizeof(double));
        o->dphi_dr = (double*)calloc((size_t)n1 + 1, sizeof(double));
        o->dphi_dalpha = (double*)calloc((size_t)n1 + 1, sizeof(double));
    } else if (o->fresh) {
        long n0 = o->max_pairs * nt * nt, n1 = o->max_pairs * nphi;
        memset(o->rho, 0, n0 * sizeof(double));
        memset(o->drho_dr, 0, n0 * sizeof(double));
        memset(o->drho_dalpha, 0, n0 * sizeof(double));
        memset(o->phi, 0, n1 * sizeof(double));
        memset(o->dphi_dr, 0, n1 * sizeof(double));
        memset(o->dphi_dalpha, 0, n1 * sizeof(double));
    }
}

static double calc_ent(double x) {
    if (x < 1.0e-8 || 0.99999999 < x) return 0.0;
    return x * log(x) + (1 - x) * log(1 - x);
}

/* embedding F(rho_bar) and F' for one (site, type); src/ff_dmd_I.cpp:324-341 */
static void embed(const orc_t* o, int type, double rho_bar, double* F, double* dF, double* ddF) {
    double p = rho_bar * o->drho_inv;
    int m = (int)p;
    m = MINV(m, o->nr - 2);
    if (m < 0) m = 0; /* reference: undefined for negative density */
    p -= m;
    p = MINV(p, 1.0);
    double p2 = p * p, p3 = p2 * p, p4 = p3 * p;
    const double* coef = o->F_arr + ((long)type * (o->nrho + 1) + m) * o->ncF;
    double tmp0 = coef[4] * p4 + coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
    double tmp1 = coef[8] * p3 + coef[7] * p2 + coef[6] * p + coef[5];
    if (rho_bar > o->rho_max) tmp0 += tmp1 * (rho_bar - o->rho_max);
    *F = tmp0;
    *dF = tmp1;
    if (ddF) *ddF = (3.0 * coef[8] * p2 + coef[7] * p + coef[6]) * o->drho_inv;
}

/* Gauss-Hermite averaged rho and r*phi primitives: src/ff_dmd_I.cpp:180-242 */
static int quadrature(const orc_t* o, double r, double r_inv, double alpha, int rslot, int pslot,
                      int want_phi, double* g, double* f) {
    double alpha_sq = sqrt(alpha);
    double inv_alph_sq = 1.0 / alpha_sq;
    double upper = (r + o->rc) * alpha_sq;
    double lower = (r - o->rc) * alpha_sq;
    if (!(lower < o->xi[o->no_i - 1])) return 0;
    double f_0 = 0.0, f_1 = 0.0, f_2 = 0.0, g_0 = 0.0, g_1 = 0.0, g_2 = 0.0;
    for (int i = 0; i < o->no_i; i++) {
        if (o->xi[i] > lower && o->xi[i] < upper) {
            double tmp0 = r - o->xi[i] * inv_alph_sq;
            double p = fabs(tmp0) * o->dr_inv;
            int m = (int)p;
            m = MINV(m, o->nr - 2);
            p -= m;
            p = MINV(p, 1.0);
            double p2 = p * p, p3 = p2 * p;
            const double* coef = o->rho_arr + ((long)rslot * o->nr + m) * 7;
            double tmp2 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
            tmp2 *= tmp0;
            g_0 += o->wi_0[i] * tmp2;
            g_1 += o->wi_1[i] * tmp2;
            g_2 += o->wi_2[i] * tmp2;
            if (want_phi) {
                coef = o->phi_arr + ((long)pslot * o->nr + m) * 7;
                double tmp1 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                if (tmp0 < 0.0) tmp1 *= -1.0;
                f_0 += o->wi_0[i] * tmp1;
                f_1 += o->wi_1[i] * tmp1;
                f_2 += o->wi_2[i] * tmp1;
            }
        }
    }
    if (want_phi) {
        f_0 *= PI_IN_SQ * r_inv;
        f_1 *= PI_IN_SQ * r_inv;
        f_2 *= PI_IN_SQ * r_inv;
        f[0] = f_0;
        f[1] = -f_0 * r_inv - 2.0 * alpha_sq * f_1;
        f[2] = (0.5 * f_0 - f_2) / alpha;
    }
    g_0 *= PI_IN_SQ * r_inv;
    g_1 *= PI_IN_SQ * r_inv;
    g_2 *= PI_IN_SQ * r_inv;
    g[0] = g_0;
    g[1] = -g_0 * r_inv - 2.0 * alpha_sq * g_1;
    g[2] = (0.5 * g_0 - g_2) / alpha;
    return 1;
}

int orc_pair_quadrature(const orc_t* o, double r, double alpha, int rho_slot, int phi_slot, double* out) {
    double g[3], f[3] = {0, 0, 0};
    int ok = quadrature(o, r, 1.0 / r, alpha, rho_slot, phi_slot, phi_slot >= 0, g, f);
    if (!ok) return 0;
    out[0] = g[0]; out[1] = g[1]; out[2] = g[2];
    out[3] = f[0]; out[4] = f[1]; out[5] = f[2];
    return 1;
}

static void force_calc_eam(orc_t* o, int st_clc, double* en_st);
static double energy_calc_eam(orc_t* o);
static void force_calc_lj(orc_t* o, int st_clc, double* en_st);
static double energy_calc_lj(orc_t* o);
static void force_calc_fs(orc_t* o, int st_clc, double* en_st);
static double energy_calc_fs(orc_t* o);
static void force_calc_fsm(orc_t* o, int st_clc, double* en_st);
static double energy_calc_fsm(orc_t* o);

/* ForceField_DMD_I::force_calc, src/ff_dmd_I.cpp:88-450 */
void orc_force_calc(orc_t* o, int st_clc, double* en_st) {
    if (o->variant == ORC_EAM) { force_calc_eam(o, st_clc, en_st); return; }
    if (o->variant == ORC_LJ) { force_calc_lj(o, st_clc, en_st); return; }
    if (o->variant == ORC_FS) { force_calc_fs(o, st_clc, en_st); return; }
    if (o->variant == ORC_FSM) { force_calc_fsm(o, st_clc, en_st); return; }
    if (o->variant == ORC_MEAM) { meam_force_calc(o, st_clc, en_st); return; }
    int nt = o->nt, xd = o->xdim, natms = o->natms, nphi = nt * (nt + 1) / 2;
    double* x = o->x;
    double* c = o->c;
    double* E = o->E;
    double* dE = o->dE;
    double* f = o->f;
    acc_t ns[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int kahan = o->compensated;
    ensure_caches(o);
    for (long i = 0; i < (long)natms * nt; i++) E[i] = 0.0;
    for (long i = 0; i < (long)natms * xd; i++) f[i] = 0.0;

    long phi_start = 0, rho_start = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)xd * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)xd * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (rsq < o->cut_sq_mod) {
                double r = sqrt(rsq);
                double r_inv = 1.0 / r;
                for (int it = 0; it < nt; it++)
                    for (int jt = 0; jt < nt; jt++) {
                        double ai = x[icomp + 3 + it], aj = x[jcomp + 3 + jt];
                        double alpha = (ai * aj) / (ai + aj);
                        int rs = type2rho(nt, it, jt), ps = comp(it, jt);
                        double cj = c[(long)jatm * nt + jt], ci = c[(long)iatm * nt + it];
                        double pairc = (c[(long)jatm * nt + jt] * c[(long)iatm * nt + it] +
                                        c[(long)jatm * nt + it] * c[(long)iatm * nt + jt]);
                        if (o->alpha_min < alpha && alpha < o->alpha_max) {
                            double g[3], fq[3];
                            if (quadrature(o, r, r_inv, alpha, rs, ps, it >= jt, g, fq)) {
                                if (it >= jt) {
                                    o->phi[phi_start + ps] = fq[0];
                                    o->dphi_dr[phi_start + ps] = fq[1];
                                    o->dphi_dalpha[phi_start + ps] = fq[2];
                                }
                                o->rho[rho_start + rs] = g[0];
                                o->drho_dr[rho_start + rs] = g[1];
                                o->drho_dalpha[rho_start + rs] = g[2];
                                E[(long)iatm * nt + it] += g[0] * cj;
                                if (jatm < natms) {
                                    E[(long)jatm * nt + jt] += g[0] * ci;
                                    if (it >= jt) acc_add(&ns[0], 0.5 * XW(o, it, jt) * pairc * fq[0], kahan);
                                } else {
                                    if (it >= jt) acc_add(&ns[0], 0.25 * XW(o, it, jt) * pairc * fq[0], kahan);
                                }
                            }
                        } else if (o->alpha_max <= alpha) {
                            /* classical branch, src/ff_dmd_I.cpp:261-310 */
                            if (rsq < o->cut_sq) {
                                double rr = sqrt(rsq);
                                double p = rr * o->dr_inv;
                                int m = (int)p;
                                m = MINV(m, o->nr - 2);
                                p -= m;
                                p = MINV(p, 1.0);
                                double p2 = p * p, p3 = p2 * p;
                                if (it >= jt) {
                                    const double* coef = o->phi_arr + ((long)ps * o->nr + m) * 7;
                                    double z2p = coef[6] * p2 + coef[5] * p + coef[4];
                                    double z2 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                                    o->phi[phi_start + ps] = z2 * r_inv;
                                    o->dphi_dalpha[phi_start + ps] = 0.0;
                                    o->dphi_dr[phi_start + ps] = (z2p - z2 * r_inv) * r_inv;
                                }
                                const double* coef = o->rho_arr + ((long)rs * o->nr + m) * 7;
                                o->rho[rho_start + rs] = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                                o->drho_dalpha[rho_start + rs] = 0.0;
                                o->drho_dr[rho_start + rs] = coef[6] * p2 + coef[5] * p + coef[4];
                                E[(long)iatm * nt + it] += o->rho[rho_start + rs] * cj;
                                if (jatm < natms) {
                                    E[(long)jatm * nt + jt] += o->rho[rho_start + rs] * ci;
                                    if (it >= jt) acc_add(&ns[0], 0.5 * XW(o, it, jt) * pairc * o->phi[phi_start + ps], kahan);
                                } else {
                                    if (it >= jt) acc_add(&ns[0], 0.25 * XW(o, it, jt) * pairc * o->phi[phi_start + ps], kahan);
                                }
                            }
                        }
                    }
            }
            phi_start += nphi;
            rho_start += nt * nt;
        }
        for (int it = 0; it < nt; it++) {
            long k = (long)iatm * nt + it;
            double Fv, dFv;
            embed(o, it, E[k], &Fv, &dFv, NULL);
            E[k] = Fv;
            dE[k] = dFv;
            acc_add(&ns[0], c[k] * E[k], kahan);
            acc_add(&ns[0], o->kbT * calc_ent(c[k]), kahan);
            acc_add(&ns[0], 1.5 * o->kbT * c[k] * log(x[icomp + 3 + it]) + c[k] * o->c_0[it], kahan);
            f[icomp + 3 + it] -= 1.5 * o->kbT * c[k] / x[icomp + 3 + it];
        }
    }

    update_vec(o, dE, nt);

    phi_start = 0;
    rho_start = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)xd * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)xd * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double r_inv = 1.0 / sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
            for (int it = 0; it < nt; it++)
                for (int jt = 0; jt < nt; jt++) {
                    long sij = rho_start + type2rho(nt, it, jt), sji = rho_start + type2rho(nt, jt, it);
                    long pp = phi_start + comp(it, jt);
                    double ci = c[(long)iatm * nt + it], cj = c[(long)jatm * nt + jt];
                    double dEi = dE[(long)iatm * nt + it], dEj = dE[(long)jatm * nt + jt];
                    double ai = x[icomp + 3 + it], aj = x[jcomp + 3 + jt];
                    double fpair = -(o->drho_dr[sji] * dEi + o->drho_dr[sij] * dEj + o->dphi_dr[pp]) * (ci * cj) * r_inv;
                    double apair = -(o->drho_dalpha[sji] * dEi + o->drho_dalpha[sij] * dEj + o->dphi_dalpha[pp]) *
                                   ci * cj / ((ai + aj) * (ai + aj));
                    if (apair != 0.0 || fpair != 0.0) {
                        f[icomp] += dx0 * fpair;
                        f[icomp + 1] += dx1 * fpair;
                        f[icomp + 2] += dx2 * fpair;
                        f[icomp + 3 + it] += apair * aj * aj;
                        if (jatm < natms) {
                            f[jcomp] -= dx0 * fpair;
                            f[jcomp + 1] -= dx1 * fpair;
                            f[jcomp + 2] -= dx2 * fpair;
                            f[jcomp + 3 + jt] += apair * ai * ai;
                            if (st_clc) {
                                acc_add(&ns[1], -(fpair * dx0 * dx0), kahan);
                                acc_add(&ns[2], -(fpair * dx1 * dx1), kahan);
                                acc_add(&ns[3], -(fpair * dx2 * dx2), kahan);
                                acc_add(&ns[4], -(fpair * dx1 * dx2), kahan);
                                acc_add(&ns[5], -(fpair * dx2 * dx0), kahan);
                                acc_add(&ns[6], -(fpair * dx0 * dx1), kahan);
                            }
                        } else if (st_clc) {
                            acc_add(&ns[1], -(0.5 * fpair * dx0 * dx0), kahan);
                            acc_add(&ns[2], -(0.5 * fpair * dx1 * dx1), kahan);
                            acc_add(&ns[3], -(0.5 * fpair * dx2 * dx2), kahan);
                            acc_add(&ns[4], -(0.5 * fpair * dx1 * dx2), kahan);
                            acc_add(&ns[5], -(0.5 * fpair * dx2 * dx0), kahan);
                            acc_add(&ns[6], -(0.5 * fpair * dx0 * dx1), kahan);
                        }
                    }
                }
            phi_start += nphi;
            rho_start += nt * nt;
        }
    }
    if (st_clc) for (int i = 0; i < 7; i++) en_st[i] = acc_value(&ns[i]);
    else en_st[0] = acc_value(&ns[0]);
}

/* ForceField_DMD_I::energy_calc, src/ff_dmd_I.cpp:454-668 (jtype from itype) */
double orc_energy_calc(orc_t* o) {
    if (o->variant == ORC_EAM) return energy_calc_eam(o);
    if (o->variant == ORC_LJ) return energy_calc_lj(o);
    if (o->variant == ORC_FS) return energy_calc_fs(o);
    if (o->variant == ORC_FSM) return energy_calc_fsm(o);
    if (o->variant == ORC_MEAM) return meam_energy_calc(o);
    int nt = o->nt, xd = o->xdim, natms = o->natms;
    double* x = o->x;
    double* c = o->c;
    double* E = o->E;
    acc_t en = {0, 0};
    const int kahan = o->compensated;
    for (long i = 0; i < (long)natms * nt; i++) E[i] = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)xd * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)xd * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (!(rsq < o->cut_sq_mod)) continue;
            double r = sqrt(rsq);
            double r_inv = 1.0 / r;
            for (int it = 0; it < nt; it++)
                for (int jt = o->energy_full ? 0 : it; jt < nt; jt++) {
                    double ai = x[icomp + 3 + it], aj = x[jcomp + 3 + jt];
                    double alpha = (ai * aj) / (ai + aj);
                    int rs = type2rho(nt, it, jt), ps = comp(it, jt);
                    double cj = c[(long)jatm * nt + jt], ci = c[(long)iatm * nt + it];
                    double pairc = (c[(long)jatm * nt + jt] * c[(long)iatm * nt + it] +
                                    c[(long)jatm * nt + it] * c[(long)iatm * nt + jt]);
                    if (o->alpha_min < alpha && alpha < o->alpha_max) {
                        double alpha_sq = sqrt(alpha);
                        double inv_alph_sq = 1.0 / alpha_sq;
                        double upper = (r + o->rc) * alpha_sq;
                        double lower = (r - o->rc) * alpha_sq;
                        double f_0 = 0.0, g_0 = 0.0;
                        if (lower < o->xi[o->no_i - 1]) {
                            for (int i = 0; i < o->no_i; i++) {
                                if (o->xi[i] > lower && o->xi[i] < upper) {
                                    double tmp0 = r - o->xi[i] * inv_alph_sq;
                                    double p = fabs(tmp0) * o->dr_inv;
                                    int m = (int)p;
                                    m = MINV(m, o->nr - 2);
                                    p -= m;
                                    p = MINV(p, 1.0);
                                    double p2 = p * p, p3 = p2 * p;
                                    if (it >= jt) {
                                        const double* coef = o->phi_arr + ((long)ps * o->nr + m) * 7;
                                        double tmp1 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                                        if (tmp0 < 0.0) tmp1 *= -1.0;
                                        f_0 += o->wi_0[i] * tmp1;
                                    }
                                    const double* coef = o->rho_arr + ((long)rs * o->nr + m) * 7;
                                    double tmp2 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                                    tmp2 *= tmp0;
                                    g_0 += o->wi_0[i] * tmp2;
                                }
                            }
                            if (it >= jt) f_0 *= PI_IN_SQ * r_inv;
                            g_0 *= PI_IN_SQ * r_inv;
                            E[(long)iatm * nt + it] += g_0 * cj;
                            if (jatm < natms) {
                                E[(long)jatm * nt + jt] += g_0 * ci;
                                if (it >= jt) acc_add(&en, 0.5 * XW(o, it, jt) * pairc * f_0, kahan);
                            } else {
                                if (it >= jt) acc_add(&en, 0.25 * XW(o, it, jt) * pairc * f_0, kahan);
                            }
                        }
                    } else if (o->alpha_max <= alpha) {
                        double f_0 = 0.0;
                        if (rsq < o->cut_sq) {
                            double rr = sqrt(rsq);
                            double p = rr * o->dr_inv;
                            int m = (int)p;
                            m = MINV(m, o->nr - 2);
                            p -= m;
                            p = MINV(p, 1.0);
                            double p2 = p * p, p3 = p2 * p;
                            if (it >= jt) {
                                const double* coef = o->phi_arr + ((long)ps * o->nr + m) * 7;
                                double z2 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                                f_0 = z2 * r_inv;
                            }
                            const double* coef = o->rho_arr + ((long)rs * o->nr + m) * 7;
                            double g_0 = coef[3] * p3 + coef[2] * p2 + coef[1] * p + coef[0];
                            E[(long)iatm * nt + it] += g_0 * cj;
                            if (jatm < natms) {
                                E[(long)jatm * nt + jt] += g_0 * ci;
                                if (it >= jt) acc_add(&en, 0.5 * XW(o, it, jt) * pairc * f_0, kahan);
                            } else {
                                if (it >= jt) acc_add(&en, 0.25 * XW(o, it, jt) * pairc * f_0, kahan);
                            }
                        }
                    }
                }
        }
        for (int it = 0; it < nt; it++) {
            long k = (long)iatm * nt + it;
            double Fv, dFv;
            embed(o, it, E[k], &Fv, &dFv, NULL);
            E[k] = Fv;
            acc_add(&en, c[k] * E[k], kahan);
            acc_add(&en, o->kbT * calc_ent(c[k]), kahan);
            acc_add(&en, 1.5 * o->kbT * c[k] * log(x[icomp + 3 + it]) + c[k] * o->c_0[it], kahan);
        }
    }
    return acc_value(&en);
}

/* ForceField_DMD_I::create_2nd_neigh_lst, src/ff_dmd_I.cpp:2138-2217 */
void orc_create_2nd_neigh_lst(orc_t* o) {
    int natms = o->natms, xd = o->xdim;
    free(o->nl2_off); free(o->nl2);
    o->nl2_off = (int*)malloc((size_t)(natms + 1) * sizeof(int));
    o->nl2 = (int*)malloc((size_t)(o->no_pairs + 1) * sizeof(int));
    long np = 0;
    for (int i = 0; i < natms + o->nghost; i++) o->crd[i] = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        o->nl2_off[iatm] = (int)np;
        long icomp = (long)xd * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)xd * jatm;
            double dx0 = o->x[icomp] - o->x[jcomp];
            double dx1 = o->x[icomp + 1] - o->x[jcomp + 1];
            double dx2 = o->x[icomp + 2] - o->x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (rsq < o->rsq_crd) {
                o->nl2[np++] = jatm;
                o->crd[iatm]++;
                if (jatm < natms) o->crd[jatm]++;
            }
        }
    }
    o->nl2_off[natms] = (int)np;
    o->no_pairs2 = np;
    update_int(o, o->crd);
}

/* E, F(E), dE (and ddE) from the cached rho slots -- the first sweep of
   c_d_calc (src/ff_dmd_I.cpp:2269-2312) and calc_g (:2447-2493) */
static void density_from_cache(orc_t* o, int want_dd) {
    int nt = o->nt, natms = o->natms;
    double *E = o->E, *c = o->c;
    long rho_start = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++)
                for (int jt = 0; jt < nt; jt++) {
                    E[icomp + it] += c[jcomp + jt] * o->rho[rho_start + type2rho(nt, jt, it)];
                    if (jatm < natms) E[jcomp + jt] += c[icomp + it] * o->rho[rho_start + type2rho(nt, it, jt)];
                }
            rho_start += nt * nt;
        }
        for (int it = 0; it < nt; it++) {
            double Fv, dFv, ddFv;
            embed(o, it, E[icomp + it], &Fv, &dFv, &ddFv);
            o->n[icomp + it] = E[icomp + it] = Fv;
            o->dE[icomp + it] = dFv;
            if (want_dd) o->ddE[icomp + it] = ddFv;
        }
    }
}

/* n += sum_j c_j (rho dE[jatm+jtype] + phi): src/ff_dmd_I.cpp:2317-2345 (index
   quirk dE[jatm+jtype] / dE[iatm+itype] kept) */
static void chem_pot_sweep(orc_t* o) {
    int nt = o->nt, natms = o->natms, nphi = nt * (nt + 1) / 2;
    double *n = o->n, *c = o->c, *dE = o->dE;
    long rho_start = 0, phi_start = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++)
                for (int jt = 0; jt < nt; jt++) {
                    double ph = o->phi[phi_start + comp(it, jt)];
                    long dj = o->fix_dE_index ? jcomp + jt : (long)jatm + jt;
                    long di = o->fix_dE_index ? icomp + it : (long)iatm + it;
                    n[icomp + it] += c[jcomp + jt] * (o->rho[rho_start + type2rho(nt, it, jt)] * dE[dj] + ph);
                    if (jatm < natms)
                        n[jcomp + jt] += c[icomp + it] * (o->rho[rho_start + type2rho(nt, jt, it)] * dE[di] + ph);
                }
            rho_start += nt * nt;
            phi_start += nphi;
        }
    }
    /* The reference closes n[iatm] (c_0, log(alpha), beta) right after the row
       of iatm; no later row writes n[iatm], so the caller closes all rows
       afterwards with identical results. */
}

/* c_d_calc for DMD / DMD_I / DMD_II */
void orc_c_d_calc(orc_t* o) {
    int nt = o->nt, natms = o->natms;
    double *c = o->c, *n = o->n, *c_d = o->c_d;
    for (long i = 0; i < (long)natms * nt; i++) c_d[i] = o->E[i] = o->dE[i] = 0.0;
    density_from_cache(o, 0);
    update_vec(o, o->dE, nt);
    chem_pot_sweep(o);
    for (int iatm = 0; iatm < natms; iatm++)
        for (int it = 0; it < nt; it++) {
            long k = (long)nt * iatm + it;
            double a = o->x[(long)o->xdim * iatm + 3 + it];
            if (o->variant == ORC_DMD_II) {
                n[k] += o->c_0[it] + 1.5 * o->kbT * log(a);
            } else {
                n[k] += o->c_0[it] + 1.5 * o->kbT * log(a);
                n[k] *= o->beta;
            }
        }
    update_vec(o, n, nt);
    double Mat = o->variant == ORC_DMD ? 5.2 * o->beta : 0.4 * o->beta;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl2_off[iatm]; jj < o->nl2_off[iatm + 1]; jj++) {
            int jatm = o->nl2[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++) {
                double s_ij;
                double ci = c[icomp + it], cj = c[jcomp + it];
                if (o->variant == ORC_DMD_II) {
                    double fi = n[icomp + it], fj = n[jcomp + it];
                    double mt = fi > fj ? fi + 0.5 : fj + 0.5;
                    double exp_fi = exp(o->beta * (fi - mt));
                    double exp_fj = exp(o->beta * (fj - mt));
                    double tmpi = ci * (1.0 - cj);
                    double tmpj = cj * (1.0 - ci);
                    s_ij = tmpj * exp_fj - tmpi * exp_fi;
                } else {
                    double tmp0, tmp1;
                    if (o->variant == ORC_DMD_I) {
                        tmp0 = exp(0.5 * (n[icomp + it] - n[jcomp + it]) + Mat);
                        tmp1 = exp(0.5 * (n[jcomp + it] - n[icomp + it]) + Mat);
                    } else {
                        tmp0 = exp(n[icomp + it] + Mat);
                        tmp1 = exp(n[jcomp + it] + Mat);
                    }
                    s_ij = ((1.0 - ci) * cj * tmp1 - (1.0 - cj) * ci * tmp0);
                }
                c_d[icomp + it] += s_ij;
                if (jatm < natms) c_d[jcomp + it] -= s_ij;
            }
        }
    }
}

/* calc_g for DMD / DMD_I / DMD_II */
double orc_calc_g(orc_t* o, int chk, double alpha, const double* a, double* g) {
    int nt = o->nt, natms = o->natms, nphi = nt * (nt + 1) / 2;
    double *c = o->c, *n = o->n, *s = o->s, *t = o->t, *v = o->v, *dE = o->dE, *ddE = o->ddE, *c_d = o->c_d;
    long N = (long)natms * nt;
    for (long i = 0; i < N; i++) g[i] = o->E[i] = 0.0;
    density_from_cache(o, 1);
    update_vec(o, dE, nt);
    chem_pot_sweep(o);
    for (int iatm = 0; iatm < natms; iatm++)
        for (int it = 0; it < nt; it++) {
            long k = (long)nt * iatm + it;
            double av = o->x[(long)o->xdim * iatm + 3 + it];
            if (o->variant == ORC_DMD_II) {
                n[k] += o->c_0[it];
                n[k] += 1.5 * o->kbT * log(av);
            } else {
                n[k] += o->c_0[it];
                n[k] *= o->beta;
                n[k] += 1.5 * log(av);
            }
        }
    update_vec(o, n, nt);
    for (long i = 0; i < N; i++) {
        s[i] = 2.0 * (c[i] + a[i]);
        if (o->variant == ORC_DMD_II) c_d[i] = 0.0;
    }
    double Mat = o->variant == ORC_DMD ? 5.2 * o->beta : 0.4 * o->beta;
    double inner = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl2_off[iatm]; jj < o->nl2_off[iatm + 1]; jj++) {
            int jatm = o->nl2[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++) {
                double ci = c[icomp + it], cj = c[jcomp + it];
                if (o->variant == ORC_DMD_II) {
                    double fi = n[icomp + it], fj = n[jcomp + it];
                    double mt = fi > fj ? fi + 0.5 : fj + 0.5;
                    double exp_fi = exp(o->beta * (fi - mt));
                    double exp_fj = exp(o->beta * (fj - mt));
                    double tmpi = ci * (1.0 - cj);
                    double tmpj = cj * (1.0 - ci);
                    double s_ij = tmpj * exp_fj - tmpi * exp_fi;
                    s[icomp + it] -= 2.0 * alpha * s_ij;
                    c_d[icomp + it] += s_ij;
                    if (jatm < natms) {
                        s[jcomp + it] += 2.0 * alpha * s_ij;
                        c_d[jcomp + it] -= s_ij;
                    }
                } else {
                    double tmp0, tmp1;
                    if (o->variant == ORC_DMD_I) {
                        tmp0 = exp(0.5 * (n[icomp + it] - n[jcomp + it]) + Mat);
                        tmp1 = exp(0.5 * (n[jcomp + it] - n[icomp + it]) + Mat);
                    } else {
                        tmp0 = exp(n[icomp + it] + Mat);
                        tmp1 = exp(n[jcomp + it] + Mat);
                    }
                    double s_ij = -2.0 * alpha * ((1.0 - ci) * cj * tmp1 - (1.0 - cj) * ci * tmp0);
                    s[icomp + it] += s_ij;
                    if (jatm < natms) s[jcomp + it] -= s_ij;
                }
            }
        }
        for (int it = 0; it < nt; it++) inner += s[icomp + it] * s[icomp + it];
    }
    double ans = inner;
    ans *= 0.25;
    if (chk) return ans;

    update_vec(o, s, nt);
    for (long i = 0; i < N; i++) {
        v[i] = t[i] = 0.0;
        g[i] = -s[i];
    }
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl2_off[iatm]; jj < o->nl2_off[iatm + 1]; jj++) {
            int jatm = o->nl2[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++) {
                double ci = c[icomp + it], cj = c[jcomp + it];
                if (o->variant == ORC_DMD_II) {
                    double s_ij = s[icomp + it] - s[jcomp + it];
                    double fi = n[icomp + it], fj = n[jcomp + it];
                    double mt = fi > fj ? fi + 0.5 : fj + 0.5;
                    double exp_fi = exp(o->beta * (fi - mt));
                    double exp_fj = exp(o->beta * (fj - mt));
                    double tmpi = ci * (1.0 - cj);
                    double tmpj = cj * (1.0 - ci);
                    double dm_ij = fi > fj ? 1.0 : 0.0;
                    double ti = tmpi * exp_fi + dm_ij * (tmpj * exp_fj - tmpi * exp_fi);
                    ti *= s_ij;
                    t[icomp + it] += ti;
                    g[icomp + it] -= alpha * ((1.0 - cj) * exp_fi + cj * exp_fj) * s_ij;
                    if (jatm < natms) {
                        double dm_ji = fj > fi ? 1.0 : 0.0;
                        double tj = tmpj * exp_fj + dm_ji * (tmpi * exp_fi - tmpj * exp_fj);
                        tj *= s_ij;
                        t[jcomp + it] -= tj;
                        g[jcomp + it] += alpha * ((1.0 - ci) * exp_fj + ci * exp_fi) * s_ij;
                    }
                } else if (o->variant == ORC_DMD_I) {
                    double tmp0 = exp(0.5 * (n[icomp + it] - n[jcomp + it]) + Mat);
                    double tmp1 = exp(0.5 * (n[jcomp + it] - n[icomp + it]) + Mat);
                    tmp0 *= (s[icomp + it] - s[jcomp + it]);
                    tmp1 *= (s[icomp + it] - s[jcomp + it]);
                    double t_ij = 0.5 * (ci * (1.0 - cj) * tmp0 + cj * (1.0 - ci) * tmp1);
                    t[icomp + it] += t_ij;
                    g[icomp + it] -= alpha * (tmp0 * (1.0 - cj) + tmp1 * cj);
                    if (jatm < natms) {
                        t[jcomp + it] -= t_ij;
                        g[jcomp + it] += alpha * (tmp1 * (1.0 - ci) + ci * tmp0);
                    }
                } else {
                    double tmp0 = (s[icomp + it] - s[jcomp + it]) * exp(Mat + n[icomp + it]);
                    double tmp1 = (s[icomp + it] - s[jcomp + it]) * exp(Mat + n[jcomp + it]);
                    t[icomp + it] += ci * tmp0 * (1.0 - cj);
                    g[icomp + it] -= alpha * (tmp0 * (1.0 - cj) + tmp1 * cj);
                    if (jatm < natms) {
                        t[jcomp + it] -= cj * tmp1 * (1.0 - ci);
                        g[jcomp + it] += alpha * (tmp1 * (1.0 - ci) + ci * tmp0);
                    }
                }
            }
        }
    }
    update_vec(o, t, nt);

    long rho_start = 0, phi_start = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++)
                for (int jt = 0; jt < nt; jt++) {
                    double rji = o->rho[rho_start + type2rho(nt, jt, it)];
                    double rij = o->rho[rho_start + type2rho(nt, it, jt)];
                    double tmp0 = o->phi[phi_start + comp(jt, it)] + rji * dE[icomp + it] + rij * dE[jcomp + jt];
                    tmp0 *= o->beta * alpha;
                    v[icomp + it] += rji * t[jcomp + jt];
                    g[icomp + it] += tmp0 * t[jcomp + jt];
                    if (jatm < natms) {
                        v[jcomp + jt] += rij * t[icomp + it];
                        g[jcomp + jt] += tmp0 * t[icomp + it];
                    }
                }
            rho_start += nt * nt;
            phi_start += nphi;
        }
    }
    update_vec(o, v, nt);
    update_vec(o, ddE, nt);
    rho_start = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = (long)nt * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            long jcomp = (long)nt * jatm;
            for (int it = 0; it < nt; it++)
                for (int jt = 0; jt < nt; jt++) {
                    g[icomp + it] += alpha * o->beta * o->rho[rho_start + type2rho(nt, it, jt)] * ddE[jcomp + jt] *
                                     c[jcomp + jt] * v[jcomp + jt];
                    if (jatm < natms)
                        g[jcomp + jt] += alpha * o->beta * o->rho[rho_start + type2rho(nt, jt, it)] * ddE[icomp + it] *
                                         c[icomp + it] * v[icomp + it];
                }
            rho_start += nt * nt;
        }
    }
    return ans;
}

/* ------------------------------------------------------------------- EAM */

/* ForceField_eam::force_calc, src/ff_eam.cpp:44-264 (c holds the atom type) */
static void force_calc_eam(orc_t* o, int st_clc, double* en_st) {
    int natms = o->natms, nt = o->nt;
    double* x = o->x;
    double* f = o->f;
    double* rho = o->E; /* per-atom rho, then F' */
    acc_t ns[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int kahan = o->compensated;
    long np = o->no_pairs;
    double* drhoi = (double*)calloc((size_t)np + 1, sizeof(double));
    double* drhoj = (double*)calloc((size_t)np + 1, sizeof(double));
    for (int i = 0; i < natms; i++) rho[i] = 0.0;
    for (long i = 0; i < (long)natms * 3; i++) f[i] = 0.0;
    long istart = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        long icomp = 3L * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            long jcomp = 3L * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            drhoi[istart] = drhoj[istart] = 0.0;
            if (rsq < o->cut_sq) {
                double r = sqrt(rsq);
                double r_inv = 1.0 / r;
                double p = r * o->dr_inv;
                int m = (int)p;
                m = MINV(m, o->nr - 2);
                p -= m;
                p = MINV(p, 1.0);
                const double* coef = o->rho_arr + ((long)type2rho(nt, jtype, itype) * o->nr + m) * 7;
                double rho_i = ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
                double drho_i_dr = (coef[6] * p + coef[5]) * p + coef[4];
                coef = o->rho_arr + ((long)type2rho(nt, itype, jtype) * o->nr + m) * 7;
                double rho_j = ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
                double drho_j_dr = (coef[6] * p + coef[5]) * p + coef[4];
                coef = o->phi_arr + ((long)comp(itype, jtype) * o->nr + m) * 7;
                double z2p = (coef[6] * p + coef[5]) * p + coef[4];
                double z2 = ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
                double phi = z2 * r_inv;
                double dphi_dr = z2p * r_inv - phi * r_inv;
                rho[iatm] += rho_i;
                double fpair = -dphi_dr * r_inv;
                f[icomp] += fpair * dx0;
                f[icomp + 1] += fpair * dx1;
                f[icomp + 2] += fpair * dx2;
                if (jatm < natms) {
                    rho[jatm] += rho_j;
                    f[jcomp] -= fpair * dx0;
                    f[jcomp + 1] -= fpair * dx1;
                    f[jcomp + 2] -= fpair * dx2;
                    acc_add(&ns[0], phi, kahan);
                    if (st_clc) {
                        acc_add(&ns[1], -(fpair * dx0 * dx0), kahan);
                        acc_add(&ns[2], -(fpair * dx1 * dx1), kahan);
                        acc_add(&ns[3], -(fpair * dx2 * dx2), kahan);
                        acc_add(&ns[4], -(fpair * dx1 * dx2), kahan);
                        acc_add(&ns[5], -(fpair * dx2 * dx0), kahan);
                        acc_add(&ns[6], -(fpair * dx0 * dx1), kahan);
                    }
                } else {
                    acc_add(&ns[0], 0.5 * phi, kahan);
                    if (st_clc) {
                        acc_add(&ns[1], -(0.5 * fpair * dx0 * dx0), kahan);
                        acc_add(&ns[2], -(0.5 * fpair * dx1 * dx1), kahan);
                        acc_add(&ns[3], -(0.5 * fpair * dx2 * dx2), kahan);
                        acc_add(&ns[4], -(0.5 * fpair * dx1 * dx2), kahan);
                        acc_add(&ns[5], -(0.5 * fpair * dx2 * dx0), kahan);
                        acc_add(&ns[6], -(0.5 * fpair * dx0 * dx1), kahan);
                    }
                }
                drhoi[istart] = -drho_i_dr * r_inv;
                drhoj[istart] = -drho_j_dr * r_inv;
            }
            istart++;
        }
        double p = rho[iatm] * o->drho_inv;
        int m = (int)p;
        m = MINV(m, o->nr - 2);
        if (m < 0) m = 0;
        p -= m;
        p = MINV(p, 1.0);
        const double* coef = o->F_arr + ((long)itype * (o->nrho + 1) + m) * 7;
        double tmp1 = (coef[6] * p + coef[5]) * p + coef[4];
        double tmp0 = ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
        if (rho[iatm] > o->rho_max) tmp0 += tmp1 * (rho[iatm] - o->rho_max);
        acc_add(&ns[0], tmp0, kahan);
        rho[iatm] = tmp1;
    }
    update_vec(o, rho, 1);
    istart = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = 3L * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            if (drhoi[istart] != 0.0 || drhoj[istart] != 0.0) {
                int jatm = o->nl[jj];
                long jcomp = 3L * jatm;
                double fpair = rho[iatm] * drhoi[istart] + rho[jatm] * drhoj[istart];
                double dx0 = x[icomp] - x[jcomp];
                double dx1 = x[icomp + 1] - x[jcomp + 1];
                double dx2 = x[icomp + 2] - x[jcomp + 2];
                f[icomp] += dx0 * fpair;
                f[icomp + 1] += dx1 * fpair;
                f[icomp + 2] += dx2 * fpair;
                if (jatm < natms) {
                    f[jcomp] -= dx0 * fpair;
                    f[jcomp + 1] -= dx1 * fpair;
                    f[jcomp + 2] -= dx2 * fpair;
                    if (st_clc) {
                        acc_add(&ns[1], -(fpair * dx0 * dx0), kahan);
                        acc_add(&ns[2], -(fpair * dx1 * dx1), kahan);
                        acc_add(&ns[3], -(fpair * dx2 * dx2), kahan);
                        acc_add(&ns[4], -(fpair * dx1 * dx2), kahan);
                        acc_add(&ns[5], -(fpair * dx2 * dx0), kahan);
                        acc_add(&ns[6], -(fpair * dx0 * dx1), kahan);
                    }
                } else if (st_clc) {
                    acc_add(&ns[1], -(0.5 * fpair * dx0 * dx0), kahan);
                    acc_add(&ns[2], -(0.5 * fpair * dx1 * dx1), kahan);
                    acc_add(&ns[3], -(0.5 * fpair * dx2 * dx2), kahan);
                    acc_add(&ns[4], -(0.5 * fpair * dx1 * dx2), kahan);
                    acc_add(&ns[5], -(0.5 * fpair * dx2 * dx0), kahan);
                    acc_add(&ns[6], -(0.5 * fpair * dx0 * dx1), kahan);
                }
            }
            istart++;
        }
    }
    free(drhoi);
    free(drhoj);
    if (st_clc) for (int i = 0; i < 7; i++) en_st[i] = acc_value(&ns[i]);
    else en_st[0] = acc_value(&ns[0]);
}

/* ForceField_eam::energy_calc, src/ff_eam.cpp:268-356 */
static double energy_calc_eam(orc_t* o) {
    int natms = o->natms, nt = o->nt;
    double* x = o->x;
    double* rho = o->E;
    double en = 0.0;
    for (int i = 0; i < natms; i++) rho[i] = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        long icomp = 3L * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            long jcomp = 3L * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (rsq < o->cut_sq) {
                double r = sqrt(rsq);
                double p = r * o->dr_inv;
                int m = (int)p;
                m = MINV(m, o->nr - 2);
                p -= m;
                p = MINV(p, 1.0);
                const double* coef = o->phi_arr + ((long)comp(itype, jtype) * o->nr + m) * 7;
                double phi = (((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0]) / r;
                coef = o->rho_arr + ((long)type2rho(nt, jtype, itype) * o->nr + m) * 7;
                rho[iatm] += ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
                if (jatm < natms) {
                    coef = o->rho_arr + ((long)type2rho(nt, itype, jtype) * o->nr + m) * 7;
                    rho[jatm] += ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
                    en += phi;
                } else
                    en += 0.5 * phi;
            }
        }
        double p = rho[iatm] * o->drho_inv;
        int m = (int)p;
        m = MINV(m, o->nr - 2);
        if (m < 0) m = 0;
        p -= m;
        p = MINV(p, 1.0);
        const double* coef = o->F_arr + ((long)itype * (o->nrho + 1) + m) * 7;
        double tmp0 = ((coef[3] * p + coef[2]) * p + coef[1]) * p + coef[0];
        if (rho[iatm] > o->rho_max) tmp0 += ((coef[6] * p + coef[5]) * p + coef[4]) * (rho[iatm] - o->rho_max);
        en += tmp0;
    }
    return en;
}

/* ------------------------------------------------------------ accessors */

/* ------------------------------------------------------------------ L-BFGS */

/* LineSearch_BackTrack::line_min without box change (src/line_search.cpp:244-341;
   c = 0.4, rho = 0.5, alpha_max = 1, d_max = 1 from the constructor :214-224),
   LineSearch::energy (:117-196), inner_f_h (:32-45), normalize_h (:64-76).
   x = x_prev + a h over all (3+nt) columns, Atoms::update_0, energy_calc. */
enum { ORC_LS_S = 0, ORC_LS_F_DOWNHILL = 1, ORC_LS_F_GRAD0 = 2, ORC_LS_F_ALPHAMIN = 3,
       ORC_MIN_F_TOLERANCE = 4, ORC_MIN_F_MAX_ITER = 5 };

/* ForceField_lj::force_calc (src/ff_lj.cpp:160-264); f is zeroed first, as the
   MD drivers do (zero_f before every force_calc, src/md_nh.cpp:512,605) */
static void force_calc_lj(orc_t* o, int st_clc, double* en_st) {
    int natms = o->natms;
    double* x = o->x;
    double* f = o->f;
    acc_t ns[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int kahan = o->compensated;
    for (long i = 0; i < (long)natms * 3; i++) f[i] = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            int curs = comp(itype, jtype);
            if (!o->md_chk[curs]) continue;
            long icomp = 3L * iatm, jcomp = 3L * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            double csq = o->md_cut_sq[curs];
            if (rsq < csq) {
                double sig = o->lj_sig[curs], eps = o->lj_eps[curs];
                double sig2 = sig * sig / rsq;
                double sig6 = sig2 * sig2 * sig2;
                double sig12 = sig6 * sig6;
                double ft = 24.0 * eps * (2.0 * sig12 - sig6) / rsq;
                f[icomp] += ft * dx0;
                f[icomp + 1] += ft * dx1;
                f[icomp + 2] += ft * dx2;
                if (jatm < natms) {
                    acc_add(&ns[0], 4.0 * eps * (sig12 - sig6) + o->lj_offset[curs], kahan);
                    f[jcomp] -= ft * dx0;
                    f[jcomp + 1] -= ft * dx1;
                    f[jcomp + 2] -= ft * dx2;
                    if (st_clc) {
                        acc_add(&ns[1], -(ft * dx0 * dx0), kahan);
                        acc_add(&ns[2], -(ft * dx1 * dx1), kahan);
                        acc_add(&ns[3], -(ft * dx2 * dx2), kahan);
                        acc_add(&ns[4], -(ft * dx1 * dx2), kahan);
                        acc_add(&ns[5], -(ft * dx2 * dx0), kahan);
                        acc_add(&ns[6], -(ft * dx0 * dx1), kahan);
                    }
                } else {
                    acc_add(&ns[0], 2.0 * eps * (sig12 - sig6) + o->lj_offset[curs] * 0.5, kahan);
                    if (st_clc) {
                        acc_add(&ns[1], -(0.5 * ft * dx0 * dx0), kahan);
                        acc_add(&ns[2], -(0.5 * ft * dx1 * dx1), kahan);
                        acc_add(&ns[3], -(0.5 * ft * dx2 * dx2), kahan);
                        acc_add(&ns[4], -(0.5 * ft * dx1 * dx2), kahan);
                        acc_add(&ns[5], -(0.5 * ft * dx2 * dx0), kahan);
                        acc_add(&ns[6], -(0.5 * ft * dx0 * dx1), kahan);
                    }
                }
            }
        }
    }
    en_st[0] = acc_value(&ns[0]);
    if (st_clc)
        for (int k = 1; k < 7; k++) en_st[k] = acc_value(&ns[k]);
}

/* ForceField_lj::energy_calc (src/ff_lj.cpp:270-341) */
static double energy_calc_lj(orc_t* o) {
    int natms = o->natms;
    double* x = o->x;
    double en = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            int curs = comp(itype, jtype);
            if (!o->md_chk[curs]) continue;
            long icomp = 3L * iatm, jcomp = 3L * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (rsq < o->md_cut_sq[curs]) {
                double sig = o->lj_sig[curs], eps = o->lj_eps[curs];
                double sig2 = sig * sig / rsq;
                double sig6 = sig2 * sig2 * sig2;
                double sig12 = sig6 * sig6;
                if (jatm < natms) en += 4.0 * eps * (sig12 - sig6) + o->lj_offset[curs];
                else en += 2.0 * eps * (sig12 - sig6) + o->lj_offset[curs] * 0.5;
            }
        }
    }
    return en;
}

/* ForceField_fs::force_calc (src/ff_fs.cpp:379-604): rho[i] is complete when
   i's own row ends (lower sites added their shares before), so the embedding
   energy and rho[i] = -A/sqrt(rho) are taken there; atoms->update(rho_n)
   forwards it to the ghosts before the density-force sweep */
static void force_calc_fs(orc_t* o, int st_clc, double* en_st) {
    int natms = o->natms;
    double* x = o->x;
    double* f = o->f;
    double* rho = o->E;
    acc_t ns[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int kahan = o->compensated;
    long np = o->no_pairs;
    double* drhoi = (double*)calloc((size_t)np + 1, sizeof(double));
    double* drhoj = (double*)calloc((size_t)np + 1, sizeof(double));
    for (long i = 0; i < (long)natms * 3; i++) f[i] = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) rho[iatm] = 0.0;
    long istart = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        long icomp = 3L * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            int curs = comp(itype, jtype);
            drhoi[istart] = drhoj[istart] = 0.0;
            if (o->md_chk[curs]) {
                long jcomp = 3L * jatm;
                double dx0 = x[icomp] - x[jcomp];
                double dx1 = x[icomp + 1] - x[jcomp + 1];
                double dx2 = x[icomp + 2] - x[jcomp + 2];
                double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
                if (rsq < o->md_cut_sq[curs]) {
                    double r = sqrt(rsq);
                    double r_inv = 1.0 / r;
                    if (r < o->fs_cut_rho[curs]) {
                        double dr_rho = r - o->fs_cut_rho[curs];
                        rho[iatm] += dr_rho * dr_rho * (o->fs_t1[jtype][itype] + o->fs_t2[jtype][itype] * dr_rho);
                        drhoi[istart] = -r_inv * dr_rho * (o->fs_t1[jtype][itype] + 1.5 * o->fs_t2[jtype][itype] * dr_rho);
                        drhoj[istart] = -r_inv * dr_rho * (o->fs_t1[itype][jtype] + 1.5 * o->fs_t2[itype][jtype] * dr_rho);
                        if (jatm < natms)
                            rho[jatm] += dr_rho * dr_rho * (o->fs_t1[itype][jtype] + o->fs_t2[itype][jtype] * dr_rho);
                    }
                    if (r < o->fs_cut_phi[curs]) {
                        double dr_phi = r - o->fs_cut_phi[curs];
                        double k1 = o->fs_k1[curs], k2 = o->fs_k2[curs], k3 = o->fs_k3[curs];
                        double phi_coef = 2.0 * dr_phi * (k1 + k2 * r + k3 * rsq) + dr_phi * dr_phi * (k2 + 2.0 * k3 * r);
                        phi_coef *= -r_inv;
                        f[icomp] += dx0 * phi_coef;
                        f[icomp + 1] += dx1 * phi_coef;
                        f[icomp + 2] += dx2 * phi_coef;
                        if (jatm < natms) {
                            f[jcomp] -= dx0 * phi_coef;
                            f[jcomp + 1] -= dx1 * phi_coef;
                            f[jcomp + 2] -= dx2 * phi_coef;
                            acc_add(&ns[0], dr_phi * dr_phi * (k1 + k2 * r + k3 * rsq), kahan);
                            if (st_clc) {
                                acc_add(&ns[1], -(phi_coef * dx0 * dx0), kahan);
                                acc_add(&ns[2], -(phi_coef * dx1 * dx1), kahan);
                                acc_add(&ns[3], -(phi_coef * dx2 * dx2), kahan);
                                acc_add(&ns[4], -(phi_coef * dx1 * dx2), kahan);
                                acc_add(&ns[5], -(phi_coef * dx2 * dx0), kahan);
                                acc_add(&ns[6], -(phi_coef * dx0 * dx1), kahan);
                            }
                        } else {
                            acc_add(&ns[0], 0.5 * dr_phi * dr_phi * (k1 + k2 * r + k3 * rsq), kahan);
                            if (st_clc) {
                                acc_add(&ns[1], -(0.5 * phi_coef * dx0 * dx0), kahan);
                                acc_add(&ns[2], -(0.5 * phi_coef * dx1 * dx1), kahan);
                                acc_add(&ns[3], -(0.5 * phi_coef * dx2 * dx2), kahan);
                                acc_add(&ns[4], -(0.5 * phi_coef * dx1 * dx2), kahan);
                                acc_add(&ns[5], -(0.5 * phi_coef * dx2 * dx0), kahan);
                                acc_add(&ns[6], -(0.5 * phi_coef * dx0 * dx1), kahan);
                            }
                        }
                    }
                }
            }
            istart++;
        }
        double rho_sqd = sqrt(rho[iatm]);
        acc_add(&ns[0], -(o->fs_A[itype] * rho_sqd), kahan);
        rho[iatm] = -o->fs_A[itype] / rho_sqd;
    }
    update_vec(o, rho, 1);
    istart = 0;
    for (int iatm = 0; iatm < natms; iatm++) {
        long icomp = 3L * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            if (drhoi[istart] != 0.0 || drhoj[istart] != 0.0) {
                int jatm = o->nl[jj];
                long jcomp = 3L * jatm;
                double dx0 = x[icomp] - x[jcomp];
                double dx1 = x[icomp + 1] - x[jcomp + 1];
                double dx2 = x[icomp + 2] - x[jcomp + 2];
                double rho_coef = drhoi[istart] * rho[iatm] + drhoj[istart] * rho[jatm];
                f[icomp] += dx0 * rho_coef;
                f[icomp + 1] += dx1 * rho_coef;
                f[icomp + 2] += dx2 * rho_coef;
                if (jatm < natms) {
                    f[jcomp] -= dx0 * rho_coef;
                    f[jcomp + 1] -= dx1 * rho_coef;
                    f[jcomp + 2] -= dx2 * rho_coef;
                    if (st_clc) {
                        acc_add(&ns[1], -(rho_coef * dx0 * dx0), kahan);
                        acc_add(&ns[2], -(rho_coef * dx1 * dx1), kahan);
                        acc_add(&ns[3], -(rho_coef * dx2 * dx2), kahan);
                        acc_add(&ns[4], -(rho_coef * dx1 * dx2), kahan);
                        acc_add(&ns[5], -(rho_coef * dx2 * dx0), kahan);
                        acc_add(&ns[6], -(rho_coef * dx0 * dx1), kahan);
                    }
                } else if (st_clc) {
                    acc_add(&ns[1], -(0.5 * rho_coef * dx0 * dx0), kahan);
                    acc_add(&ns[2], -(0.5 * rho_coef * dx1 * dx1), kahan);
                    acc_add(&ns[3], -(0.5 * rho_coef * dx2 * dx2), kahan);
                    acc_add(&ns[4], -(0.5 * rho_coef * dx1 * dx2), kahan);
                    acc_add(&ns[5], -(0.5 * rho_coef * dx2 * dx0), kahan);
                    acc_add(&ns[6], -(0.5 * rho_coef * dx0 * dx1), kahan);
                }
            }
            istart++;
        }
    }
    free(drhoi);
    free(drhoj);
    en_st[0] = acc_value(&ns[0]);
    if (st_clc)
        for (int k = 1; k < 7; k++) en_st[k] = acc_value(&ns[k]);
}

/* ForceField_fs::energy_calc (src/ff_fs.cpp:609-686) */
static double energy_calc_fs(orc_t* o) {
    int natms = o->natms;
    double* x = o->x;
    double* rho = o->E;
    double en = 0.0;
    for (int i = 0; i < natms; i++) rho[i] = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        long icomp = 3L * iatm;
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            int curs = comp(itype, jtype);
            if (!o->md_chk[curs]) continue;
            long jcomp = 3L * jatm;
            double dx0 = x[icomp] - x[jcomp];
            double dx1 = x[icomp + 1] - x[jcomp + 1];
            double dx2 = x[icomp + 2] - x[jcomp + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (rsq < o->md_cut_sq[curs]) {
                double r = sqrt(rsq);
                if (r < o->fs_cut_rho[curs]) {
                    double dr_rho = r - o->fs_cut_rho[curs];
                    rho[iatm] += dr_rho * dr_rho * (o->fs_t1[jtype][itype] + o->fs_t2[jtype][itype] * dr_rho);
                    if (jatm < natms)
                        rho[jatm] += dr_rho * dr_rho * (o->fs_t1[itype][jtype] + o->fs_t2[itype][jtype] * dr_rho);
                }
                if (r < o->fs_cut_phi[curs]) {
                    double dr_phi = r - o->fs_cut_phi[curs];
                    double k1 = o->fs_k1[curs], k2 = o->fs_k2[curs], k3 = o->fs_k3[curs];
                    if (jatm < natms) en += dr_phi * dr_phi * (k1 + k2 * r + k3 * rsq);
                    else en += 0.5 * dr_phi * dr_phi * (k1 + k2 * r + k3 * rsq);
                }
            }
        }
        en -= o->fs_A[itype] * sqrt(rho[iatm]);
    }
    return en;
}

/* virial share of a pair force coefficient c along d (w: 1, or 0.5 for a
   ghost partner) */
static void pair_virial(acc_t* ns, double w, double c, double d0, double d1, double d2, int kahan) {
    acc_add(&ns[1], -(w * c * d0 * d0), kahan);
    acc_add(&ns[2], -(w * c * d1 * d1), kahan);
    acc_add(&ns[3], -(w * c * d2 * d2), kahan);
    acc_add(&ns[4], -(w * c * d1 * d2), kahan);
    acc_add(&ns[5], -(w * c * d2 * d0), kahan);
    acc_add(&ns[6], -(w * c * d0 * d1), kahan);
}

/* the density sweep of ForceField_fsm (src/ff_fsm.cpp:407-434 in force_calc,
   :616-643 in energy_calc): shares inside cut_rho^2 with the half list's
   partner shares to local sites only; with en != NULL (energy_calc) also the
   pair energy inside cut_phi^2 (ghost partners at half weight) and the
   embedding energy -A sqrt(rho) once a site's row ends, when rho is complete
   (lower sites added their shares before) */
static void fsm_density(orc_t* o, double* rho, double* en) {
    int natms = o->natms;
    const double* x = o->x;
    for (int i = 0; i < natms; i++) rho[i] = 0.0;
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            int curs = comp(itype, jtype);
            if (!o->md_chk[curs]) continue;
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = x[3L * iatm + k] - x[3L * jatm + k];
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (en && !(rsq < o->md_cut_sq[curs])) continue;
            double r = sqrt(rsq);
            if (rsq < o->fs_cut_sq_rho[curs]) {
                double dr = r - sqrt(o->fs_cut_sq_rho[curs]);
                rho[iatm] += dr * dr * (o->fs_t1[jtype][itype] + o->fs_t2[jtype][itype] * dr);
                if (jatm < natms) rho[jatm] += dr * dr * (o->fs_t1[itype][jtype] + o->fs_t2[itype][jtype] * dr);
            }
            if (en && rsq < o->fs_cut_sq_phi[curs]) {
                double dp = r - sqrt(o->fs_cut_sq_phi[curs]);
                double e = dp * dp * (o->fs_k1[curs] + o->fs_k2[curs] * r + o->fs_k3[curs] * rsq);
                *en += jatm < natms ? e : 0.5 * e;
            }
        }
        if (en) *en += -o->fs_A[itype] * sqrt(rho[iatm]);
    }
}

/* ForceField_fsm::force_calc (src/ff_fsm.cpp:371-576): the density sweep,
   atoms->update(rho_n) (:436) forwards the raw densities to the ghosts, then
   one sweep takes the embedding energy of each site (:448) and per pair the
   density force A_i g_ji/sqrt(rho_i) + A_j g_ij/sqrt(rho_j) over r, with
   g = dr (t1 + 1.5 t2 dr) (:466-476), and the pair force and energy inside
   cut_phi^2 (:518-558); ghost partners take half the virial and energy */
static void force_calc_fsm(orc_t* o, int st_clc, double* en_st) {
    int natms = o->natms;
    const double* x = o->x;
    double* f = o->f;
    double* rho = o->E;
    acc_t ns[7] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const int kahan = o->compensated;
    for (long i = 0; i < (long)natms * 3; i++) f[i] = 0.0;
    fsm_density(o, rho, NULL);
    update_vec(o, rho, 1);
    for (int iatm = 0; iatm < natms; iatm++) {
        int itype = (int)o->c[iatm];
        acc_add(&ns[0], -o->fs_A[itype] * sqrt(rho[iatm]), kahan);
        for (int jj = o->nl_off[iatm]; jj < o->nl_off[iatm + 1]; jj++) {
            int jatm = o->nl[jj];
            int jtype = (int)o->c[jatm];
            int curs = comp(itype, jtype);
            if (!o->md_chk[curs]) continue;
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = x[3L * iatm + k] - x[3L * jatm + k];
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(rsq < o->md_cut_sq[curs])) continue;
            double r = sqrt(rsq);
            int local = jatm < natms;
            double w = local ? 1.0 : 0.5;
            double coef[2] = {0.0, 0.0};
            int have[2] = {0, 0};
            if (rsq < o->fs_cut_sq_rho[curs]) {
                double dr = r - sqrt(o->fs_cut_sq_rho[curs]);
                double c = o->fs_A[itype] * (dr * (1.0 * o->fs_t1[jtype][itype] + 1.5 * o->fs_t2[jtype][itype] * dr)) /
                           sqrt(rho[iatm]);
                c += o->fs_A[jtype] * (dr * (1.0 * o->fs_t1[itype][jtype] + 1.5 * o->fs_t2[itype][jtype] * dr)) /
                     sqrt(rho[jatm]);
                c *= 1.0 / r;
                coef[0] = c;
                have[0] = 1;
            }
            if (rsq < o->fs_cut_sq_phi[curs]) {
                double dp = r - sqrt(o->fs_cut_sq_phi[curs]);
                double k1 = o->fs_k1[curs], k2 = o->fs_k2[curs], k3 = o->fs_k3[curs];
                double c = 2.0 * dp * (k1 + k2 * r + k3 * rsq) + dp * dp * (k2 + 2.0 * k3 * r);
                c *= -1.0 / r;
                coef[1] = c;
                have[1] = 1;
                double e = dp * dp * (k1 + k2 * r + k3 * rsq);
                acc_add(&ns[0], local ? e : 0.5 * e, kahan);
            }
            for (int t = 0; t < 2; t++) {
                if (!have[t]) continue;
                for (int k = 0; k < 3; k++) {
                    f[3L * iatm + k] += d[k] * coef[t];
                    if (local) f[3L * jatm + k] -= d[k] * coef[t];
                }
                if (st_clc) pair_virial(ns, w, coef[t], d[0], d[1], d[2], kahan);
            }
        }
    }
    en_st[0] = acc_value(&ns[0]);
    if (st_clc)
        for (int k = 1; k < 7; k++) en_st[k] = acc_value(&ns[k]);
}

/* ForceField_fsm::energy_calc (src/ff_fsm.cpp:582-664) */
static double energy_calc_fsm(orc_t* o) {
    double en = 0.0;
    fsm_density(o, o->E, &en);
    return en;
}

static double ls_energy(orc_t* o, const double* x_prev, const double* h, double a, double* xbuf, long n) {
    for (long i = 0; i < n; i++) xbuf[i] = x_prev[i] + a * h[i];
    orc_set_x_update(o, xbuf, NULL);
    return orc_energy_calc(o);
}

static int line_min(orc_t* o, const double* x_prev, double* h, double* xbuf, long n, double* nrgy,
                    double* alph, int* n_energy) {
    const double c = 0.4, rho = 0.5, a_max = 1.0, d_max = 1.0;
    double a_min = 0.0;
    double min_h = INFINITY;
    for (int i = 0; i < o->natms; i++)
        for (int j = 0; j < o->nt; j++) {
            double v = o->x[(long)i * o->xdim + 3 + j];
            min_h = v < min_h ? v : min_h;
        }
    min_h *= 0.99;
    a_min = a_min < min_h ? a_min : min_h;
    *alph = 0.0;
    double hh = 0.0;
    for (long i = 0; i < n; i++) hh += h[i] * h[i];
    hh = d_max / sqrt(hh);
    for (long i = 0; i < n; i++) h[i] *= hh;
    double inner = 0.0;
    for (long i = 0; i < n; i++) inner += o->f[i] * h[i];
    if (inner <= 0) return ORC_LS_F_DOWNHILL;
    double max_h = 0.0;
    for (long i = 0; i < n; i++) max_h = fabs(h[i]) > max_h ? fabs(h[i]) : max_h;
    if (max_h == 0) return ORC_LS_F_GRAD0;
    double a = a_max < max_h ? a_max : max_h;
    if (a <= a_min) return ORC_LS_F_ALPHAMIN;
    while (1) {
        double ideal = *nrgy - a * c * inner;
        double cur = ls_energy(o, x_prev, h, a, xbuf, n);
        (*n_energy)++;
        if (cur <= ideal) {
            *nrgy = cur;
            *alph = a;
            return ORC_LS_S;
        }
        a *= rho;
        if (a <= a_min) {
            *nrgy = ls_energy(o, x_prev, h, 0.0, xbuf, n);
            (*n_energy)++;
            return ORC_LS_F_ALPHAMIN;
        }
    }
}

/* Min_LBFGS::init tail + Min_LBFGS::run, fixed box (src/min_l-bfgs.cpp:274-289,
   :302-323, :606-787): two-loop recursion over m_it pairs, ratio
   -<s,y>/<y,y>, y = f - f_prev, s = alpha_m h; stops on a failed line search,
   an energy decrease below e_tol or max_iter steps.  energies[k] is the
   force_calc energy after step k (energies[0]: initial); returns the exit code. */
int orc_min_lbfgs(orc_t* o, int max_iter, int m_it, double e_tol, double* energies, int* nsteps, int* n_force,
                  int* n_energy) {
    long n = (long)o->natms * o->xdim;
    double* h = (double*)calloc(n, sizeof(double));
    double* x0 = (double*)calloc(n, sizeof(double));
    double* f0 = (double*)calloc(n, sizeof(double));
    double* xb = (double*)calloc(n, sizeof(double));
    double* S = (double*)calloc((size_t)n * (m_it > 0 ? m_it : 1), sizeof(double));
    double* Y = (double*)calloc((size_t)n * (m_it > 0 ? m_it : 1), sizeof(double));
    int* sl = (int*)calloc(m_it > 0 ? m_it : 1, sizeof(int));
    double* rh = (double*)calloc(m_it > 0 ? m_it : 1, sizeof(double));
    double* al = (double*)calloc(m_it > 0 ? m_it : 1, sizeof(double));
    for (int i = 0; i < m_it; i++) sl[i] = i;
    double en[7];
    *n_force = 0;
    *n_energy = 0;
    orc_force_calc(o, 1, en);
    (*n_force)++;
    double curr = en[0], prev, ratio = 1.0, a_m = 0.0;
    energies[0] = curr;
    int k_it = 0, istp = 0, err = ORC_LS_S;
    while (err == ORC_LS_S) {
        memcpy(h, o->f, n * sizeof(double));
        memcpy(x0, o->x, n * sizeof(double));
        memcpy(f0, o->f, n * sizeof(double));
        for (int i = k_it - 1; i > -1; i--) {
            const double* s = S + (size_t)sl[i] * n;
            const double* y = Y + (size_t)sl[i] * n;
            double in = 0.0;
            for (long j = 0; j < n; j++) in += h[j] * s[j];
            al[i] = in * rh[i];
            for (long j = 0; j < n; j++) h[j] -= al[i] * y[j];
        }
        for (long j = 0; j < n; j++) h[j] *= ratio;
        for (int i = 0; i < k_it; i++) {
            const double* s = S + (size_t)sl[i] * n;
            const double* y = Y + (size_t)sl[i] * n;
            double in = 0.0;
            for (long j = 0; j < n; j++) in += h[j] * y[j];
            in *= rh[i];
            for (long j = 0; j < n; j++) h[j] -= (in + al[i]) * s[j];
        }
        prev = curr;
        err = line_min(o, x0, h, xb, n, &curr, &a_m, n_energy);
        if (err == ORC_LS_S && prev - curr < e_tol) err = ORC_MIN_F_TOLERANCE;
        if (err == ORC_LS_S && istp == max_iter) err = ORC_MIN_F_MAX_ITER;
        orc_force_calc(o, 1, en);
        (*n_force)++;
        curr = en[0];
        energies[istp + 1] = curr;
        if (err) break;
        double in = 0.0, in2 = 0.0;
        if (m_it) {
            if (k_it == m_it) {
                int t = sl[0];
                for (int i = 0; i < m_it - 1; i++) {
                    sl[i] = sl[i + 1];
                    rh[i] = rh[i + 1];
                }
                sl[m_it - 1] = t;
            } else {
                k_it++;
            }
            double* s = S + (size_t)sl[k_it - 1] * n;
            double* y = Y + (size_t)sl[k_it - 1] * n;
            for (long i = 0; i < n; i++) {
                s[i] = a_m * h[i];
                y[i] = o->f[i] - f0[i];
                in += s[i] * y[i];
                in2 += y[i] * y[i];
            }
            ratio = -in / in2;
            rh[k_it - 1] = 1.0 / in;
        } else {
            for (long i = 0; i < n; i++) {
                in += (a_m * h[i]) * (o->f[i] - f0[i]);
                in2 += (o->f[i] - f0[i]) * (o->f[i] - f0[i]);
            }
            ratio = -in / in2;
        }
        istp++;
    }
    *nsteps = istp;
    free(h); free(x0); free(f0); free(xb); free(S); free(Y); free(sl); free(rh); free(al);
    return err;
}


/* -------------------------------------------------------------------- BDF */

/* coefficients of Clock_BDF::interpolate (src/clock_bdf.cpp:295-341) for
   order q and the step to t[0] + del_t: out = {beta, alpha_dy_0,
   dalpha_dy_0, alpha_y[0..q-1], dalpha_y[0..q-1]}.  fix = 1 corrects the sign
   of the (t_new - t[0]) c1 c2 term of dalpha_y[0] (:327), which the reference
   adds instead of subtracting, so that the derivative weights sum to zero. */
void orc_bdf_coef(int q, double del_t, const double* t, int fix, double* out) {
    double alpha_y[8], dalpha_y[8];
    double t_new = t[0] + del_t, c0 = 1.0, c1 = 0.0, c2 = 0.0;
    for (int i = 1; i < q; i++) {
        double tmp0 = (t_new - t[0]) / (t[i] - t[0]);
        double tmp1 = tmp0 * tmp0;
        for (int j = 1; j < q; j++)
            if (i != j) tmp1 *= (t_new - t[j]) / (t[i] - t[j]);
        alpha_y[i] = tmp1;
        c0 *= (t_new - t[i]) / (t[0] - t[i]);
        c1 += 1.0 / (t[0] - t[i]);
        c2 += 1.0 / (t_new - t[i]);
    }
    alpha_y[0] = (1.0 - (t_new - t[0]) * c1) * c0;
    double alpha_dy_0 = c0 * (t_new - t[0]);
    double c3 = 2.0 / (t_new - t[0]);
    double dalpha_dy_0 = alpha_dy_0 * (c2 + 0.5 * c3);
    if (fix) dalpha_y[0] = c0 * (c2 - c1 - (t_new - t[0]) * c1 * c2);
    else dalpha_y[0] = c0 * (c2 - c1 + (t_new - t[0]) * c1 * c2);
    for (int i = 1; i < q; i++) dalpha_y[i] = alpha_y[i] * (c2 + c3 - 1.0 / (t_new - t[i]));
    double tmp0 = 0.0;
    for (int i = 0; i < q; i++) tmp0 += 1.0 / (double)(i + 1);
    out[0] = del_t / tmp0;
    out[1] = alpha_dy_0;
    out[2] = dalpha_dy_0;
    for (int i = 0; i < q; i++) {
        out[3 + i] = alpha_y[i];
        out[3 + q + i] = dalpha_y[i];
    }
}

/* Clock_BDF::interpolate (src/clock_bdf.cpp:295-369): the predictor y_0 and
   the constant part a of the residual; -1 when the predictor leaves [0,1]. */
static int bdf_interpolate(int q, double del_t, const double* t, double* const* y, const double* dy, long n,
                           double* a, double* y0, double* beta, int fix) {
    double cf[3 + 2 * 8];
    orc_bdf_coef(q, del_t, t, fix, cf);
    *beta = cf[0];
    const double alpha_dy_0 = cf[1], dalpha_dy_0 = cf[2];
    const double* alpha_y = cf + 3;
    const double* dalpha_y = cf + 3 + q;
    for (long k = 0; k < n; k++) {
        a[k] = *beta * dalpha_dy_0 * dy[k];
        y0[k] = alpha_dy_0 * dy[k];
        for (int j = 0; j < q; j++) {
            a[k] += *beta * dalpha_y[j] * y[j][k];
            y0[k] += alpha_y[j] * y[j][k];
        }
        a[k] -= y0[k];
        if (y0[k] < 0.0 || y0[k] > 1.0) return -1;
    }
    return 1;
}

/* Clock_BDF::solve (src/clock_bdf.cpp:677-883) with curr_t = t[0] + del_t:
   bounded predictor, then nonlinear conjugate gradients on calc_g's cost with a
   backtracking step that keeps c inside [0,1].  Returns cost / (m_tol dof) or
   -1 if c left [0,1]; *err is the local error estimate, *iters the CG count. */
/* The cells a BDF run integrates as one system: their dof concatenated in
   cell order, calc_g's cost summed over the cells, every dot product, norm and
   bound taken over the concatenation.  One cell is Clock_BDF itself; several
   are the batched run's semantics (one step size and order for the batch). */
typedef struct {
    orc_t** cell;
    int ncell;
    long n;      /* dof of all cells */
    long* off;   /* first dof of each cell */
    double* c;   /* current concentrations of all cells */
    double* c_d; /* current dc/dt of all cells */
} bsys_t;

static void bsys_set_c(bsys_t* S, const double* c) {
    for (int k = 0; k < S->ncell; k++) orc_set_c(S->cell[k], c + S->off[k]);
    memcpy(S->c, c, S->n * sizeof(double));
}

static void bsys_gather(bsys_t* S) {
    for (int k = 0; k < S->ncell; k++) {
        long m = S->off[k + 1] - S->off[k];
        memcpy(S->c + S->off[k], S->cell[k]->c, m * sizeof(double));
        memcpy(S->c_d + S->off[k], S->cell[k]->c_d, m * sizeof(double));
    }
}

static double bsys_calc_g(bsys_t* S, int chk, double alpha, const double* a, double* g, int* n_calc_g) {
    double cost = 0.0;
    for (int k = 0; k < S->ncell; k++) cost += orc_calc_g(S->cell[k], chk, alpha, a + S->off[k], g + S->off[k]);
    (*n_calc_g)++;
    return cost;
}

static double bdf_solve(bsys_t* S, int q, double del_t, const double* t, double* const* y, const double* dy,
                        const double* a, const double* y0, double beta, double* err, int* iters, int* n_calc_g) {
    const double min_gamma = 1.0e-10, gamma_red = 0.8, slope = 0.4, m_tol = 1.0e-10, a_tol = 1.0e-6;
    const int max_iter = 50;
    long n = S->n;
    double* c = (double*)calloc(n, sizeof(double));
    double* g = (double*)calloc(n, sizeof(double));
    double* g0 = (double*)calloc(n, sizeof(double));
    double* h = (double*)calloc(n, sizeof(double));
    double* cc0 = (double*)calloc(n, sizeof(double));
    double ratio = 1.0;
    for (long i = 0; i < n; i++) {
        double tmp = dy[i] * del_t + y[0][i];
        if (tmp > 1.0) ratio = MINV((1.0 - y[0][i]) / (tmp - y[0][i]), ratio);
        else if (tmp < 0.0) ratio = MINV((y[0][i] - 0.0) / (y[0][i] - tmp), ratio);
    }
    for (long i = 0; i < n; i++) c[i] = y[0][i] + dy[i] * del_t * ratio;
    bsys_set_c(S, c);
    double curr_cost = bsys_calc_g(S, 0, beta, a, g, n_calc_g);
    memcpy(h, g, n * sizeof(double));
    double g0_g0 = 0.0;
    for (long i = 0; i < n; i++) g0_g0 += g[i] * g[i];
    double g_h = g0_g0, max_gamma = 1.0;
    int iter = 0;
    while (curr_cost > m_tol * (double)n && iter < max_iter && max_gamma > min_gamma) {
        memcpy(g0, g, n * sizeof(double));
        memcpy(cc0, S->c, n * sizeof(double));
        double gamma = 0.99;
        for (long i = 0; i < n; i++) {
            double tmp = cc0[i] + h[i];
            if (tmp > 1.0) gamma = MINV((1.0 - cc0[i]) / (tmp - cc0[i]), gamma);
            else if (tmp < 0.0) gamma = MINV((cc0[i] - 0.0) / (cc0[i] - tmp), gamma);
        }
        max_gamma = gamma * 0.9;
        int chk = 1;
        double cost = curr_cost;
        while (chk && max_gamma > min_gamma) {
            for (long i = 0; i < n; i++) c[i] = cc0[i] + max_gamma * h[i];
            bsys_set_c(S, c);
            curr_cost = bsys_calc_g(S, 0, beta, a, g, n_calc_g);
            double ideal = cost - slope * max_gamma * g_h;
            if (curr_cost < ideal) chk = 0;
            max_gamma *= gamma_red;
            if (max_gamma <= min_gamma) {
                bsys_set_c(S, cc0);
                curr_cost = bsys_calc_g(S, 1, beta, a, g, n_calc_g);
            }
        }
        if (chk == 0) {
            double g_g0 = 0.0, g_g = 0.0;
            for (long i = 0; i < n; i++) g_g0 += g[i] * g0[i];
            for (long i = 0; i < n; i++) g_g += g[i] * g[i];
            double r = (g_g - g_g0) / g0_g0;
            g0_g0 = g_g;
            double inner = 0.0;
            for (long i = 0; i < n; i++) {
                h[i] *= r;
                h[i] += g[i];
                inner += h[i] * g[i];
            }
            g_h = inner;
            if (g_h < 0.0) {
                memcpy(h, g, n * sizeof(double));
                g_h = g_g;
            }
        }
        iter++;
    }
    *iters = iter;
    double curr_t = t[0] + del_t, a_1 = 1.0, l1_q = 0.0;
    for (int i = 0; i < q; i++) {
        l1_q += 1.0 / (curr_t - t[i]);
        if (i != 0) a_1 *= (curr_t - t[i]) / (t[0] - t[i]);
    }
    double err_pre_fac = -1.0 / (l1_q * (1.0 + a_1));
    double e = 0.0;
    for (long i = 0; i < n; i++) {
        double d = y0[i] - S->c[i];
        e += d * d;
    }
    e = sqrt(e / ((double)n * a_tol));
    *err = e * fabs(err_pre_fac);
    double ret = curr_cost / (m_tol * (double)n);
    for (long i = 0; i < n; i++)
        if (S->c[i] < 0.0 || S->c[i] > 1.0) ret = -1.0;
    free(c); free(g); free(g0); free(h); free(cc0);
    return ret;
}

/* Clock_BDF::hi_lo_err_est (src/clock_bdf.cpp:372-478): error estimates of
   the orders q+1 and q-1 for the step that starts at the current t[0]. */
typedef struct {
    double* e_n;
    double lwr_alpha[8];
    double lo_err, hi_err;
    int lo_avail, hi_avail;
} bdf_est_t;

static void bdf_hi_lo(bdf_est_t* E, int q, double del_t, const double* t, double* const* y, const double* c,
                      const double* c_d, const double* y0, long n) {
    const double a_tol = 1.0e-6;
    double curr_t = t[0] + del_t, a_0 = 1.0, a_1 = 1.0, l1_q = 0.0;
    double hi_a_0, hi_a_1, Q_n = 1.0, hi_err_pre_fac = 1.0, lo_err_pre_fac = 0.0, c0 = 0.0, c1 = 1.0;
    double lwr_alpha_dy = 0.0, lwr_alpha_y = 0.0;
    for (int i = 0; i < q; i++) {
        l1_q += 1.0 / (curr_t - t[i]);
        a_0 *= curr_t - t[i];
        if (i != 0) a_1 *= (curr_t - t[i]) / (t[0] - t[i]);
    }
    if (E->hi_avail) {
        hi_a_0 = hi_a_1 = 1.0;
        for (int i = 0; i < q; i++) {
            hi_a_0 *= t[0] - t[i + 1];
            if (i != 0) hi_a_1 *= (t[0] - t[i + 1]) / (t[1] - t[i + 1]);
        }
        Q_n = (a_0 / hi_a_0) * (1.0 + a_1) / (1.0 + hi_a_1) * (del_t / (t[0] - t[1]));
        double hi_l1_q = l1_q + 1.0 / (curr_t - t[q]);
        hi_err_pre_fac = -(curr_t - t[q]) / (q + 2);
        hi_err_pre_fac *= 1.0 / (hi_l1_q * (1.0 + a_1) * del_t * del_t);
    }
    if (E->lo_avail) {
        lo_err_pre_fac = -del_t * a_0 / (curr_t - t[q - 1]);
        c0 = 0.0;
        c1 = 1.0;
        for (int i = 0; i < q - 1; i++) {
            double tmp0 = 1.0 / (curr_t - t[i]);
            c0 += tmp0;
            c1 *= tmp0;
            E->lwr_alpha[i] = tmp0 * tmp0;
            for (int j = 0; j < q - 1; j++)
                if (i != j) E->lwr_alpha[i] *= 1.0 / (t[i] - t[j]);
            E->lwr_alpha[i] *= lo_err_pre_fac;
        }
        lwr_alpha_dy = c1 * lo_err_pre_fac;
        lwr_alpha_y = -c0 * c1 * lo_err_pre_fac;
    }
    double lo = 0.0, hi = 0.0;
    for (long i = 0; i < n; i++) {
        if (E->hi_avail) {
            double tmp0 = (c[i] - y0[i]) - Q_n * E->e_n[i];
            tmp0 *= hi_err_pre_fac;
            hi += tmp0 * tmp0;
        }
        if (E->lo_avail) {
            double tmp0 = lwr_alpha_y * c[i];
            tmp0 += lwr_alpha_dy * c_d[i];
            for (int j = 0; j < q - 1; j++) tmp0 += E->lwr_alpha[j] * y[j][i];
            tmp0 *= c1 / c0;
            lo += tmp0 * tmp0;
        }
        E->e_n[i] = c[i] - y0[i];
    }
    E->lo_err = sqrt(lo / ((double)n * a_tol));
    E->hi_err = sqrt(hi / ((double)n * a_tol));
}

/* Clock_BDF::init + Clock_BDF::run (src/clock_bdf.cpp:201-275, :479-672):
   variable step and order up to max_order, retries with a smaller step while
   the error estimate or the solver's cost ratio is >= 1, order changes from
   hi_lo_err_est.  initial_del_t <= 0 takes the reference's a_tol / rms(c_d)
   rule.  solve's uninitialised curr_t is taken as t[0] + del_t.  A step that
   needs more than max_retry retries ends the run (the reference would loop).
   report[step*6 + {0..5}] = order, del_t, cost ratio, error, CG iterations,
   solve calls.  Returns the number of steps, -1 if c left [0,1], -2 on too
   many retries. */
static int bdf_run_sys(bsys_t* S, int nsteps, int max_order, double initial_del_t, int max_retry, int fix,
                       double* report, int* n_calc_g) {
    const double a_tol = 1.0e-6, min_del_t = 1.0e-12, max_del_t = 1.0e4;
    long n = S->n;
    if (max_order < 1 || max_order > 7) return -3;
    double en[7];
    for (int k = 0; k < S->ncell; k++) {
        orc_force_calc(S->cell[k], 1, en);
        orc_create_2nd_neigh_lst(S->cell[k]);
        orc_c_d_calc(S->cell[k]);
    }
    bsys_gather(S);
    double t[8] = {0};
    double* y[8];
    for (int i = 0; i < max_order; i++) y[i] = (double*)calloc(n, sizeof(double));
    double* dy = (double*)calloc(n, sizeof(double));
    double* a = (double*)calloc(n, sizeof(double));
    double* y0 = (double*)calloc(n, sizeof(double));
    bdf_est_t E;
    memset(&E, 0, sizeof(E));
    E.e_n = (double*)calloc(n, sizeof(double));
    if (initial_del_t <= 0.0) {
        double sum = 0.0;
        for (long i = 0; i < n; i++) sum += S->c_d[i] * S->c_d[i];
        sum = sqrt(sum / (double)n);
        initial_del_t = a_tol / sum > min_del_t ? a_tol / sum : min_del_t;
        initial_del_t = initial_del_t < max_del_t ? initial_del_t : max_del_t;
    }
    memcpy(y[0], S->c, n * sizeof(double));
    memcpy(dy, S->c_d, n * sizeof(double));
    *n_calc_g = 0;
    double del_t = initial_del_t, beta = 0.0, err = 0.0, cost;
    int ord = 1, const_stp = 0, istep = 0, rc = 0;
    while (istep < nsteps) {
        int chk = bdf_interpolate(ord, del_t, t, y, dy, n, a, y0, &beta, fix);
        while (chk == -1) {
            del_t *= 0.8;
            del_t = del_t > min_del_t ? del_t : min_del_t;
            chk = bdf_interpolate(ord, del_t, t, y, dy, n, a, y0, &beta, fix);
        }
        int iters, it_sum = 0, solves = 1, retries = 0;
        cost = bdf_solve(S, ord, del_t, t, y, dy, a, y0, beta, &err, &iters, n_calc_g);
        it_sum += iters;
        while (cost >= 0 && (err >= 1.0 || cost >= 1.0)) {
            if (++retries > max_retry) {
                rc = -2;
                break;
            }
            double err1 = err > cost ? err : cost;
            double ratio = pow(0.5 / err1, 1.0 / (double)(ord + 1));
            ratio = ratio < 0.5 ? 0.5 : (ratio > 0.9 ? 0.9 : ratio);
            double tmp = del_t * ratio;
            del_t = tmp < min_del_t ? min_del_t : tmp;
            chk = bdf_interpolate(ord, del_t, t, y, dy, n, a, y0, &beta, fix);
            while (chk == -1) {
                del_t *= 0.8;
                chk = bdf_interpolate(ord, del_t, t, y, dy, n, a, y0, &beta, fix);
            }
            cost = bdf_solve(S, ord, del_t, t, y, dy, a, y0, beta, &err, &iters, n_calc_g);
            it_sum += iters;
            solves++;
        }
        report[istep * 6 + 0] = ord;
        report[istep * 6 + 1] = del_t;
        report[istep * 6 + 2] = cost;
        report[istep * 6 + 3] = err;
        report[istep * 6 + 4] = it_sum;
        report[istep * 6 + 5] = solves;
        if (cost < 0) rc = -1;
        if (rc) break;
        double* tmpy = y[max_order - 1];
        for (int i = max_order - 1; i > 0; i--) {
            t[i] = t[i - 1];
            y[i] = y[i - 1];
        }
        y[0] = tmpy;
        t[0] += del_t;
        bsys_gather(S);
        memcpy(y[0], S->c, n * sizeof(double));
        memcpy(dy, S->c_d, n * sizeof(double));
        E.lo_avail = ord != 1 && const_stp > ord + 1;
        E.hi_avail = ord < max_order && istep != 0 && const_stp > ord + 1;
        bdf_hi_lo(&E, ord, del_t, t, y, S->c, S->c_d, y0, n);
        double ratio, lo_ratio, hi_ratio;
        if (E.hi_avail && E.lo_avail) {
            lo_ratio = pow(0.5 / E.lo_err, 1.0 / (double)ord);
            hi_ratio = pow(0.5 / E.hi_err, 1.0 / (double)(ord + 2));
            ratio = pow(0.5 / err, 1.0 / (double)(ord + 1));
            if (hi_ratio > lo_ratio && hi_ratio > ratio) {
                ord++;
                ratio = hi_ratio;
            } else if (lo_ratio > hi_ratio && lo_ratio > ratio) {
                ord--;
                ratio = lo_ratio;
            }
        } else if (E.hi_avail && !E.lo_avail) {
            hi_ratio = pow(0.5 / E.hi_err, 1.0 / (double)(ord + 2));
            ratio = pow(0.5 / err, 1.0 / (double)(ord + 1));
            if (hi_ratio > ratio) {
                ord++;
                ratio = hi_ratio;
            }
        } else if (!E.hi_avail && E.lo_avail) {
            lo_ratio = pow(0.5 / E.lo_err, 1.0 / (double)ord);
            ratio = pow(0.5 / err, 1.0 / (double)(ord + 1));
            if (lo_ratio > ratio) {
                ord--;
                ratio = lo_ratio;
            }
        } else {
            ratio = pow(0.5 / err, 1.0 / (double)(ord + 1));
        }
        if (ratio >= 2.0) {
            ratio = 2.0;
            const_stp = 0;
        } else if (ratio <= 0.5) {
            ratio = 0.5;
            const_stp = 0;
        } else {
            ratio = 1.0;
            const_stp++;
        }
        double tmp = ratio * del_t;
        del_t = tmp > max_del_t ? max_del_t : (tmp < min_del_t ? min_del_t : tmp);
        istep++;
    }
    for (int i = 0; i < max_order; i++) free(y[i]);
    free(dy); free(a); free(y0); free(E.e_n);
    return rc ? rc : istep;
}

int orc_bdf_run_cells(orc_t** cells, int ncell, int nsteps, int max_order, double initial_del_t, int max_retry,
                      int fix, double* report, int* n_calc_g) {
    if (ncell < 1) return -3;
    bsys_t S;
    S.cell = cells;
    S.ncell = ncell;
    S.off = (long*)malloc((ncell + 1) * sizeof(long));
    S.off[0] = 0;
    for (int k = 0; k < ncell; k++) S.off[k + 1] = S.off[k] + (long)cells[k]->natms * cells[k]->nt;
    S.n = S.off[ncell];
    S.c = (double*)calloc(S.n, sizeof(double));
    S.c_d = (double*)calloc(S.n, sizeof(double));
    int rc = bdf_run_sys(&S, nsteps, max_order, initial_del_t, max_retry, fix, report, n_calc_g);
    free(S.off);
    free(S.c);
    free(S.c_d);
    return rc;
}

int orc_bdf_run(orc_t* o, int nsteps, int max_order, double initial_del_t, int max_retry, int fix,
                double* report, int* n_calc_g) {
    return orc_bdf_run_cells(&o, 1, nsteps, max_order, initial_del_t, max_retry, fix, report, n_calc_g);
}

int orc_natms(const orc_t* o) { return o->natms; }
int orc_nghost(const orc_t* o) { return o->nghost; }
long orc_npairs(const orc_t* o) { return o->no_pairs; }
long orc_npairs2(const orc_t* o) { return o->no_pairs2; }
const int* orc_list_offsets(const orc_t* o) { return o->nl_off; }
const int* orc_list(const orc_t* o) { return o->nl; }
const int* orc_list2_offsets(const orc_t* o) { return o->nl2_off; }
const int* orc_list2(const orc_t* o) { return o->nl2; }
const double* orc_x(const orc_t* o) { return o->x; }
const double* orc_c(const orc_t* o) { return o->c; }
const int* orc_ghost_src(const orc_t* o) { return o->ghost_src; }
int orc_nswap(const orc_t* o) { return o->nswap; }
const int* orc_swap_offsets(const orc_t* o) { return o->swap_off; }
const int* orc_crd(const orc_t* o) { return o->crd; }

double* orc_vec(orc_t* o, const char* name) {
    if (!strcmp(name, "f")) return o->f;
    if (!strcmp(name, "E")) return o->E;
    if (!strcmp(name, "dE")) return o->dE;
    if (!strcmp(name, "ddE")) return o->ddE;
    if (!strcmp(name, "n")) return o->n;
    if (!strcmp(name, "s")) return o->s;
    if (!strcmp(name, "t")) return o->t;
    if (!strcmp(name, "v")) return o->v;
    if (!strcmp(name, "c_d")) return o->c_d;
    if (!strcmp(name, "x")) return o->x;
    if (!strcmp(name, "c")) return o->c;
    return NULL;
}

void orc_scalars(const orc_t* o, double* out) {
    out[0] = o->rc; out[1] = o->cut_sq; out[2] = o->mod_rc; out[3] = o->cut_sq_mod;
    out[4] = o->cut_sk_sq; out[5] = o->kbT; out[6] = o->beta; out[7] = o->rho_max;
    out[8] = o->dr_inv; out[9] = o->drho_inv;
    for (int i = 0; i < o->nt; i++) out[10 + i] = o->c_0[i];
}

void orc_gh(const orc_t* o, double* xi, double* w0, double* w1, double* w2) {
    for (int i = 0; i < o->no_i; i++) {
        xi[i] = o->xi[i]; w0[i] = o->wi_0[i]; w1[i] = o->wi_1[i]; w2[i] = o->wi_2[i];
    }
}

const double* orc_table(const orc_t* o, int kind, int slot, int* nrows, int* ncoef) {
    if (kind == 0) { *nrows = o->nrho + 1; *ncoef = o->ncF; return o->F_arr + (long)slot * (o->nrho + 1) * o->ncF; }
    *nrows = o->nr; *ncoef = 7;
    if (kind == 1) return o->rho_arr + (long)slot * o->nr * 7;
    return o->phi_arr + (long)slot * o->nr * 7;
}

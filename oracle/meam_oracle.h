/*
 * TEST INFRASTRUCTURE ONLY -- the oracle's restatement of ForceField_meam
 * (src/ff_meam.cpp), the modified embedded-atom method with angular partial
 * densities and many-body screening, in the reference's operation order.
 * Only tests load it; the product never links or calls it.
 */
#ifndef MAPP_MEAM_ORACLE_H
#define MAPP_MEAM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#include "dmd_oracle.h"

#define ORC_MEAM_MAXT 4

/* reference lattices (src/ff_meam.cpp:13) */
enum { ORC_FCC = 0, ORC_BCC, ORC_HCP, ORC_DIM, ORC_DIA, ORC_B1, ORC_C11, ORC_L12, ORC_B2 };

/* The members ForceField_meam holds after read_global and read_local
   (src/ff_meam.cpp:2660-3191), before coef applies augt1 and the alloy rules.
   Per-element values are in type order; an off-diagonal lattice of -1 is
   unset.  delr_inv stays 10 whatever delr the local file gives (the reader
   sets delr only, :3090). */
typedef struct {
    int nt;
    double Z[ORC_MEAM_MAXT], A[ORC_MEAM_MAXT], rho0[ORC_MEAM_MAXT];
    double beta0[ORC_MEAM_MAXT], beta1[ORC_MEAM_MAXT], beta2[ORC_MEAM_MAXT], beta3[ORC_MEAM_MAXT];
    double t0[ORC_MEAM_MAXT], t1[ORC_MEAM_MAXT], t2[ORC_MEAM_MAXT], t3[ORC_MEAM_MAXT];
    int ielt[ORC_MEAM_MAXT], ibar[ORC_MEAM_MAXT];
    int lattice[ORC_MEAM_MAXT][ORC_MEAM_MAXT], nn2[ORC_MEAM_MAXT][ORC_MEAM_MAXT], zbl[ORC_MEAM_MAXT][ORC_MEAM_MAXT];
    double Ec[ORC_MEAM_MAXT][ORC_MEAM_MAXT], alpha[ORC_MEAM_MAXT][ORC_MEAM_MAXT], re[ORC_MEAM_MAXT][ORC_MEAM_MAXT];
    double delta[ORC_MEAM_MAXT][ORC_MEAM_MAXT], attrac[ORC_MEAM_MAXT][ORC_MEAM_MAXT];
    double repuls[ORC_MEAM_MAXT][ORC_MEAM_MAXT], ebound[ORC_MEAM_MAXT][ORC_MEAM_MAXT];
    double cmin[ORC_MEAM_MAXT][ORC_MEAM_MAXT][ORC_MEAM_MAXT], cmax[ORC_MEAM_MAXT][ORC_MEAM_MAXT][ORC_MEAM_MAXT];
    double rc, delr, delr_inv, gsmooth;
    int augt1, ialloy, mix_ref_t, erose_form, emb_lin_neg, bkgd_dyn;
} orc_meam_params;

/* ForceField_meam::coef after the two readers (src/ff_meam.cpp:2593-2627):
   augt1, alloy_params, compute_reference_density, the pair table.  Returns 0,
   or -1 for an unset or unknown lattice, an unknown ibar, mix_ref_t = 1 (the
   reference's pair potential then reads an uninitialised background density,
   :3747-3748) or more than ORC_MEAM_MAXT types. */
int orc_meam_coef(orc_t* o, const orc_meam_params* p);
/* 1: the corrected force loop -- the rho2 angular argument carries its
   direction factor (the reference adds rho2 components without it,
   :1692-1693), the screening derivative keeps the cutoff taper when the
   screening product is 1 (:1262-1263 drop it), the ialloy = 1 screening
   derivatives of the t averages take the atomic density (:1896-1906 take its
   r-derivative) and square t2, t3 (:1910-1913 use t1), and the two
   three-body derivatives of the screening go to their own atoms (the
   reference's dCfunc2 call swaps them, :2052); with it f = -grad U
   wherever no screening atom is a ghost of the pair's owner (the reference
   drops those three-body forces, :2076).  0 (default): the reference. */
void orc_meam_set_fix(orc_t* o, int fix);
/* the pair table phirar[COMP][row][0..6] (nr = 1000 rows of 7) and its step */
const double* orc_meam_table(orc_t* o, int* nr, double* dr);
/* per-site records (natms + ghosts) x 38 after force_calc or energy_calc, and
   the reference densities rho_ref[nt] */
const double* orc_meam_records(orc_t* o);
const double* orc_meam_rho_ref(orc_t* o);
/* the pair potential of the reference structure at r (phi_meam, :3629-3842) */
double orc_meam_phi(orc_t* o, double r, int a, int b);

void meam_force_calc(orc_t* o, int st_clc, double* en_st);
double meam_energy_calc(orc_t* o);

#ifdef __cplusplus
}
#endif
#endif

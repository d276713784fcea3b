/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * FP64 CPU restatement of MAPP's diffusive-molecular-dynamics force field
 * (ForceField_DMD, ForceField_DMD_I, ForceField_DMD_II), its plain EAM loop
 * (ForceField_eam), the single-rank periodic ghost construction and the cell
 * list / half neighbor list, in the reference's operation order.  It is
 * compiled with -O2 -ffp-contract=off so every product and sum is a separately
 * rounded IEEE binary64 operation, as in the reference build.
 *
 * Only tests, the smoke check and the CPU-baseline leg of bench.py may load it;
 * the product never links or calls it.
 *
 * Function-by-function citations are in dmd_oracle.c.
 */
#ifndef MAPP_DMD_ORACLE_H
#define MAPP_DMD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* variants of the DMD force field (they differ only in c_d_calc / calc_g) */
enum { ORC_DMD = 0, ORC_DMD_I = 1, ORC_DMD_II = 2, ORC_EAM = 3, ORC_LJ = 4, ORC_FS = 5, ORC_FSM = 6, ORC_MEAM = 7 };

/* MD-mode potentials (x rows [3], c holds the type): Lennard-Jones
   (ForceField_lj, src/ff_lj.cpp) and Finnis-Sinclair (ForceField_fs,
   src/ff_fs.cpp).  Their neighbour lists use a cutoff per type pair,
   cut_sk_sq[COMP(type_i, type_j)] (src/neighbor.cpp:205-208). */
/* ForceField_lj::coef "itype jtype epsilon sigma cutoff" (src/ff_lj.cpp:72-125) */
int orc_lj_coef(orc_t* o, int itype, int jtype, double epsilon, double sigma, double cut);
/* ForceField_fs::coef after reading its file (src/ff_fs.cpp:113-335): A[nt],
   t1[nt*nt], t2[nt*nt] as mat_t_1[i][j], and per type pair in COMP order
   cut_rho, cut_phi, k1, k2, k3 (nt(nt+1)/2 each) */
int orc_fs_coef(orc_t* o, const double* A, const double* t1, const double* t2, const double* cut_rho,
                const double* cut_phi, const double* k1, const double* k2, const double* k3);

orc_t* orc_create(int nt, int variant);
void orc_destroy(orc_t* o);

/* Finnis-Sinclair tables in file order: F[a][nrho], rho[a][b][nr], rphi[i][j<=i][nr] */
int orc_set_fs(orc_t* o, int nrho, double drho, int nr, double dr,
               const double* F, const double* rho, const double* rphi);
/* ff_coef r_crd no_i alpha_min alpha_max T; masses in amu */
int orc_set_params(orc_t* o, double r_crd, int no_i, double alpha_min,
                   double alpha_max, double T, const double* mass);
/* H lower triangular (row-major 3x3); x: natms*(3+nt) with reduced coordinates
   in the first three slots and alpha after them; c: natms*nt.  For ORC_EAM:
   x is natms*3 reduced coordinates and c holds the type of each atom. */
int orc_set_system(orc_t* o, const double* H, int natms, const double* x,
                   const double* c);
/* force-field init + ghosts + cell list + half list (Min_LBFGS::init order) */
int orc_init(orc_t* o, double skin);
/* replace owned x (cartesian + alpha), then Atoms::update_0 (rebuild or ghost update) */
int orc_set_x_update(orc_t* o, const double* x_owned, int* rebuilt);
/* replace owned c and forward it to the ghosts */
void orc_set_c(orc_t* o, const double* c_owned);
/* 1: pair caches are zeroed before every force_calc ("fresh" semantics);
   0: reference semantics (zeroed only when the caches grow) */
void orc_set_fresh_caches(orc_t* o, int fresh);
/* 1: index dE by [atom*nt+type] in c_d_calc / calc_g instead of the reference's
   dE[jatm+jtype] (src/ff_dmd_I.cpp:2336,2339) -- a switch for tests only */
void orc_set_fix_dE_index(orc_t* o, int fix);
/* 1: force_calc / energy_calc of every potential (DMD, EAM, LJ, FS, FSM,
   MEAM) sum energy and virial with Neumaier's compensated summation (same
   terms, same order); 0 (default): the reference's plain running sum */
void orc_set_compensated(orc_t* o, int on);
int orc_compensated(const orc_t* o);
/* 1: consistent energy -- energy_calc sums every ordered type pair and both
   energies weight cross-type pair terms as the force loop does, so for
   symmetric density tables and one alpha per site f = -grad U;
   0 (default): the reference (jtype >= itype loop in energy_calc,
   src/ff_dmd_I.cpp:508; cross-type pair energy 0.5 (c c + c c) phi, :244-255) */
void orc_set_energy_full(orc_t* o, int full);

void orc_force_calc(orc_t* o, int st_clc, double* en_st);
/* for the MEAM restatement (meam_oracle.c): its state hangs on the oracle;
   orc_update_vec is Atoms::update of a per-atom vector of dimension d */
void* orc_ext(orc_t* o);
void orc_set_ext(orc_t* o, void* ext, void (*free_ext)(void*));
void orc_update_vec(orc_t* o, double* v, int d);
void orc_mark_params(orc_t* o, double rc);
double orc_energy_calc(orc_t* o);
void orc_create_2nd_neigh_lst(orc_t* o);
void orc_c_d_calc(orc_t* o);
double orc_calc_g(orc_t* o, int chk, double alpha, const double* a, double* g);

/* Min_LBFGS::run with LineSearch_BackTrack on a fixed box; energies has
   max_iter + 2 slots; returns the exit code (1 downhill, 2 grad0, 3 alpha_min,
   4 energy tolerance, 5 max iterations) */
int orc_min_lbfgs(orc_t* o, int max_iter, int m_it, double e_tol, double* energies, int* nsteps, int* n_force,
                  int* n_energy);

/* Clock_BDF::init + run: variable step and order (up to max_order), retries
   while the error estimate or cost ratio is >= 1; initial_del_t <= 0 takes the
   reference's a_tol / rms(c_d) rule; fix as in orc_bdf_coef.  report[6*nsteps]: per step order, del_t,
   cost ratio, error, CG iterations, solve calls.  Returns the steps taken,
   -1 if c left [0,1], -2 after more than max_retry retries of one step. */
int orc_bdf_run(orc_t* o, int nsteps, int max_order, double initial_del_t, int max_retry, int fix,
                double* report, int* n_calc_g);
/* the same run over several independent cells integrated as one system (the
   batched run): dof concatenated in cell order, calc_g's cost summed over the
   cells, one step size and order for all of them */
int orc_bdf_run_cells(orc_t** cells, int ncell, int nsteps, int max_order, double initial_del_t, int max_retry,
                      int fix, double* report, int* n_calc_g);
/* coefficients of Clock_BDF::interpolate: out[3 + 2q] = beta, alpha_dy_0,
   dalpha_dy_0, alpha_y[q], dalpha_y[q]; fix = 1 corrects the sign of the
   c1 c2 term of dalpha_y[0] (src/clock_bdf.cpp:327) */
void orc_bdf_coef(int q, double del_t, const double* t, int fix, double* out);

/* introspection */
int orc_natms(const orc_t* o);
int orc_nghost(const orc_t* o);
long orc_npairs(const orc_t* o);
long orc_npairs2(const orc_t* o);
const int* orc_list_offsets(const orc_t* o);   /* natms+1 */
const int* orc_list(const orc_t* o);           /* npairs */
const int* orc_list2_offsets(const orc_t* o);
const int* orc_list2(const orc_t* o);
const double* orc_x(const orc_t* o);           /* (natms+nghost)*xdim */
const double* orc_c(const orc_t* o);           /* (natms+nghost)*nt */
const int* orc_ghost_src(const orc_t* o);      /* nghost: source index of each ghost */
int orc_nswap(const orc_t* o);
const int* orc_swap_offsets(const orc_t* o);   /* nswap+1, ghost-relative */
double* orc_vec(orc_t* o, const char* name);   /* f, E, dE, ddE, n, s, t, v, c_d */
const int* orc_crd(const orc_t* o);
void orc_scalars(const orc_t* o, double* out); /* rc, cut_sq, mod_rc, cut_sq_mod, cut_sk_sq, kbT, beta, rho_max, dr_inv, drho_inv, c_0[0..nt-1] */
void orc_gh(const orc_t* o, double* xi, double* w0, double* w1, double* w2);
/* spline rows: kind 0=F (9 per row), 1=rho (7), 2=r*phi (7) */
const double* orc_table(const orc_t* o, int kind, int slot, int* nrows, int* ncoef);
/* box helpers (restated M3INV, x2s, s2x) */
void orc_m3inv(const double* H, double* B);
void orc_x2s(const double* B, int n, int xdim, double* x);
void orc_s2x(const double* H, int n, int xdim, double* x);

/* Gaussian-averaged radial primitives for one pair (r, alpha) and table slots:
   out = {rho, drho_dr, drho_dalpha, phi, dphi_dr, dphi_dalpha}; returns 1 if computed */
int orc_pair_quadrature(const orc_t* o, double r, double alpha, int rho_slot,
                        int phi_slot, double* out);

#ifdef __cplusplus
}
#endif
#endif

/*
 * mapp_dmd.h -- C ABI of the MI355X-native DMD force field.
 *
 * One context holds one MPI rank's whole domain on one GPU: tables, atoms
 * (owned + periodic ghosts), the half neighbor list and every per-atom
 * vector stay resident in HBM between calls.  Every function returns an int
 * status (0 = success, negative = error, message from mapp_dmd_last_error)
 * where the reference calls error->abort (src/error.cpp:29-43).
 *
 * Each entry point names the reference interface it replaces.  The reference
 * force field is the C++ class ForceField (src/ff.h:12-29) with the DMD
 * implementations ForceField_DMD / ForceField_DMD_I / ForceField_DMD_II
 * (src/ff_dmd.cpp, src/ff_dmd_I.cpp, src/ff_dmd_II.cpp).
 *
 * Layouts follow the reference's per-atom vectors (src/atoms.h:21-131):
 *   x   [natms][3+nt]  position (Angstrom) then alpha_t (1/Angstrom^2)
 *   c   [natms][nt]    concentrations
 *   f   [natms][3+nt]  -dU/dx, -dU/dalpha
 *   E, dE, n, c_d ...  [natms][nt]
 */
#ifndef MAPP_DMD_H
#define MAPP_DMD_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mapp_dmd mapp_dmd_t;

enum {
    MAPP_DMD_OK = 0,
    MAPP_DMD_E_ARG = -1,   /* invalid argument */
    MAPP_DMD_E_STATE = -2, /* call out of order (e.g. force before init) */
    MAPP_DMD_E_HIP = -3,   /* HIP runtime error */
    MAPP_DMD_E_RANGE = -4, /* value outside what the tables / kernels cover */
    MAPP_DMD_E_IO = -5,    /* file could not be read or parsed */
    MAPP_DMD_E_INTERNAL = -6 /* unexpected host-side failure (message in mapp_dmd_last_error / err) */
};

/* variants differ in the mobility of c_d_calc / calc_g only */
enum {
    MAPP_DMD_VARIANT_DMD = 0, /* ForceField_DMD    (src/ff_dmd.cpp)    Mat = 5.2 beta, rate exp(n + Mat) */
    MAPP_DMD_VARIANT_I = 1,   /* ForceField_DMD_I  (src/ff_dmd_I.cpp)  Mat = 0.4 beta */
    MAPP_DMD_VARIANT_II = 2,  /* ForceField_DMD_II (src/ff_dmd_II.cpp) mat = max(f_i,f_j) + 0.5 */
    /* plain EAM, ForceField_eam (src/ff_eam.cpp): x rows are [3] (positions),
       c holds the type of each site, coef ignores its DMD arguments, and
       force_calc / energy_calc follow src/ff_eam.cpp:44-356 */
    MAPP_DMD_VARIANT_EAM = 3,
    /* MD-mode pair potentials with the EAM layout (x rows [3], c = type) and a
       list cutoff per type pair, cut_sk_sq[COMP(type_i, type_j)]
       (src/neighbor.cpp:205-208); up to 4 types.  Coefficients come from
       mapp_dmd_lj_coef / mapp_dmd_fs_coef / mapp_dmd_read_fs_params instead of
       set_fs + coef. */
    MAPP_DMD_VARIANT_LJ = 4, /* ForceField_lj (src/ff_lj.cpp), force_calc :160-264, energy_calc :270-341 */
    MAPP_DMD_VARIANT_FS = 5, /* ForceField_fs (src/ff_fs.cpp), force_calc :379-604, energy_calc :609-686 */
    /* ForceField_fsm (src/ff_fsm.cpp), force_calc :371-576, energy_calc
       :582-664: FS with the cutoffs kept squared (:281-282) and compared as
       rsq < cut^2; same coefficient entry points and file format as FS */
    MAPP_DMD_VARIANT_FSM = 6,
    /* ForceField_meam (src/ff_meam.cpp): modified EAM with angular partial
       densities and three-body screening over the FULL neighbour list
       (pair_wise = 0, src/ff_meam.cpp:675; src/neighbor.cpp:240-284), one
       cutoff rc for every type pair; force_calc :699-2142, energy_calc
       :2148-2589.  Coefficients from mapp_dmd_meam_coef /
       mapp_dmd_read_meam_params. */
    MAPP_DMD_VARIANT_MEAM = 7
};

/* arithmetic of the pair quadrature (per-site sums, embedding and the
   energy total stay FP64 in every mode) */
enum {
    MAPP_DMD_PREC_FP64 = 0,     /* quadrature in FP64 */
    MAPP_DMD_PREC_MIXED = 1,    /* quadrature sums in FP32 with FP32 accumulation */
    MAPP_DMD_PREC_SURROGATE = 2 /* BF16-MFMA MLP surrogate for the averaged primitives */
};

/* per-atom vectors for mapp_dmd_get_vec / mapp_dmd_set_vec / mapp_dmd_device_ptr */
enum {
    MAPP_DMD_VEC_X = 0,   /* [natms][3+nt] */
    MAPP_DMD_VEC_C = 1,   /* [natms][nt] */
    MAPP_DMD_VEC_F = 2,   /* [natms][3+nt] */
    MAPP_DMD_VEC_E = 3,   /* [natms][nt]  (after force_calc: F(rho_bar)) */
    MAPP_DMD_VEC_DE = 4,  /* [natms][nt]  F'(rho_bar) */
    MAPP_DMD_VEC_N = 5,   /* [natms][nt]  chemical potential of c_d_calc / calc_g */
    MAPP_DMD_VEC_CD = 6,  /* [natms][nt]  dc/dt */
    MAPP_DMD_VEC_DDE = 7, /* [natms][nt]  F''(rho_bar) (calc_g) */
    MAPP_DMD_VEC_S = 8,   /* [natms][nt]  calc_g residual s */
    MAPP_DMD_VEC_T = 9,   /* [natms][nt]  calc_g t */
    MAPP_DMD_VEC_V = 10,  /* [natms][nt]  calc_g v */
    MAPP_DMD_VEC_A = 11,  /* [natms][nt]  device buffer for calc_g's a (caller-filled) */
    MAPP_DMD_VEC_G = 12   /* [natms][nt]  device buffer for calc_g's g */
};

/* Constructor: ForceField_DMD_I::ForceField_DMD_I (src/ff_dmd_I.cpp:10-40).
   no_types = nt (1..4; plain EAM 1..2, MAPP_DMD_E_RANGE beyond); device = HIP
   device ordinal.  The surrogate path needs nt = 2. */
int mapp_dmd_create(mapp_dmd_t** ctx, int no_types, int variant, int device);
/* Destructor: ForceField_DMD_I::~ForceField_DMD_I (src/ff_dmd_I.cpp:44-85). */
int mapp_dmd_destroy(mapp_dmd_t* ctx);
const char* mapp_dmd_last_error(const mapp_dmd_t* ctx);

/* Finnis-Sinclair tables, element order = type order:
   ForceField_DMD_II::set_fs (src/ff_dmd_II.cpp:1253-1480) then set_arrays
   (src/ff_dmd_I.cpp:1729-1757).  F: [nt][nrho]; rho: [a][b][nr] (density at
   an a site from a b neighbour, the file's order); rphi: [i][j<=i][nr] holding
   r*phi.  nr must equal nrho (the reference clamps F rows with nr-2,
   src/ff_dmd_I.cpp:326). */
int mapp_dmd_set_fs(mapp_dmd_t* ctx, int nrho, double drho, int nr, double dr,
                    const double* F, const double* rho, const double* rphi);
/* FS file reader, same format as ForceField_DMD_II::set_fs
   (src/ff_dmd_II.cpp:1253-1480): elements[nt] names the type order. */
int mapp_dmd_read_fs(mapp_dmd_t* ctx, const char* path, const char* const* elements);
/* ff_coef r_crd no_i alpha_min alpha_max T: ForceField_DMD_I::coef
   (src/ff_dmd_I.cpp:735-796); mass[nt] in amu (AtomTypes::mass). */
int mapp_dmd_coef(mapp_dmd_t* ctx, double r_crd, int no_i, double alpha_min,
                  double alpha_max, double T, const double* mass);
int mapp_dmd_set_precision(mapp_dmd_t* ctx, int mode);

/* ff_coef "itype jtype epsilon sigma cutoff" of ForceField_lj::coef
   (src/ff_lj.cpp:72-125), one call per type pair.  Negative values are
   rejected; a zero epsilon, sigma or cutoff leaves the pair without forces but
   with the list cutoff (cutoff + skin)^2; a type pair never given a call lists
   nothing.  The potential is not shifted (the constructor's shift, :35). */
int mapp_dmd_lj_coef(mapp_dmd_t* ctx, int itype, int jtype, double epsilon, double sigma, double cut);
/* Finnis-Sinclair coefficients of ForceField_fs::coef (src/ff_fs.cpp:113-335)
   in type order: A[nt]; t1[nt*nt], t2[nt*nt] as mat_t_1[i][j], mat_t_2[i][j];
   per type pair in COMP order (i <= j: j(j+1)/2 + i) cut_rho, cut_phi, k1, k2,
   k3 [nt(nt+1)/2].  Also for the FSM variant (ForceField_fsm::coef,
   src/ff_fsm.cpp:111-333), which keeps cut_rho^2 and cut_phi^2. */
int mapp_dmd_fs_coef(mapp_dmd_t* ctx, const double* A, const double* t1, const double* t2, const double* cut_rho,
                     const double* cut_phi, const double* k1, const double* k2, const double* k3);
/* ff_coef FS <file>: the parameter file of ForceField_fs::coef (the element
   line, nt A lines, nt*nt "t1 t2" lines, nt(nt+1)/2 "cut_rho cut_phi k1 k2 k3"
   lines, in the file's element order; '#' starts a comment); elements[nt]
   names the type order.  mapp_dmd_fs_params_read parses it without a context
   (arrays as for mapp_dmd_fs_coef, in elements' order). */
int mapp_dmd_read_fs_params(mapp_dmd_t* ctx, const char* path, const char* const* elements);
int mapp_dmd_fs_params_read(const char* path, int nt, const char* const* elements, double* A, double* t1,
                            double* t2, double* cut_rho, double* cut_phi, double* k1, double* k2, double* k3,
                            char* err, int err_cap);

/* MEAM parameters: the members ForceField_meam holds after read_global and
   read_local (src/ff_meam.cpp:2660-3191), before coef applies augt1, the
   alloy rules, the reference densities and the pair table (:2593-2627).
   Per-element values in type order; lattice codes 0 fcc, 1 bcc, 2 hcp,
   3 dim, 4 dia, 5 b1, 6 c11, 7 l12, 8 b2 (src/ff_meam.cpp:13), -1 unset.
   delr_inv stays 10 whatever delr the local file gives (the reader sets delr
   only, :3090). */
#define MAPP_MEAM_MAXT 4
typedef struct {
    int nt;
    double Z[MAPP_MEAM_MAXT], A[MAPP_MEAM_MAXT], rho0[MAPP_MEAM_MAXT];
    double beta0[MAPP_MEAM_MAXT], beta1[MAPP_MEAM_MAXT], beta2[MAPP_MEAM_MAXT], beta3[MAPP_MEAM_MAXT];
    double t0[MAPP_MEAM_MAXT], t1[MAPP_MEAM_MAXT], t2[MAPP_MEAM_MAXT], t3[MAPP_MEAM_MAXT];
    int ielt[MAPP_MEAM_MAXT], ibar[MAPP_MEAM_MAXT];
    int lattice[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT], nn2[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    int zbl[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double Ec[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT], alpha[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double re[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT], delta[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double attrac[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT], repuls[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double ebound[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double cmin[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double cmax[MAPP_MEAM_MAXT][MAPP_MEAM_MAXT][MAPP_MEAM_MAXT];
    double rc, delr, delr_inv, gsmooth;
    int augt1, ialloy, mix_ref_t, erose_form, emb_lin_neg, bkgd_dyn;
} mapp_meam_params;

/* ff_coef MEAM <library> <local>: ForceField_meam::coef (src/ff_meam.cpp:
   2593-2627).  mapp_dmd_meam_params_read parses the library file (per element
   the quoted name and lattice, z, ielement, atwt, alpha, b0..b3, alat, esub,
   asub, t0..t3, rozero, ibar; quotes read as blanks; elements not among the
   types are skipped, :2660-2857) and the local file ("keyword(i,j,k) = value"
   with quotes, commas, parentheses and '=' read as blanks; indices 1-based in
   library order; "NULL" skips it, :2858-3191) into p, for the types named by
   elements[nt].  The reader's own behaviour is kept: "delr" leaves delr_inv at
   10, "bkgd_dyn" sets emb_lin_neg, "zbl" is an error.  mapp_dmd_meam_coef
   derives the tables; it rejects mix_ref_t = 1 (the reference's pair potential
   then reads an uninitialised background density, :3747-3748), an unset pair
   lattice and an unknown ibar. */
int mapp_dmd_meam_params_read(const char* library, const char* local, int nt, const char* const* elements,
                              mapp_meam_params* p, char* err, int err_cap);
int mapp_dmd_meam_coef(mapp_dmd_t* ctx, const mapp_meam_params* p);
int mapp_dmd_read_meam_params(mapp_dmd_t* ctx, const char* library, const char* local,
                              const char* const* elements);
/* 1: a corrected MEAM force loop -- the rho2 angular argument carries its
   direction factor (the reference drops it, src/ff_meam.cpp:1692-1693), the
   screening derivative keeps the cutoff taper when the screening product is 1
   (:1262-1263 drop it), the ialloy = 1 screening derivatives of the t
   averages take the atomic density (:1896-1906 take its r-derivative) and
   square t2, t3 (:1910-1913 use t1), and the two three-body derivatives of
   the screening go to their own atoms (the dCfunc2 call swaps them, :2052).
   0 (default): the reference.  Forces on a screening atom that is a ghost in
   the pair owner's list are dropped in both (:2076). */
int mapp_dmd_meam_set_fix(mapp_dmd_t* ctx, int fix);

/* System: lower-triangular H (row-major 3x3, rows are the box vectors,
   src/read_cfg.cpp:425-453), natms sites whose x rows hold the reduced
   coordinates s in [0,1) followed by alpha (what read_cfg leaves in x,
   src/read_cfg.cpp:562-625), and c.  Host pointers. */
int mapp_dmd_set_system(mapp_dmd_t* ctx, const double* H, int natms,
                        const double* x_reduced, const double* c);
/* ForceField_DMD_I::init (src/ff_dmd_I.cpp:672-699) + Atoms::ph_setup
   (src/atoms.cpp:317-372) + Neighbor::create_list(0,1)
   (src/neighbor.cpp:137-292) + Atoms::store_0, on the GPU. */
int mapp_dmd_init(mapp_dmd_t* ctx, double skin);
/* ForceField_DMD_I::fin (src/ff_dmd_I.cpp:703-725): drops lists and caches. */
int mapp_dmd_fin(mapp_dmd_t* ctx);

/* New owned x (cartesian + alpha; host pointer, or NULL when the device
   vector was written in place), then Atoms::update_0 (src/atoms.cpp:1780-1865):
   rebuild ghosts and lists if a site moved more than skin, else forward-update
   ghost x.  *rebuilt (may be NULL) reports which. */
int mapp_dmd_set_x(mapp_dmd_t* ctx, const double* x_host, int* rebuilt);
/* New owned c (host pointer or NULL for in-place device writes), forwarded to
   ghosts (Atoms::update of c). */
int mapp_dmd_set_c(mapp_dmd_t* ctx, const double* c_host);

/* virtual void force_calc(int st_clc, TYPE0* en_st)  (src/ff.h:20,
   src/ff_dmd_I.cpp:88-450).  en_st (host, >= 7 doubles): [0] = energy;
   when st_clc, [1..6] = virial xx yy zz yz zx xy. Writes f, E, dE. */
int mapp_dmd_force_calc(mapp_dmd_t* ctx, int st_clc, double* en_st);
/* virtual TYPE0 energy_calc()  (src/ff.h:21, src/ff_dmd_I.cpp:454-668). */
int mapp_dmd_energy_calc(mapp_dmd_t* ctx, double* energy);
/* virtual void create_2nd_neigh_lst()  (src/ff.h:27, src/ff_dmd_I.cpp:2138-2217). */
int mapp_dmd_create_2nd_neigh_lst(mapp_dmd_t* ctx);
/* virtual void c_d_calc()  (src/ff.h:29; src/ff_dmd_I.cpp:2241-2398,
   src/ff_dmd_II.cpp:2186-2337, src/ff_dmd.cpp:2241-2398).  Writes c_d, E, dE, n. */
int mapp_dmd_c_d_calc(mapp_dmd_t* ctx);
/* virtual TYPE0 calc_g(int chk, TYPE0 alpha, TYPE0* a, TYPE0* g)  (src/ff.h:28;
   src/ff_dmd_I.cpp:2402-2674, src/ff_dmd_II.cpp:2341-2630).  a and g are
   DEVICE pointers of natms*nt doubles, or NULL for the library's own buffers
   (MAPP_DMD_VEC_A / MAPP_DMD_VEC_G); g is written only when chk == 0. */
int mapp_dmd_calc_g(mapp_dmd_t* ctx, int chk, double alpha, const double* a_dev,
                    double* g_dev, double* cost);
/* energy mode (no counterpart in the reference; a switch):
   0 (default) the reference: energy_calc loops over type pairs jtype >= itype
   (src/ff_dmd_I.cpp:508), so its energy differs from force_calc's for two or
   more types, and both weight cross-type pair terms by 0.5 (c c + c c)
   (:244-255) while the force loop weights each ordered pair by c_i c_j
   (:375-380);
   1 consistent energy: energy_calc over every ordered type pair and cross-type
   pair terms weighted as in the forces, in both energies; then, for symmetric
   density tables, f = -grad U in x and in a per-site alpha (sum of the types'
   alpha forces). */
int mapp_dmd_set_energy_mode(mapp_dmd_t* ctx, int full);

/* Min_LBFGS::run on a fixed box with LineSearch_BackTrack (replaces the
   minimiser's host loops, src/min_l-bfgs.cpp:302-787, src/line_search.cpp:
   32-341): x / f / h and the m_it (s, y) pairs stay on the device, dot
   products are FP64 deterministic device reductions; energy_calc, force_calc
   and Atoms::update_0 (skin test, rebuild or ghost update) as in the reference.
   energies[max_iter + 2]: force_calc energy after each step (energies[0]
   initial); counts[4]: steps, force_calc calls, energy_calc calls, exit code
   (1 not downhill, 2 zero gradient, 3 step below alpha_min, 4 energy change
   below e_tol, 5 max_iter reached). */
int mapp_dmd_min_lbfgs(mapp_dmd_t* ctx, int max_iter, int m_it, double e_tol, double* energies, long* counts);

/* Clock_BDF::init + Clock_BDF::run (src/clock_bdf.cpp:201-275, :479-672)
   on the device: force_calc, create_2nd_neigh_lst and c_d_calc, then nsteps
   steps of the variable-step, variable-order (up to max_order) BDF method on
   c with the nonlinear conjugate-gradient solve of calc_g's cost
   (Clock_BDF::solve, :677-883) and the order estimates of hi_lo_err_est
   (:372-478).  History, residual and CG vectors stay in HBM; dot products
   are FP64 deterministic device reductions.  initial_del_t <= 0 takes the
   reference's a_tol / rms(c_d) rule.  Deviations: solve's uninitialised
   curr_t is t[0] + del_t; a step retried more than max_retry times ends the
   run (the reference loops).  flags bit 0 corrects the sign of the
   (t_new - t[0]) c1 c2 term of dalpha_y[0] (:327; the reference's corrector
   then drops y_n for orders >= 2); default off.  report[6*nsteps]: per step order, del_t, cost
   ratio, error estimate, CG iterations, solve calls; counts[2]: steps taken
   (or -1 when c left [0,1], -2 after too many retries), calc_g calls. */
int mapp_dmd_bdf_run(mapp_dmd_t* ctx, int nsteps, int max_order, double initial_del_t, int max_retry, int flags,
                     double* report, long* counts);

/* A batch of independent periodic cells in one context (no counterpart in
   the reference, which runs one domain per process): cell k has the lower-
   triangular box H[9k..9k+8] and natms[k] sites; x (reduced positions +
   alpha) and c are the cells' rows concatenated in cell order.  init builds
   each cell's ghosts and lists as set_system + init would, then concatenates
   them (owned sites of all cells, then all ghosts), so force_calc,
   energy_calc, c_d_calc, calc_g, create_2nd_neigh_lst and bdf_run cover every
   cell in one launch per kernel; per-site outputs come back in cell order and
   energies / virials are batch totals.  Positions are fixed (set_x and
   min_lbfgs are refused). */
int mapp_dmd_set_cells(mapp_dmd_t* ctx, int ncell, const double* H, const int* natms, const double* x,
                       const double* c);
/* number of cells, and per cell owned sites and ghosts (up to cap) */
int mapp_dmd_cells(mapp_dmd_t* ctx, int* ncell, int* natms, int* nghost, int cap);

/* Launch schedules of the two force passes (no counterpart in the reference):
   {pass-1 lanes per (pair, type pair) (1, 2 or 4), pass-1 waves per block
   (1, 2, 4), pass-1 total waves (0 = one wavefront per site, otherwise a
   persistent grid whose waves stride through XCD-contiguous site ranges),
   pass-2 waves per block, pass-2 total waves}.  mapp_dmd_autotune times every
   candidate on the context's own system with HIP events (reps runs each
   after one warm-up force_calc) and keeps the fastest per pass (until then:
   {2, 4, 0, 2, 0}); report
   [cap][5] receives {lanes per pair, waves per block, total waves, pass-1 ms,
   pass-2 ms}. */
int mapp_dmd_autotune(mapp_dmd_t* ctx, int reps, double* report, int cap, int* ncand);
int mapp_dmd_get_schedule(mapp_dmd_t* ctx, int* out);
int mapp_dmd_set_schedule(mapp_dmd_t* ctx, const int* in);

/* Extended CFG files of DMD runs (host only, no context, no GPU).
   mapp_dmd_cfg_read follows ReadCFG (src/read_cfg.cpp:15-178 header and atom
   loop, :228-350 read_header, :356-418 set_box without eta, :562-657
   read_atom_dmd/add_atom_read_x): H comes back lower triangular (row-major),
   x as [natms][3+nt] reduced position wrapped into [0,1) and alpha, c as
   [natms][nt], mass[nt], names '\n'-separated.  With x == NULL only natms and
   nt are filled.  mapp_dmd_cfg_write writes the layout of
   WriteCFG::write_file_dmd (src/write_cfg.cpp:262-433) with %.17g, so a round
   trip is exact.  err receives the message of a failed call. */
int mapp_dmd_cfg_read(const char* path, int cap_atoms, int* natms, int* nt, double* H, double* x, double* c,
                      double* mass, char* names, int names_cap, char* err, int err_cap);
int mapp_dmd_cfg_write(const char* path, const double* H, int natms, int nt, const double* x, const double* c,
                       int ntypes, const double* mass, const char* const* names, char* err, int err_cap);

/* Learned surrogate (no counterpart in the reference):
   one MLP 3->64->64->64->6 (SiLU) per ordered type pair maps (r, alpha_i,
   alpha_j) to {rho_bar, drho_bar/dr, drho_bar/dalpha, phi_bar, dphi_bar/dr,
   dphi_bar/dalpha} as baseline + s(r) * net: the baseline is the second-order
   Gaussian smoothing of the tabulated functions at t = r (one table row), s(r)
   an envelope of the residual.  Trains on the GPU on n_per_tp samples per type
   pair labelled by the FP32 quadrature (current rule and tables), Adam with a
   cosine schedule.  The quadrature then judges the inference path (BF16 MFMA
   networks) on validation samples, and a cell of (r, alpha_i, alpha_j) is
   served by the networks only where every sample's error is within half of the
   per-pair tolerances that a relative force bound tol (max |f_net - f| /
   max |f| over the current system) implies (see mapp_dmd_surrogate_set_bound);
   every other pair takes the FP32 quadrature.  tol < 0 serves no cell,
   tol >= 1e30 every cell.  Needs an initialised two-type system (the bound's
   scales are taken from it).
   box = {r_lo, r_hi, alpha_lo, alpha_hi} or NULL for {1.8, quadrature cut, 60, 120}.
   report (8 doubles): final train loss, validation mean and max error of the
   networks in their standardised units, fraction of cells falling back, steps,
   seconds, training samples, validation samples.
   Afterwards MAPP_DMD_PREC_SURROGATE runs pass 1 of force_calc on BF16 MFMA. */
int mapp_dmd_surrogate_train(mapp_dmd_t* ctx, long n_per_tp, int epochs, int batch_per_tp, double lr,
                             unsigned long long seed, double tol, const double* box, double* report);
/* gradient of each network's mean squared (standardised) error over n_per_tp
   samples (a multiple of 64) by the training kernel: params [4][8966], norm
   [56] (box, input scaling, output means, output stds), X [4*n][3], T [4*n][6]
   -> grad [4][8966], loss [4] */
int mapp_dmd_surrogate_gradient(mapp_dmd_t* ctx, const float* params, const float* norm, long n_per_tp,
                                const float* X, const float* T, float* grad, float* loss);
/* trained parameters [4][8966] and normalisation [56] */
int mapp_dmd_surrogate_params(mapp_dmd_t* ctx, float* params, float* norm);
/* number of (pair, type pair) samples of the last surrogate force_calc that fell back to the quadrature
   (when the mask serves no cell, force_calc runs the mixed path's pass 1 and every in-cut evaluation
   counts) */
int mapp_dmd_surrogate_fallbacks(mapp_dmd_t* ctx, long* count);
/* calc_g without any allocation: no counterpart in the reference.  For callers
   that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc): the same computation and
   result as mapp_dmd_calc_g, with the cost reduction's scratch in a
   file-scope device array of the library (about 2.1 MB per GPU, for systems
   of up to 16,777,216 owned sites; larger ones get MAPP_DMD_E_RANGE).  a_dev,
   g_dev: device vectors [natms][nt] (both required).  The pair values must be
   current: a mapp_dmd_c_d_calc or mapp_dmd_calc_g since the last change of
   positions or widths (else MAPP_DMD_E_STATE); concentrations may change
   (mapp_dmd_set_c).  Calls on one GPU share the array, so calls that overlap
   on one GPU race. */
int mapp_dmd_calc_g_noalloc(mapp_dmd_t* ctx, int chk, double alpha, const double* a_dev, double* g_dev,
                            double* cost);

/* re-derives the fallback mask of the trained networks for another relative
   force bound without retraining.  Per-pair tolerances, from first-order error
   propagation with the triangle inequality: with N the largest number of
   neighbours of a site inside the quadrature cut, P_r, P_a the largest
   per-site sums of |drho_bar/dr|, |drho_bar/dalpha| over them, F = max |f|,
   F' = max |dE|, F'' = max |F''| over the densities present, U the energy:
     phi_r, phi_alpha: bound F / (3N);  rho_r, rho_alpha: bound F / (6 N F');
     rho: bound F / (6 N F'' max(P_r, P_a)) (and the energy's share);
     phi: bound |U| / (natms N).
   report (16 doubles): fraction of cells served, the six tolerances in output
   order, the bound, then the scales F, F', F'', N, P_r, P_a, |U|, natms. */
int mapp_dmd_surrogate_set_bound(mapp_dmd_t* ctx, double bound, double* report);
/* the smallest relative force bound at which each cell of the trained networks
   is served (every validation sample within half of each tolerance; infinity
   for a cell without samples): out[tp][64 r cells][4 alpha_i cells][4 alpha_j
   cells], cap >= 4096 doubles */
int mapp_dmd_surrogate_cell_bounds(mapp_dmd_t* ctx, double* out, long cap);

/* host <-> device copies of per-atom vectors (owned rows only) */
int mapp_dmd_get_vec(mapp_dmd_t* ctx, int which, double* host, long count);
int mapp_dmd_device_ptr(mapp_dmd_t* ctx, int which, void** dev_ptr);
/* host -> device copy into the owned rows of X (staging for set_x(NULL)),
   A or G; count = number of doubles supplied */
int mapp_dmd_set_vec(mapp_dmd_t* ctx, int which, const double* host, long count);
int mapp_dmd_sync(mapp_dmd_t* ctx);

/* sizes: out[0] natms, out[1] nghost, out[2] half-list pairs, out[3] 2nd-list
   pairs, out[4] number of ghost swaps, out[5] pairs inside the quadrature cut */
int mapp_dmd_sizes(mapp_dmd_t* ctx, long* out);
/* cutoffs after coef (ForceField_DMD_I::coef, src/ff_dmd_I.cpp:786-792):
   out[0] rc = (nr-1) dr, out[1] mod_rc = rc + xi[no_i-1]/sqrt(alpha_min) (the
   quadrature cut; init's list cut is mod_rc + skin, :676-678), out[2] r_crd */
int mapp_dmd_get_cutoffs(mapp_dmd_t* ctx, double* out);
/* Neighbor::neighbor_list as CSR (offsets natms+1, entries npairs) */
int mapp_dmd_get_list(mapp_dmd_t* ctx, int* offsets, int* entries);
int mapp_dmd_get_list2(mapp_dmd_t* ctx, int* offsets, int* entries);
/* ghost rows: x (cartesian + alpha) [nghost][3+nt] and source index [nghost] */
int mapp_dmd_get_ghosts(mapp_dmd_t* ctx, double* x, int* src);
/* Gauss-Hermite rule in use: xi, wi_0, wi_1, wi_2 (no_i each) */
int mapp_dmd_get_gh(mapp_dmd_t* ctx, double* xi, double* w0, double* w1, double* w2);
/* device time in ms of the kernels of the last force_calc:
   [0] pass 1, [1] embedding, [2] pass 2, [3] gathers + reduction, [4] total */
int mapp_dmd_last_timing(mapp_dmd_t* ctx, double* ms);
/* enable (1) / disable (0) per-kernel event timing of force_calc */
int mapp_dmd_enable_timing(mapp_dmd_t* ctx, int on);

#ifdef __cplusplus
}
#endif
#endif

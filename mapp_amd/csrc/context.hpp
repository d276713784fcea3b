// The context behind the C ABI: everything one MPI rank's domain needs, resident on one GPU.
#pragma once
#include <string>
#include <vector>

#include "dmd_common.hpp"
#include "meam_setup.hpp"
#include "surrogate_types.hpp"

struct mapp_dmd {
    int nt = 0, variant = 1, device = 0, prec = MAPP_DMD_PREC_FP64;
    std::string err;
    hipStream_t stream = nullptr;

    // tables (host copies, reference layouts)
    bool have_tables = false, have_params = false, have_system = false, inited = false;
    int nr = 0, nrho = 0;
    double dr = 0, drho = 0;
    std::vector<double> F9, rho7, phi7, F7;  // [nt][nrho+1][9], [nt*nt][nr][7], [nphi][nr][7], EAM F [nt][nrho+1][7]

    // parameters
    mappdmd::FFParams<double> P{};
    mappdmd::MdParams md{};  // LJ / FS coefficients per type pair
    mappdmd::meam_host::Setup meam;  // MEAM: parameters after coef, pair table (host)
    int meam_fix = 0;
    mappdmd::DevBuf<double> d_mtab, d_mrec;  // MEAM pair table, per-site records [ntot][38]
    mappdmd::DevBuf<double4> d_mpairs;       // MEAM per list entry {fc, S, dS, dU/dS}
    mappdmd::DevBuf<int> d_mactive, d_mact;  // MEAM: per site the count and list positions of its own
                                             // entries (j > i) with dU/dS != 0 and 0 < S fc < 1
    mappdmd::FFParams<float> Pf{};
    double rsq_crd = 0, mod_rc = 0;
    double wi1_ref[mappdmd::kMaxPoints]{}, wi2_ref[mappdmd::kMaxPoints]{};

    // box and atoms
    mappdmd::BoxParams bp{};
    double skin = 0.5;
    int natms = 0, nghost = 0;
    std::vector<double> h_s, h_alpha, h_c;  // owned input rows until init
    // batch of independent cells (empty: one cell from set_system)
    std::vector<int> cell_natms, cell_nghost;
    std::vector<double> cell_H, cell_x, cell_c;
    std::vector<int> swap_off, swap_dim, swap_dir;

    // device tables
    mappdmd::DevBuf<double> d_quad64, d_rho7, d_phi7, d_F9, d_F7, d_drho, d_pairF2;
    mappdmd::DevBuf<float> d_quad32;
    // device atoms (owned rows first, ghosts after)
    mappdmd::DevBuf<double4> d_pos, d_x0;
    mappdmd::DevBuf<double> d_alpha, d_conc;
    mappdmd::DevBuf<int> d_ghost_src, d_ghost_owner;
    // binning and lists
    mappdmd::DevBuf<int> d_flag, d_offs, d_bin_of, d_bin_cnt, d_bin_start, d_bin_atoms;
    mappdmd::DevBuf<int> d_cnt, d_off, d_nl, d_roff, d_rev;
    mappdmd::DevBuf<int> d_off2, d_nl2, d_roff2, d_rev2, d_crd;
    mappdmd::DevBuf<int> d_sort_tmp;  // out-of-place segment sort
    mappdmd::DevBuf<int> d_pair_i, d_pair_i2, d_rpos;
    mappdmd::DevBuf<double> d_inva, d_deriv;
    mappdmd::DevBuf<double4> d_arec;  // [ntot][nt] {alpha, c, 1/alpha, dE}  // deriv: [npairs][DerivRec<nt>::kLen] of the quadrature type
    long nrev = 0;
    std::vector<mappdmd::DevBuf<int>*> scan_levels;
    long npairs = 0, npairs2 = 0;
    bool have_list2 = false;
    // per-atom vectors ([ntot][nt] unless noted) and pair buffers
    mappdmd::DevBuf<double> d_E, d_dE, d_ddE, d_n, d_s, d_t, d_v, d_cd, d_fa, d_Eown, d_a, d_g;
    mappdmd::DevBuf<double> d_f, d_fown;        // [natms][3+nt]
    mappdmd::DevBuf<double> d_pairE, d_pairF;   // [npairs][nt], [npairs][3+nt]
    // per half-list entry {partner, reverse position}; the partner's top bit: inside the
    // quadrature cut at the last pass 1 (the shares of entries outside it stay zero)
    mappdmd::DevBuf<int2> d_nlrp;
    mappdmd::DevBuf<double> d_rhoc, d_phic;     // pair values for c_d_calc / calc_g, per half-list pair
    mappdmd::DevBuf<int> d_toff, d_tidx;        // site table: both ends of every contributing pair
    mappdmd::DevBuf<double> d_trho, d_tphi;
    long ntable = 0;
    mappdmd::DevBuf<int> d_foff2, d_fidx2;      // 2nd list of both ends (~owner for the partner end)
    bool cache_valid = false;
    mappdmd::DevBuf<double> d_part, d_out, d_xaos, d_stage;
    mappdmd::DevBuf<int> d_iflag;
    mappdmd::DevBuf<unsigned long long> d_total;  // 64-bit count total of counts_to_offsets
    long n_inrange = 0;

    // learned surrogate
    mappdmd::DevBuf<float> d_sp, d_sm, d_sv, d_sX, d_sT, d_sXv, d_sTv, d_sgpart, d_slpart, d_serr;
    mappdmd::DevBuf<int> d_sidx, d_nfall;
    mappdmd::DevBuf<unsigned short> d_sfrag;  // bf16 MFMA weight fragments
    mappdmd::DevBuf<unsigned> d_smask;
    mappdmd::DevBuf<float> d_sB, d_sBv, d_senv;  // sample baselines, log2 residual envelope
    mappdmd::SurNorm snorm{};
    bool sur_trained = false;
    // the validation errors of the quadrature check (k_sur_judge: |prediction -
    // quadrature| per output) and the cell of each validation sample, kept so
    // the fallback mask can be re-derived for another force bound without
    // retraining
    std::vector<float> h_sjudge;
    std::vector<int> h_scell;
    long sur_nval = 0;
    double sur_scale[8]{};  // force-bound scales of the system the networks were judged for
    double sur_eps[6]{};    // per-pair tolerances of the current mask
    double sur_bound = 0.0;
    double sur_served = 0.0;  // fraction of cells the networks serve
    bool sur_last_quadrature = false;  // the last surrogate force_calc served no cell and ran the mixed pass 1
    bool energy_full = false;  // energy_calc over every ordered type pair
    // launch shapes of the two force passes (waves per block; total waves, 0 = one per site);
    // until autotuned, the shapes the autotuner keeps on MI355X for the 64^3 workloads
    struct Shape {
        int wpb = 4;
        long waves = 0;
        int parts = 2;  // pass 1: lanes per (pair, type pair)
    };
    Shape sh1{4, 0, 2}, sh2{2, 0, 2};
    mappdmd::DevBuf<double> d_lh, d_lx0, d_lf0, d_lS, d_lY;  // minimiser vectors
    mappdmd::DevBuf<double> d_bdf;  // BDF history and solver vectors
    mappdmd::DevBuf<int> d_qidx;    // element read by the dE[jatm + jtype] index quirk

    // timing
    bool timing = false;
    hipEvent_t ev[8]{};
    double last_ms[5]{};

    ~mapp_dmd();
};

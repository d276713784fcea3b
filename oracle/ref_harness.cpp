// TEST INFRASTRUCTURE ONLY -- drives the reference's own force field classes
// (compiled from /root/reference/src by oracle/Makefile.ref into oracle/_ref/)
// on one input and dumps every result of the hot path, to pin the oracle's
// restatement against the reference bit for bit, and, with "time_force N",
// to time the reference's own force_calc as the CPU baseline of bench.py
// (where the build travels with the tree).  Nothing on the GPU path links or
// loads it.
//
// Sequence (the one Clock_BDF::init runs, src/clock_bdf.cpp:201-228): an
// input script reads the extended CFG (mode dmd, skin, read cfg), the force
// field is created and given its ff_coef line, then f is added, the
// communication list set, init, ph_setup, neighbor init + create_list(0,1),
// store_0, create_2nd_neigh_lst, force_calc(1), energy_calc, c_d_calc,
// calc_g(0, beta, a, g) and calc_g(1, beta, a, g).
//
// Tables: DMD_II reads the FS file with set_fs (all nt^2 density slots,
// src/ff_dmd_II.cpp:1253-1480).  DMD_I and DMD map "FS" to a setfl-style
// reader that leaves density slots 2 and 3 allocated but unset for two types
// (src/ff_dmd_I.cpp:751-755, :1069-1298); for those the harness runs their own
// coef on a setfl-style file of the same grid and then copies every spline
// row from a DMD_II instance that read the FS file, so all three variants run
// on identical tables.
//
// usage: ref_harness <job file>; the job file holds "key value..." lines:
//   cfg <path>  fs <path>  setfl <path>  variant DMD|DMD_I|DMD_II
//   coef r_crd no_i alpha_min alpha_max T  skin s  a <path>  beta b  out <path>
//   variant EAM | LJ | FS | FSM | MEAM: MD mode (no setfl, coef, a or beta):
//   the sequence of Min_LBFGS::init up to force_calc (src/min_l-bfgs.cpp:101-178),
//   then force_calc(1) and energy_calc; output x_read, type, H, x_all,
//   list_off, list, en_st, f, energy_calc.  Coefficients: EAM the FS file
//   ("fs"), LJ the ff_coef lines ("lj": five tokens a line, one line a type
//   pair), FS and FSM the parameter file ("params"), MEAM the library and
//   local files ("lib", "local")
//   time_force N (optional: after create_2nd_neigh_lst, N calls of
//   force_calc(1), each timed on the steady clock; output t_force [N] seconds
//   and en_st only)  go <path> (optional, with time_force: create
//   "<out>.ready" once the system is built, then start the timed calls when
//   <path> exists, so that several processes time their calls together)
// Output: for every array a line "name rows cols" followed by rows*cols
// little-endian doubles.
#include <chrono>
#include <unistd.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "MAPP.h"
#include "atom_types.h"
#include "atoms.h"
#include "error.h"
#include "ff_dmd.h"
#include "ff_dmd_I.h"
#include "ff_dmd_II.h"
#include "ff_eam.h"
#include "ff_fs.h"
#include "ff_fsm.h"
#include "ff_lj.h"
#include "ff_meam.h"
#include "neighbor.h"

using namespace MAPP_NS;

namespace {

// table access to a DMD_II that read the FS file
struct DonorII : ForceField_DMD_II {
    explicit DonorII(MAPP* m) : ForceField_DMD_II(m) {}
    TYPE0*** rho_t() { return rho_arr; }
    TYPE0*** phi_t() { return phi_r_arr; }
    TYPE0*** F_t() { return F_arr; }
    int nr_() const { return nr; }
    int nrho_() const { return nrho; }
    TYPE0 dr_() const { return dr; }
    TYPE0 drho_() const { return drho; }
};

// a DMD_I / DMD force field whose tables are replaced by the donor's, and
// read access to its 2nd neighbour list
template <class FF>
struct Taker : FF {
    explicit Taker(MAPP* m) : FF(m) {}
    void take(DonorII* d, int nt) {
        if (this->nr != d->nr_() || this->nrho != d->nrho_() || this->dr != d->dr_() || this->drho != d->drho_()) {
            fprintf(stderr, "ref_harness: setfl and FS grids differ\n");
            exit(3);
        }
        for (int s = 0; s < nt * nt; s++)
            for (int r = 0; r < this->nr; r++)
                for (int k = 0; k < 7; k++) this->rho_arr[s][r][k] = d->rho_t()[s][r][k];
        for (int s = 0; s < nt * (nt + 1) / 2; s++)
            for (int r = 0; r < this->nr; r++)
                for (int k = 0; k < 7; k++) this->phi_r_arr[s][r][k] = d->phi_t()[s][r][k];
        for (int t = 0; t < nt; t++)
            for (int r = 0; r < this->nrho + 1; r++)
                for (int k = 0; k < 9; k++) this->F_arr[t][r][k] = d->F_t()[t][r][k];
    }
    int list2_size(int i) { return this->neigh_lst_sz[i]; }
    int* list2(int i) { return this->neigh_lst[i]; }
};
// the same read access for DMD_II
struct ListII : ForceField_DMD_II {
    explicit ListII(MAPP* m) : ForceField_DMD_II(m) {}
    int list2_size(int i) { return neigh_lst_sz[i]; }
    int* list2(int i) { return neigh_lst[i]; }
};

struct Out {
    FILE* fp;
    void put(const char* name, int rows, int cols, const double* v) {
        fprintf(fp, "%s %d %d\n", name, rows, cols);
        fwrite(v, sizeof(double), (size_t)rows * cols, fp);
        fputc('\n', fp);
    }
    template <typename T>
    void put_vec(const char* name, int rows, int cols, const T* v) {
        std::vector<double> d((size_t)rows * cols);
        for (size_t k = 0; k < d.size(); k++) d[k] = (double)v[k];
        put(name, rows, cols, d.data());
    }
};

std::vector<double> read_values(const std::string& path) {
    std::vector<double> v;
    FILE* fp = fopen(path.c_str(), "r");
    if (!fp) {
        fprintf(stderr, "ref_harness: cannot read %s\n", path.c_str());
        exit(2);
    }
    double x;
    while (fscanf(fp, "%lf", &x) == 1) v.push_back(x);
    fclose(fp);
    return v;
}

char* dup(const std::string& s) {
    char* p = new char[s.size() + 1];
    memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

}  // namespace

// MD mode: plain EAM (ForceField_eam with its FS reader, src/ff_eam.cpp:400-430),
// LJ (src/ff_lj.cpp:70-100), FS / FSM (src/ff_fs.cpp:113, src/ff_fsm.cpp:111),
// MEAM (src/ff_meam.cpp:2593)
int run_md(std::map<std::string, std::vector<std::string>>& job, const std::string& variant) {
    const std::string script = job["out"].at(0) + ".in";
    {
        FILE* fp = fopen(script.c_str(), "w");
        fprintf(fp, "mode md\nskin %s\nread cfg %s\n", job["skin"].at(0).c_str(), job["cfg"].at(0).c_str());
        fclose(fp);
    }
    char* margs[] = {dup("ref_harness"), dup("-i"), dup(script), dup("-o"), dup(job["out"].at(0) + ".log")};
    MAPP* mapp = new MAPP(5, margs, MPI_COMM_WORLD);
    Atoms* atoms = mapp->atoms;
    const int natms = atoms->natms;
    const int xdim = atoms->vectors[0].dim;
    const int type_n = atoms->find("type");
    Out out{fopen(job["out"].at(0).c_str(), "wb")};
    {
        TYPE0* x;
        atoms->vectors[0].ret(x);
        out.put("x_read", natms, xdim, x);
        int* type;
        atoms->vectors[type_n].ret(type);
        out.put_vec("type", natms, 1, type);
        double H[9];
        for (int i = 0; i < 9; i++) H[i] = atoms->H[i / 3][i % 3];
        out.put("H", 3, 3, H);
    }
    ForceField* ff = nullptr;
    if (variant == "EAM") {
        ForceField_eam* f = new ForceField_eam(mapp);
        mapp->forcefield = ff = f;
        char* a[] = {dup("ff_coef"), dup("FS"), dup(job["fs"].at(0))};
        f->coef(3, a);
    } else if (variant == "LJ") {
        ForceField_lj* f = new ForceField_lj(mapp);
        mapp->forcefield = ff = f;
        const std::vector<std::string>& t = job["lj"];
        for (size_t k = 0; k + 5 <= t.size(); k += 5) {
            char* a[] = {dup("ff_coef"), dup(t[k]), dup(t[k + 1]), dup(t[k + 2]), dup(t[k + 3]), dup(t[k + 4])};
            f->coef(6, a);
        }
    } else if (variant == "FS" || variant == "FSM") {
        char* a[] = {dup("ff_coef"), dup(job["params"].at(0))};
        if (variant == "FS") {
            ForceField_fs* f = new ForceField_fs(mapp);
            mapp->forcefield = ff = f;
            f->coef(2, a);
        } else {
            ForceField_fsm* f = new ForceField_fsm(mapp);
            mapp->forcefield = ff = f;
            f->coef(2, a);
        }
    } else if (variant == "MEAM") {
        ForceField_meam* f = new ForceField_meam(mapp);
        mapp->forcefield = ff = f;
        char* a[] = {dup("ff_coef"), dup(job["lib"].at(0)), dup(job["local"].at(0))};
        f->coef(3, a);
    } else {
        fprintf(stderr, "ref_harness: unknown variant %s\n", variant.c_str());
        return 2;
    }
    if (atoms->find_exist("f") < 0) atoms->add<TYPE0>(0, xdim, "f");
    VecLst* vecs_comm = new VecLst(mapp, 2, 0, type_n);
    vecs_comm->add_update(0);
    atoms->reset_comm(vecs_comm);
    ff->init();
    atoms->ph_setup(1, vecs_comm);
    mapp->neighbor->init();
    mapp->neighbor->create_list(0, 1);
    atoms->store_0();
    {
        TYPE0* x;
        atoms->vectors[0].ret(x);
        out.put("x_all", natms + atoms->natms_ph, xdim, x);
        std::vector<double> off(natms + 1, 0.0), ent;
        for (int i = 0; i < natms; i++) {
            off[i + 1] = off[i] + mapp->neighbor->neighbor_list_size[i];
            for (int k = 0; k < mapp->neighbor->neighbor_list_size[i]; k++)
                ent.push_back(mapp->neighbor->neighbor_list[i][k]);
        }
        out.put("list_off", natms + 1, 1, off.data());
        out.put("list", (int)ent.size(), 1, ent.data());
    }
    // the MD force fields add to f; the drivers clear it before every call
    // (src/min_l-bfgs.cpp:232-236, :690-696)
    TYPE0* f;
    atoms->vectors[atoms->find("f")].ret(f);
    for (int i = 0; i < natms * xdim; i++) f[i] = 0.0;
    double en[7] = {0, 0, 0, 0, 0, 0, 0};
    ff->force_calc(1, en);
    out.put("en_st", 1, 7, en);
    atoms->vectors[atoms->find("f")].ret(f);
    out.put("f", natms, xdim, f);
    const double ue = ff->energy_calc();
    out.put("energy_calc", 1, 1, &ue);
    fclose(out.fp);
    MPI_Finalize();
    return 0;
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    if (argc != 2) {
        fprintf(stderr, "usage: ref_harness <job file>\n");
        return 2;
    }
    std::map<std::string, std::vector<std::string>> job;
    {
        FILE* fp = fopen(argv[1], "r");
        if (!fp) return 2;
        char line[4096];
        while (fgets(line, sizeof line, fp)) {
            std::vector<std::string> tok;
            for (char* t = strtok(line, " \t\r\n"); t; t = strtok(nullptr, " \t\r\n")) tok.push_back(t);
            if (tok.empty()) continue;
            job[tok[0]] = std::vector<std::string>(tok.begin() + 1, tok.end());
        }
        fclose(fp);
    }
    const std::string variant = job["variant"].at(0);
    if (variant == "EAM" || variant == "LJ" || variant == "FS" || variant == "FSM" || variant == "MEAM")
        return run_md(job, variant);
    const std::vector<std::string>& coef = job["coef"];

    // the input script: mode, skin, read cfg
    const std::string script = job["out"].at(0) + ".in";
    {
        FILE* fp = fopen(script.c_str(), "w");
        fprintf(fp, "mode dmd\nskin %s\nread cfg %s\n", job["skin"].at(0).c_str(), job["cfg"].at(0).c_str());
        fclose(fp);
    }
    char* margs[] = {dup("ref_harness"), dup("-i"), dup(script), dup("-o"), dup(job["out"].at(0) + ".log")};
    MAPP* mapp = new MAPP(5, margs, MPI_COMM_WORLD);
    Atoms* atoms = mapp->atoms;
    const int nt = mapp->atom_types->no_types;
    const int natms = atoms->natms;
    const int xdim = atoms->vectors[0].dim;

    const int n_time = job.count("time_force") ? atoi(job["time_force"].at(0).c_str()) : 0;
    Out out{fopen(job["out"].at(0).c_str(), "wb")};
    if (n_time == 0) {
        TYPE0* x;
        atoms->vectors[0].ret(x);
        out.put("x_read", natms, xdim, x);
        TYPE0* c;
        atoms->vectors[atoms->find("c")].ret(c);
        out.put("c_read", natms, nt, c);
        double H[9];
        for (int i = 0; i < 9; i++) H[i] = atoms->H[i / 3][i % 3];
        out.put("H", 3, 3, H);
    }

    auto coef_args = [&](const std::string& kind, const std::string& file) {
        std::vector<char*> a;
        a.push_back(dup("ff_coef"));
        for (const auto& s : coef) a.push_back(dup(s));
        a.push_back(dup(kind));
        a.push_back(dup(file));
        return a;
    };

    ForceField* ff = nullptr;
    int (*l2size)(ForceField*, int) = nullptr;
    int* (*l2)(ForceField*, int) = nullptr;
    if (variant == "DMD_II") {
        ListII* f = new ListII(mapp);
        mapp->forcefield = f;
        auto a = coef_args("FS", job["fs"].at(0));
        f->coef((int)a.size(), a.data());
        ff = f;
        l2size = [](ForceField* p, int i) { return static_cast<ListII*>(p)->list2_size(i); };
        l2 = [](ForceField* p, int i) { return static_cast<ListII*>(p)->list2(i); };
    } else {
        DonorII* donor = new DonorII(mapp);
        auto a = coef_args("FS", job["fs"].at(0));
        donor->coef((int)a.size(), a.data());
        auto b = coef_args("FS", job["setfl"].at(0));
        if (variant == "DMD_I") {
            Taker<ForceField_DMD_I>* f = new Taker<ForceField_DMD_I>(mapp);
            mapp->forcefield = f;
            f->coef((int)b.size(), b.data());
            f->take(donor, nt);
            ff = f;
            l2size = [](ForceField* p, int i) { return static_cast<Taker<ForceField_DMD_I>*>(p)->list2_size(i); };
            l2 = [](ForceField* p, int i) { return static_cast<Taker<ForceField_DMD_I>*>(p)->list2(i); };
        } else if (variant == "DMD") {
            Taker<ForceField_DMD>* f = new Taker<ForceField_DMD>(mapp);
            mapp->forcefield = f;
            f->coef((int)b.size(), b.data());
            f->take(donor, nt);
            ff = f;
            l2size = [](ForceField* p, int i) { return static_cast<Taker<ForceField_DMD>*>(p)->list2_size(i); };
            l2 = [](ForceField* p, int i) { return static_cast<Taker<ForceField_DMD>*>(p)->list2(i); };
        } else {
            fprintf(stderr, "ref_harness: unknown variant %s\n", variant.c_str());
            return 2;
        }
    }

    // Clock_BDF::init up to force_calc
    int f_n = atoms->find_exist("f");
    if (f_n < 0) f_n = atoms->add<TYPE0>(0, xdim, "f");
    const int c_n = atoms->find("c");
    const int c_d_n = atoms->find("c_d");
    VecLst* vecs_comm = new VecLst(mapp, 3, 0, c_n, c_d_n);
    vecs_comm->add_update(0);
    atoms->reset_comm(vecs_comm);
    ff->init();
    atoms->ph_setup(1, vecs_comm);
    mapp->neighbor->init();
    mapp->neighbor->create_list(0, 1);
    atoms->store_0();
    ff->create_2nd_neigh_lst();

    if (n_time > 0) {
        if (job.count("go")) {
            const std::string ready = job["out"].at(0) + ".ready";
            FILE* fr = fopen(ready.c_str(), "w");
            if (fr) fclose(fr);
            const std::string go = job["go"].at(0);
            while (access(go.c_str(), F_OK) != 0) usleep(1000);
        }
        double en[7] = {0, 0, 0, 0, 0, 0, 0};
        std::vector<double> t(n_time);
        for (int k = 0; k < n_time; k++) {
            const auto t0 = std::chrono::steady_clock::now();
            ff->force_calc(1, en);
            t[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        out.put("t_force", n_time, 1, t.data());
        out.put("en_st", 1, 7, en);
        fclose(out.fp);
        MPI_Finalize();
        return 0;
    }

    const int nph = atoms->natms_ph;
    {
        TYPE0* x;
        atoms->vectors[0].ret(x);
        out.put("x_all", natms + nph, xdim, x);
        std::vector<double> off(natms + 1, 0.0), ent;
        for (int i = 0; i < natms; i++) {
            off[i + 1] = off[i] + mapp->neighbor->neighbor_list_size[i];
            for (int k = 0; k < mapp->neighbor->neighbor_list_size[i]; k++)
                ent.push_back(mapp->neighbor->neighbor_list[i][k]);
        }
        out.put("list_off", natms + 1, 1, off.data());
        out.put("list", (int)ent.size(), 1, ent.data());
        std::vector<double> off2(natms + 1, 0.0), ent2;
        for (int i = 0; i < natms; i++) {
            off2[i + 1] = off2[i] + l2size(ff, i);
            for (int k = 0; k < l2size(ff, i); k++) ent2.push_back(l2(ff, i)[k]);
        }
        out.put("list2_off", natms + 1, 1, off2.data());
        out.put("list2", (int)ent2.size(), 1, ent2.data());
        int* crd;
        atoms->vectors[atoms->find("crd")].ret(crd);
        out.put_vec("crd", natms, 1, crd);
    }

    auto put_vec = [&](const char* name, const char* vec, int cols) {
        TYPE0* v;
        atoms->vectors[atoms->find(vec)].ret(v);
        out.put(name, natms, cols, v);
    };

    double en[7] = {0, 0, 0, 0, 0, 0, 0};
    ff->force_calc(1, en);
    out.put("en_st", 1, 7, en);
    put_vec("f", "f", xdim);
    put_vec("E_force", "E", nt);
    put_vec("dE_force", "dE", nt);

    const double ue = ff->energy_calc();
    out.put("energy_calc", 1, 1, &ue);

    ff->c_d_calc();
    put_vec("c_d", "c_d", nt);
    put_vec("n_cd", "n", nt);
    put_vec("E_cd", "E", nt);
    put_vec("dE_cd", "dE", nt);

    std::vector<double> a = read_values(job["a"].at(0));
    if ((int)a.size() != natms * nt) {
        fprintf(stderr, "ref_harness: a has %zu values, expected %d\n", a.size(), natms * nt);
        return 2;
    }
    const double beta = atof(job["beta"].at(0).c_str());
    std::vector<double> g(natms * nt, 0.0);
    const double cost0 = ff->calc_g(0, beta, a.data(), g.data());
    out.put("cost0", 1, 1, &cost0);
    out.put("g", natms, nt, g.data());
    put_vec("s_g", "s", nt);
    put_vec("n_g", "n", nt);
    put_vec("c_d_after_g", "c_d", nt);
    std::vector<double> g1(natms * nt, 0.0);
    const double cost1 = ff->calc_g(1, beta, a.data(), g1.data());
    out.put("cost1", 1, 1, &cost1);
    fclose(out.fp);
    MPI_Finalize();
    return 0;
}

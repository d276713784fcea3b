// MEAM set-up on the host: the library and local parameter files of
// ForceField_meam::read_global / read_local (src/ff_meam.cpp:2660-3191) and
// coef (:2593-2627) -- augt1, alloy_params (:3192-3249), the reference
// densities (:3844-3880) and the pair table of compute_pair_meam
// (:3252-3410) with phi_meam (:3629-3842).  The kernels read MeamDev and the
// pair table.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mapp_dmd.h"

namespace mappdmd {

constexpr int kMeamT = MAPP_MEAM_MAXT;
constexpr int kMeamNr = 1000;  // rows of the pair table (src/ff_meam.cpp:45)
constexpr int kMeamRec = 38;   // per-site record (src/ff_meam.h:46-100)
enum { kLatFcc = 0, kLatBcc, kLatHcp, kLatDim, kLatDia, kLatB1, kLatC11, kLatL12, kLatB2 };

// what the kernels read (a kernel argument)
struct MeamDev {
    int nt, ialloy, emb_lin_neg, bkgd_dyn, fix;
    double rc, delr_inv, dr_inv;
    double Z[kMeamT], A[kMeamT], rho0[kMeamT], beta[kMeamT][4], t[kMeamT][3], rho_ref[kMeamT];
    double Ec[kMeamT], re[kMeamT], inv_re[kMeamT], shape[kMeamT][3], gsmooth;
    int ibar[kMeamT];
    double ebound[kMeamT][kMeamT], cmin[kMeamT][kMeamT][kMeamT], cmax[kMeamT][kMeamT][kMeamT];
};

namespace meam_host {

inline int lattice_code(const char* s) {
    static const char* names[] = {"fcc", "bcc", "hcp", "dim", "dia", "b1", "c11", "l12", "b2"};
    for (int i = 0; i < 9; ++i)
        if (!std::strcmp(s, names[i])) return i;
    return -1;
}

// MAPP::parse_line (src/MAPP.cpp:468-530): blank-separated tokens, '#' at the
// start of a token ends the line
inline std::vector<std::string> tokens(const std::string& line) {
    std::vector<std::string> out;
    size_t c = 0;
    while (c < line.size() && line[c] != '#') {
        if (std::isspace((unsigned char)line[c])) {
            ++c;
            continue;
        }
        size_t b = c;
        while (c < line.size() && !std::isspace((unsigned char)line[c])) ++c;
        out.push_back(line.substr(b, c - b));
    }
    return out;
}

inline bool read_lines(const char* path, std::vector<std::string>* lines) {
    FILE* fp = std::fopen(path, "r");
    if (!fp) return false;
    char buf[4096];
    while (std::fgets(buf, sizeof buf, fp)) lines->push_back(buf);
    std::fclose(fp);
    return true;
}

// read_global: 19-token entries, quotes read as blanks; entries of elements
// that are not among the types are skipped; then the defaults of the local
// settings.  Returns "" or an error.
inline std::string read_library(const char* path, int nt, const char* const* elements, mapp_meam_params* p,
                                std::vector<int>* type_ref) {
    std::vector<std::string> lines;
    if (!read_lines(path, &lines)) return std::string("file ") + path + " not found";
    std::memset(p, 0, sizeof *p);
    p->nt = nt;
    for (int i = 0; i < kMeamT; ++i)
        for (int j = 0; j < kMeamT; ++j) p->lattice[i][j] = -1;
    std::vector<std::string> args;
    std::vector<int> seen(nt, 0);
    for (std::string line : lines) {
        for (char& ch : line)
            if (ch == '\'') ch = ' ';
        for (auto& t : tokens(line)) args.push_back(t);
        if (args.size() > 19) return "wrong number of arguments";
        if (args.size() < 19) continue;
        int it = -1;
        for (int i = 0; i < nt; ++i)
            if (args[0] == elements[i]) it = i;
        if (it >= 0) {
            type_ref->push_back(it);
            seen[it] = 1;
            const int lat = lattice_code(args[1].c_str());
            if (lat < 0 || lat > kLatDia) return "unknown lattice";
            p->lattice[it][it] = lat;
            p->Z[it] = std::atof(args[2].c_str());
            p->ielt[it] = std::atoi(args[3].c_str());
            p->alpha[it][it] = std::atof(args[5].c_str());
            p->beta0[it] = std::atof(args[6].c_str());
            p->beta1[it] = std::atof(args[7].c_str());
            p->beta2[it] = std::atof(args[8].c_str());
            p->beta3[it] = std::atof(args[9].c_str());
            const double alat = std::atof(args[10].c_str());
            // nearest-neighbour distance of the element's lattice (:2766-2776)
            if (lat == kLatFcc) p->re[it][it] = alat / std::sqrt(2.0);
            else if (lat == kLatBcc) p->re[it][it] = alat * 0.5 * std::sqrt(3.0);
            else if (lat == kLatDia) p->re[it][it] = alat * 0.25 * std::sqrt(3.0);
            else p->re[it][it] = alat;
            p->Ec[it][it] = std::atof(args[11].c_str());
            p->A[it] = std::atof(args[12].c_str());
            p->t0[it] = std::atof(args[13].c_str());
            p->t1[it] = std::atof(args[14].c_str());
            p->t2[it] = std::atof(args[15].c_str());
            p->t3[it] = std::atof(args[16].c_str());
            p->rho0[it] = std::atof(args[17].c_str());
            p->ibar[it] = std::atoi(args[18].c_str());
        }
        args.clear();
    }
    for (int i = 0; i < nt; ++i)
        if (!seen[i]) return std::string("element ") + elements[i] + " is not in the library";
    for (int i = 0; i < kMeamT; ++i)
        for (int j = 0; j < kMeamT; ++j) {
            for (int k = 0; k < kMeamT; ++k) {
                p->cmin[i][j][k] = 2.0;
                p->cmax[i][j][k] = 2.8;
            }
            p->ebound[i][j] = 1.96 / 1.8;
            p->zbl[i][j] = 1;
        }
    p->augt1 = 1;
    p->delr_inv = 10.0;
    p->delr = 0.1;
    p->rc = 4.0;
    p->gsmooth = 99.0;
    return "";
}

// read_local: keyword(indices) = value lines
inline std::string read_local(const char* path, const std::vector<int>& type_ref, mapp_meam_params* p) {
    if (!std::strcmp(path, "NULL")) return "";
    std::vector<std::string> lines;
    if (!read_lines(path, &lines)) return std::string("file ") + path + " not found";
    const int nref = (int)type_ref.size();
    for (std::string line : lines) {
        for (char& ch : line)
            if (ch == '\'' || ch == ',' || ch == '(' || ch == ')' || ch == '=') ch = ' ';
        const auto a = tokens(line);
        if (a.empty()) continue;
        const std::string& key = a[0];
        const int n = (int)a.size();
        int ti[3] = {0, 0, 0};
        auto idx = [&](int count) -> bool {
            for (int q = 0; q < count; ++q) {
                const int v = std::atoi(a[1 + q].c_str()) - 1;
                if (v < 0 || v >= nref) return false;
                ti[q] = type_ref[v];
            }
            return true;
        };
        auto need = [&](int count) { return n == count; };
        const double val = std::atof(a[n - 1].c_str());
        if (key == "Ec" || key == "alpha" || key == "re") {
            if (!need(4) || !idx(2)) return "wrong command: " + key;
            double(&m)[kMeamT][kMeamT] = key == "Ec" ? p->Ec : key == "alpha" ? p->alpha : p->re;
            m[ti[0]][ti[1]] = val;
            m[ti[1]][ti[0]] = val;
        } else if (key == "rho0") {
            if (!need(3) || !idx(1)) return "wrong command: " + key;
            p->rho0[ti[0]] = val;
        } else if (key == "delta" || key == "attrac" || key == "repuls") {
            if (!need(4) || !idx(2)) return "wrong command: " + key;
            double(&m)[kMeamT][kMeamT] = key == "delta" ? p->delta : key == "attrac" ? p->attrac : p->repuls;
            m[ti[0]][ti[1]] = val;
        } else if (key == "lattce") {
            if (!need(4) || !idx(2)) return "wrong command: " + key;
            const int lat = lattice_code(a[3].c_str());
            if (lat < 0) return "unknown lattice " + a[3];
            p->lattice[ti[0]][ti[1]] = lat;
            p->lattice[ti[1]][ti[0]] = lat;
        } else if (key == "nn2") {
            if (!need(4) || !idx(2)) return "wrong command: " + key;
            p->nn2[ti[0]][ti[1]] = std::atoi(a[3].c_str());
            p->nn2[ti[1]][ti[0]] = p->nn2[ti[0]][ti[1]];
        } else if (key == "Cmin" || key == "Cmax") {
            if (!need(5) || !idx(3)) return "wrong command: " + key;
            double(&m)[kMeamT][kMeamT][kMeamT] = key == "Cmin" ? p->cmin : p->cmax;
            m[ti[0]][ti[1]][ti[2]] = val;
            m[ti[1]][ti[0]][ti[2]] = val;
        } else if (key == "rc" || key == "delr" || key == "gsmooth_factor") {
            if (!need(2)) return "wrong command: " + key;
            (key == "rc" ? p->rc : key == "delr" ? p->delr : p->gsmooth) = val;
        } else if (key == "augt1" || key == "ialloy" || key == "mixture_ref_t" || key == "erose_form" ||
                   key == "emb_lin_neg" || key == "bkgd_dyn") {
            if (!need(2)) return "wrong command: " + key;
            const int v = std::atoi(a[1].c_str());
            if (key == "augt1") p->augt1 = v;
            else if (key == "ialloy") p->ialloy = v;
            else if (key == "mixture_ref_t") p->mix_ref_t = v;
            else if (key == "erose_form") p->erose_form = v;
            else p->emb_lin_neg = v;  // bkgd_dyn lands in emb_lin_neg (:3158-3163)
        } else {
            return "wrong input: " + key;  // zbl included (:3138-3141)
        }
    }
    return "";
}

// ---------------------------------------------------------------- derived tables
struct Setup {
    mapp_meam_params p;  // after augt1 and the alloy rules
    double rho_ref[kMeamT] = {0, 0, 0, 0};
    double dr = 0.0, dr_inv = 0.0;
    std::vector<double> table;  // [ncomp][kMeamNr][7]
    bool bad = false;

    static int comp(int i, int j) { return i < j ? (j + 1) * j / 2 + i : (i + 1) * i / 2 + j; }

    double G(double gam, int ibar, double* dG = nullptr) {
        const double gs = p.gsmooth;
        double g = 0.0, d = 0.0;
        if (ibar == 0 || ibar == 4) {
            const double sw = -gs / (gs + 1.0);
            g = gam < sw ? std::sqrt(std::pow(sw / gam, gs) / (gs + 1.0)) : std::sqrt(1.0 + gam);
            d = 0.5 / g;
        } else if (ibar == 1) {
            g = std::exp(gam * 0.5);
            d = 0.5 * g;
        } else if (ibar == 3) {
            g = 2.0 / (1.0 + std::exp(-gam));
            d = g * (1.0 - 0.5 * g);
        } else if (ibar == -5) {
            if ((1.0 + gam) >= 0.0) {
                g = std::sqrt(1.0 + gam);
                d = 0.5 / g;
            } else {
                g = -std::sqrt(-1.0 - gam);
                d = -0.5 / g;
            }
        } else {
            bad = true;
        }
        if (dG) *dG = d;
        return g;
    }
    static double fcut(double x) {
        if (x >= 1.0) return 1.0;
        if (x <= 0.0) return 0.0;
        double a = 1.0 - x;
        a *= a;
        a *= a;
        a = 1.0 - a;
        return a * a;
    }
    bool shape(int lat, double* s) const {
        s[0] = s[1] = s[2] = 0.0;
        if (lat == kLatFcc || lat == kLatBcc || lat == kLatB1 || lat == kLatB2) return true;
        if (lat == kLatHcp) s[2] = 1.0 / 3.0;
        else if (lat == kLatDia) s[2] = 32.0 / 9.0;
        else if (lat == kLatDim) {
            s[0] = 1.0;
            s[1] = 2.0 / 3.0;
            s[2] = 0.4;
        } else return false;
        return true;
    }
    static double z1(int lat) {
        static const double z[9] = {12.0, 8.0, 12.0, 1.0, 4.0, 6.0, 10.0, 12.0, 8.0};
        return lat >= 0 && lat < 9 ? z[lat] : 0.0;
    }
    bool z2(int lat, double cmin, double cmax, double* Z2, double* arat, double* S) const {
        int num = 4;
        switch (lat) {
            case kLatFcc: case kLatHcp: case kLatL12: *Z2 = 6.0; *arat = std::sqrt(2.0); break;
            case kLatBcc: case kLatB2: *Z2 = 6.0; *arat = 2.0 / std::sqrt(3.0); break;
            case kLatDim: *Z2 = 0.0; *arat = 1.0; *S = 0.0; return true;
            case kLatDia: *Z2 = 0.0; *arat = std::sqrt(8.0 / 3.0); break;
            case kLatB1: *Z2 = 12.0; *arat = std::sqrt(2.0); num = 2; break;
            default: return false;
        }
        const double C = 4.0 / (*arat * *arat) - 1.0;
        *S = std::pow(fcut((C - cmin) / (cmax - cmin)), num);
        return true;
    }
    double sijk(double C, int i, int j, int k) const {
        return fcut((C - p.cmin[i][j][k]) / (p.cmax[i][j][k] - p.cmin[i][j][k]));
    }
    static double zbl(double r, double z1_, double z2_) {
        const double a = 0.4685 / (std::pow(z1_, 0.23) + std::pow(z2_, 0.23));
        const double x = r / a;
        double v = 0.028171 * std::exp(-0.20162 * x) + 0.28022 * std::exp(-0.40290 * x) +
                   0.50986 * std::exp(-0.94229 * x) + 0.18175 * std::exp(-3.1998 * x);
        if (r > 0.0) v *= z1_ * z2_ * 14.3997 / r;
        return v;
    }
    double erose(double r, int a, int b) const {
        if (!(r > 0.0)) return 0.0;
        const double re = p.re[a][b], Ec = p.Ec[a][b], rep = p.repuls[a][b], att = p.attrac[a][b];
        const double as = p.alpha[a][b] * (r / re - 1.0);
        const double a3 = as >= 0.0 ? att : rep;
        if (p.erose_form == 1) return -Ec * (1.0 + as + (rep / r - att) * as * as * as) * std::exp(-as);
        if (p.erose_form == 2) return -Ec * (1.0 + as + a3 * as * as * as) * std::exp(-as);
        return -Ec * (1.0 + as + a3 * (as * as * as) / (r / re)) * std::exp(-as);
    }

    // partial densities of the a-b reference structure (get_dens_ref); d[0..3]
    // at a, d[4..7] at b
    void dens_ref(double r, int a, int b, double* d) {
        double a1 = r / p.re[a][a] - 1.0, a2 = r / p.re[b][b] - 1.0;
        const double* ba[4] = {&p.beta0[a], &p.beta1[a], &p.beta2[a], &p.beta3[a]};
        const double* bb[4] = {&p.beta0[b], &p.beta1[b], &p.beta2[b], &p.beta3[b]};
        double ra[4], rb[4];
        for (int q = 0; q < 4; ++q) {
            ra[q] = p.rho0[a] * std::exp(-*ba[q] * a1);
            rb[q] = p.rho0[b] * std::exp(-*bb[q] * a2);
        }
        for (int q = 0; q < 8; ++q) d[q] = 0.0;
        const int lat = p.lattice[a][b];
        switch (lat) {
            case kLatFcc: d[0] = 12.0 * rb[0]; d[4] = 12.0 * ra[0]; break;
            case kLatBcc: case kLatB2: d[0] = 8.0 * rb[0]; d[4] = 8.0 * ra[0]; break;
            case kLatB1: d[0] = 6.0 * rb[0]; d[4] = 6.0 * ra[0]; break;
            case kLatDia:
                d[0] = 4.0 * rb[0]; d[4] = 4.0 * ra[0];
                d[3] = 32.0 / 9.0 * rb[3] * rb[3]; d[7] = 32.0 / 9.0 * ra[3] * ra[3];
                break;
            case kLatHcp:
                d[0] = 12.0 * rb[0]; d[4] = 12.0 * ra[0];
                d[3] = 1.0 / 3.0 * rb[3] * rb[3]; d[7] = 1.0 / 3.0 * ra[3] * ra[3];
                break;
            case kLatDim: {
                double s[3];
                shape(lat, s);
                d[0] = rb[0]; d[4] = ra[0];
                d[1] = s[0] * rb[1] * rb[1]; d[5] = s[0] * ra[1] * ra[1];
                d[2] = s[1] * rb[2] * rb[2]; d[6] = s[1] * ra[2] * ra[2];
                d[3] = s[2] * rb[3] * rb[3]; d[7] = s[2] * ra[3] * ra[3];
                break;
            }
            case kLatC11:
                for (int q = 0; q < 4; ++q) {
                    d[q] = ra[q];
                    d[4 + q] = rb[q];
                }
                break;
            case kLatL12:
                d[0] = 8.0 * ra[0] + 4.0 * rb[0];
                d[4] = 12.0 * ra[0];
                if (p.ialloy == 1) {
                    d[2] = (8.0 / 3.0) * (ra[2] * p.t2[a] - rb[2] * p.t2[b]) * (ra[2] * p.t2[a] - rb[2] * p.t2[b]);
                    const double den = 8.0 * ra[0] * p.t2[a] * p.t2[a] + 4.0 * rb[0] * p.t2[b] * p.t2[b];
                    if (den > 0.0) d[2] *= d[0] / den;
                } else {
                    d[2] = 8.0 / 3.0 * (ra[2] - rb[2]) * (ra[2] - rb[2]);
                }
                break;
            default: bad = true; return;
        }
        if (p.nn2[a][b] == 1) {
            double Z2, arat, scrn;
            if (!z2(lat, p.cmin[a][a][b], p.cmax[a][a][b], &Z2, &arat, &scrn)) { bad = true; return; }
            a1 = arat * r / p.re[a][a] - 1.0;
            a2 = arat * r / p.re[b][b] - 1.0;
            const double na = p.rho0[a] * std::exp(-p.beta0[a] * a1), nb = p.rho0[b] * std::exp(-p.beta0[b] * a2);
            if (lat == kLatL12) {
                const double s111 = sijk(1.0, a, a, a), s112 = sijk(1.0, a, a, b), s221 = sijk(1.0, b, b, a);
                const double S11 = s111 * s111 * s112 * s112, S22 = s221 * s221 * s221 * s221;
                d[0] += 6.0 * S11 * na;
                d[4] += 6.0 * S22 * nb;
            } else {
                d[0] += Z2 * scrn * na;
                if (!z2(lat, p.cmin[b][b][a], p.cmax[b][b][a], &Z2, &arat, &scrn)) { bad = true; return; }
                d[4] += Z2 * scrn * nb;
            }
        }
    }

    double emb(int a, double rb) const {
        if (rb == 0.0) return 0.0;
        if (p.emb_lin_neg == 1 && rb <= 0.0) return -p.A[a] * p.Ec[a][a] * rb;
        return p.A[a] * p.Ec[a][a] * rb * std::log(rb);
    }

    double phi(double r, int a, int b) {
        const int lat = p.lattice[a][b];
        const double Z12 = z1(lat);
        double d[8];
        dens_ref(r, a, b, d);
        if (d[0] <= 1.0e-14 && d[4] <= 1.0e-14) return 0.0;
        double t[6] = {0, 0, 0, 0, 0, 0};  // t11 t21 t31 t12 t22 t32
        const double ta[3] = {p.t1[a], p.t2[a], p.t3[a]}, tb[3] = {p.t1[b], p.t2[b], p.t3[b]};
        if (p.ialloy == 2) {
            for (int q = 0; q < 3; ++q) {
                t[q] = ta[q];
                t[3 + q] = tb[q];
            }
        } else if (lat == kLatC11) {
            const double sc = 1.0 / (d[0] + d[4]);
            for (int q = 0; q < 3; ++q) t[q] = t[3 + q] = sc * (ta[q] * d[0] + tb[q] * d[4]);
        } else if (lat == kLatL12) {
            const double a1 = r / p.re[a][a] - 1.0, a2 = r / p.re[b][b] - 1.0;
            const double ra0 = p.rho0[a] * std::exp(-p.beta0[a] * a1), rb0 = p.rho0[b] * std::exp(-p.beta0[b] * a2);
            const double r01 = 8.0 * ra0 + 4.0 * rb0;
            for (int q = 0; q < 3; ++q) {
                t[q] = (8.0 * ta[q] * ra0 + 4.0 * tb[q] * rb0) / r01;
                t[3 + q] = ta[q];
            }
        } else if (lat >= 0 && lat <= kLatB2) {
            for (int q = 0; q < 3; ++q) {
                t[q] = tb[q];
                t[3 + q] = ta[q];
            }
        } else {
            bad = true;
        }
        double rb1, rb2;
        if (lat == kLatC11) {
            if (p.lattice[a][a] == kLatDia) {
                rb1 = 0.25 * Z12 * Z12 * (d[4] + d[0]) * (d[4] + d[0]) + t[0] * (d[5] - d[1]) * (d[5] - d[1]) +
                      t[1] / 6.0 * (d[6] + d[2]) * (d[6] + d[2]) + 3.025 * t[2] * (d[7] - d[3]) * (d[7] - d[3]);
                rb1 = std::sqrt(rb1);
                rb2 = Z12 * d[0] * Z12 * d[0] + 2.0 / 3.0 * t[1] * d[2] * d[2];
                rb2 = std::sqrt(rb2);
            } else {
                rb2 = 0.25 * Z12 * Z12 * (d[0] + d[4]) * (d[0] + d[4]) + t[3] * (d[5] - d[1]) * (d[5] - d[1]) +
                      t[4] / 6.0 * (d[6] + d[2]) * (d[6] + d[2]) + 3.025 * t[5] * (d[7] - d[3]) * (d[7] - d[3]);
                rb2 = std::sqrt(rb2);
                rb1 = Z12 * d[4] * Z12 * d[4] + 2.0 / 3.0 * t[4] * d[6] * d[6];
                rb1 = std::sqrt(rb1);
            }
        } else {
            double g1 = t[0] * d[1] + t[1] * d[2] + t[2] * d[3];
            g1 = g1 / (d[0] * d[0]);
            double g2 = t[3] * d[5] + t[4] * d[6] + t[5] * d[7];
            g2 = g2 / (d[4] * d[4]);
            const double G1 = G(g1, p.ibar[a]), G2 = G(g2, p.ibar[b]);
            const double bk1 = p.bkgd_dyn == 1 ? p.rho0[a] * p.Z[a] : rho_ref[a];
            const double bk2 = p.bkgd_dyn == 1 ? p.rho0[b] * p.Z[b] : rho_ref[b];
            rb1 = d[0] / bk1 * G1;
            rb2 = d[4] / bk2 * G2;
        }
        const double F1 = emb(a, rb1), F2 = emb(b, rb2);
        const double Eu = erose(r, a, b);
        double v;
        if (lat == kLatC11) {
            const double pp = p.lattice[a][a] == kLatDia ? phi(r, a, a) : phi(r, b, b);
            v = (3.0 * Eu - F1 - 2.0 * F2 - 5.0 * pp) / Z12;
        } else if (lat == kLatL12) {
            double paa = phi(r, a, a);
            double Z2, arat, scrn;
            if (!z2(p.lattice[a][a], p.cmin[a][a][a], p.cmax[a][a][a], &Z2, &arat, &scrn)) bad = true;
            const double Z1 = z1(p.lattice[a][a]);
            if (scrn > 0.0) {
                double rr = r;
                const double ratio = -Z2 * scrn / Z1;
                for (int n = 0; n < 10; ++n) {
                    rr *= arat;
                    paa += ratio * phi(rr, a, a);  // the ratio, not its power (:3827)
                }
            }
            v = Eu / 3.0 - F1 / 4.0 - F2 / 12.0 - paa;
        } else {
            v = (2.0 * Eu - F1 - F2) / Z12;
        }
        return r == 0.0 ? 0.0 : v;
    }

    // second-neighbour series sum_n ratio^n phi(r arat^n) added to v
    double series(double v, double r, double ratio, double arat, int a, int b) {
        double pw = 1.0, rr = r;
        for (int n = 0; n < 10; ++n) {
            pw *= ratio;
            rr *= arat;
            v += pw * phi(rr, a, b);
        }
        return v;
    }

    std::string build(const mapp_meam_params& in) {
        if (in.nt < 1 || in.nt > kMeamT) return "MEAM takes 1 to 4 types";
        if (in.mix_ref_t == 1)
            return "mixture_ref_t = 1 is not supported: the reference's pair potential then reads an "
                   "uninitialised background density";
        p = in;
        const int nt = p.nt;
        for (int a = 0; a < nt; ++a)
            for (int b = 0; b < nt; ++b)
                if (p.lattice[a][b] < 0 || p.lattice[a][b] > kLatB2) return "a type pair has no lattice (lattce)";
        for (int a = 0; a < nt; ++a)
            if (!(p.ibar[a] == 0 || p.ibar[a] == 1 || p.ibar[a] == 3 || p.ibar[a] == 4 || p.ibar[a] == -5))
                return "ibar error in MEAM";
        for (int a = 0; a < nt; ++a) p.t1[a] += (double)p.augt1 * 0.6 * p.t3[a];
        // alloy_params
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j < nt; ++j) {
                if (p.Ec[i][j] == 0.0) {
                    if (p.lattice[i][j] == kLatL12) p.Ec[i][j] = 0.25 * (3.0 * p.Ec[i][i] + p.Ec[j][j]) - p.delta[i][j];
                    else if (p.lattice[i][j] == kLatC11)
                        p.Ec[i][j] = p.lattice[i][i] == kLatDia ? (2.0 * p.Ec[i][i] + p.Ec[j][j]) / 3.0 - p.delta[i][j]
                                                                : (p.Ec[i][i] + 2.0 * p.Ec[j][j]) / 3.0 - p.delta[i][j];
                    else p.Ec[i][j] = 0.5 * (p.Ec[i][i] + p.Ec[j][j]);
                    p.Ec[j][i] = p.Ec[i][j];
                }
                if (p.alpha[i][j] == 0.0) {
                    p.alpha[i][j] = 0.5 * (p.alpha[i][i] + p.alpha[j][j]);
                    p.alpha[j][i] = p.alpha[i][j];
                }
                if (p.re[i][j] == 0.0) {
                    p.re[i][j] = 0.5 * (p.re[i][i] + p.re[j][j]);
                    p.re[j][i] = p.re[i][j];
                }
            }
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j < nt; ++j)
                for (int k = 0; k < nt; ++k) {
                    const double c = p.cmax[i][j][k];
                    const double e = 0.25 * (c * c) / (c - 1.0);
                    p.ebound[i][j] = p.ebound[i][j] > e ? p.ebound[i][j] : e;
                }
        // compute_reference_density
        for (int t = 0; t < nt; ++t) {
            const double Z = p.Z[t];
            double Gbar = 1.0;
            if (p.ibar[t] > 0) {
                double s[3];
                if (!shape(p.lattice[t][t], s)) bad = true;
                Gbar = G((p.t1[t] * s[0] + p.t2[t] * s[1] + p.t3[t] * s[2]) / (Z * Z), p.ibar[t]);
            }
            double r0 = p.rho0[t] * Z;
            if (p.nn2[t][t] == 1) {
                double Z2, arat, scrn;
                if (!z2(p.lattice[t][t], p.cmin[t][t][t], p.cmax[t][t][t], &Z2, &arat, &scrn)) bad = true;
                r0 += Z2 * (p.rho0[t] * std::exp(-p.beta0[t] * (arat - 1.0))) * scrn;
            }
            rho_ref[t] = r0 * Gbar;
        }
        // compute_pair_meam
        dr = 1.1 * p.rc / (double)kMeamNr;
        dr_inv = 1.0 / dr;
        table.assign((size_t)(nt * (nt + 1) / 2) * kMeamNr * 7, 0.0);
        for (int a = 0; a < nt; ++a)
            for (int b = a; b < nt; ++b) {
                double* T = table.data() + (size_t)comp(a, b) * kMeamNr * 7;
                const int lat = p.lattice[a][b];
                for (int n = 0; n < kMeamNr; ++n) {
                    const double r = (double)n * dr;
                    double v = phi(r, a, b);
                    if (p.nn2[a][b] == 1) {
                        double Z1 = z1(lat), Z2, arat, scrn, scrn2;
                        if (!z2(lat, p.cmin[a][a][b], p.cmax[a][a][b], &Z2, &arat, &scrn)) bad = true;
                        if (lat == kLatB1 || lat == kLatB2 || lat == kLatL12) {
                            const double rarat = r * arat;
                            double paa = phi(rarat, a, a);
                            Z1 = z1(p.lattice[a][a]);
                            if (!z2(p.lattice[a][a], p.cmin[a][a][a], p.cmax[a][a][a], &Z2, &arat, &scrn)) bad = true;
                            if (scrn > 0.0) paa = series(paa, rarat, -Z2 * scrn / Z1, arat, a, a);
                            double pbb = phi(rarat, b, b);
                            Z1 = z1(p.lattice[b][b]);
                            if (!z2(p.lattice[b][b], p.cmin[b][b][b], p.cmax[b][b][b], &Z2, &arat, &scrn)) bad = true;
                            if (scrn > 0.0) pbb = series(pbb, rarat, -Z2 * scrn / Z1, arat, b, b);
                            if (lat == kLatB1 || lat == kLatB2) {
                                Z1 = z1(lat);
                                z2(lat, p.cmin[a][a][b], p.cmax[a][a][b], &Z2, &arat, &scrn);
                                v -= 0.5 * Z2 * scrn / Z1 * paa;
                                z2(lat, p.cmin[b][b][a], p.cmax[b][b][a], &Z2, &arat, &scrn2);
                                v -= 0.5 * Z2 * scrn2 / Z1 * pbb;
                            } else {
                                const double s111 = sijk(1.0, a, a, a), s112 = sijk(1.0, a, a, b);
                                const double s221 = sijk(1.0, b, b, a);
                                const double S11 = s111 * s111 * s112 * s112, S22 = s221 * s221 * s221 * s221;
                                v += -0.75 * S11 * paa - 0.25 * S22 * pbb;
                            }
                        } else {
                            v = series(v, r, -Z2 * scrn / Z1, arat, a, b);
                        }
                    }
                    if (p.zbl[a][b] == 1) {
                        const double as = p.alpha[a][b] * (r / p.re[a][b] - 1.0);
                        if (as <= -3.0) {
                            v = zbl(r, (double)p.ielt[a], (double)p.ielt[b]);
                        } else if (as > -3.0 && as < -1.0) {
                            const double fr = fcut(1.0 + 0.5 * (as + 1.0));
                            v = fr * v + (1.0 - fr) * zbl(r, (double)p.ielt[a], (double)p.ielt[b]);
                        }
                    }
                    T[(size_t)n * 7] = v;
                }
                // 5-point derivatives and Hermite rows (:3374-3407)
                auto V = [&](int n) { return T[(size_t)n * 7]; };
                auto D = [&](int n) -> double& { return T[(size_t)n * 7 + 1]; };
                const int N = kMeamNr;
                D(0) = V(1) - V(0);
                D(1) = 0.5 * (V(2) - V(0));
                D(N - 2) = 0.5 * (V(N - 1) - V(N - 3));
                D(N - 1) = 0.0;
                for (int n = 2; n < N - 2; ++n) D(n) = ((V(n - 2) - V(n + 2)) + 8.0 * (V(n + 1) - V(n - 1))) / 12.0;
                T[(size_t)(N - 1) * 7 + 2] = T[(size_t)(N - 1) * 7 + 3] = 0.0;
                for (int n = 0; n < N - 1; ++n) {
                    T[(size_t)n * 7 + 2] = 3.0 * (V(n + 1) - V(n)) - 2.0 * D(n) - D(n + 1);
                    T[(size_t)n * 7 + 3] = D(n) + D(n + 1) - 2.0 * (V(n + 1) - V(n));
                }
                for (int n = 0; n < N; ++n) {
                    T[(size_t)n * 7 + 4] = D(n) * dr_inv;
                    T[(size_t)n * 7 + 5] = 2.0 * T[(size_t)n * 7 + 2] * dr_inv;
                    T[(size_t)n * 7 + 6] = 3.0 * T[(size_t)n * 7 + 3] * dr_inv;
                }
            }
        if (bad) return "MEAM set-up reached an unsupported lattice or ibar";
        return "";
    }

    MeamDev device(int fix) const {
        MeamDev d{};
        d.nt = p.nt;
        d.ialloy = p.ialloy;
        d.emb_lin_neg = p.emb_lin_neg;
        d.bkgd_dyn = p.bkgd_dyn;
        d.fix = fix;
        d.rc = p.rc;
        d.delr_inv = p.delr_inv;
        d.dr_inv = dr_inv;
        d.gsmooth = p.gsmooth;
        for (int t = 0; t < p.nt; ++t) {
            d.Z[t] = p.Z[t];
            d.A[t] = p.A[t];
            d.rho0[t] = p.rho0[t];
            d.beta[t][0] = p.beta0[t];
            d.beta[t][1] = p.beta1[t];
            d.beta[t][2] = p.beta2[t];
            d.beta[t][3] = p.beta3[t];
            d.t[t][0] = p.t1[t];
            d.t[t][1] = p.t2[t];
            d.t[t][2] = p.t3[t];
            d.rho_ref[t] = rho_ref[t];
            d.Ec[t] = p.Ec[t][t];
            d.re[t] = p.re[t][t];
            d.inv_re[t] = 1.0 / p.re[t][t];
            double s[3];
            shape(p.lattice[t][t], s);
            for (int q = 0; q < 3; ++q) d.shape[t][q] = s[q];
            d.ibar[t] = p.ibar[t];
        }
        for (int i = 0; i < kMeamT; ++i)
            for (int j = 0; j < kMeamT; ++j) {
                d.ebound[i][j] = p.ebound[i][j];
                for (int k = 0; k < kMeamT; ++k) {
                    d.cmin[i][j][k] = p.cmin[i][j][k];
                    d.cmax[i][j][k] = p.cmax[i][j][k];
                }
            }
        return d;
    }
};

}  // namespace meam_host
}  // namespace mappdmd

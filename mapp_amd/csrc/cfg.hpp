// Extended-CFG files of DMD runs (host only).
//
//   reader  ReadCFG (header, set_box, read_atom_dmd, add_atom_read_x),
//           src/read_cfg.cpp:15-178, :228-418, :562-694
//   writer  WriteCFG::write_file_dmd, src/write_cfg.cpp:262-433
//
// A DMD atom line holds entry_count = 3 + 2 nt numbers: reduced position,
// nt Gaussian widths alpha, nt concentrations.  Type blocks are a mass line
// followed by an element-name line.  The writer emits the same layout with
// %.17g instead of the reference's %lf, so a write/read round trip is exact.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tables.hpp"

namespace mappdmd {

struct CfgData {
    int natms = 0, nt = 0;
    double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    std::vector<double> x;  // [natms][3+nt] reduced position in [0,1), alpha
    std::vector<double> c;  // [natms][nt]
    std::vector<double> mass;
    std::vector<std::string> names;
};

namespace cfgdetail {

inline bool parse_index2(const std::string& tok, const char* head, int* i, int* j) {
    // "H0(1,2)" -> (0, 1)
    const size_t n = std::strlen(head);
    if (tok.compare(0, n, head) != 0) return false;
    return std::sscanf(tok.c_str() + n, "(%d,%d)", i, j) == 2;
}

inline bool number(const std::string& s, double* v) {
    char* end = nullptr;
    *v = std::strtod(s.c_str(), &end);
    return end && end != s.c_str();
}

// ReadCFG::set_box (src/read_cfg.cpp:356-418) without eta: H_x = A (H0 . T),
// then the lower-triangular box with the same edge lengths and angles.
inline void set_box(const double H0[3][3], const double T[3][3], double A, double out[3][3]) {
    double Hx[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Hx[i][j] = 0.0;
            for (int k = 0; k < 3; ++k) Hx[i][j] += H0[i][k] * T[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Hx[i][j] *= A;
    double sq[3], Ht[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3];
    for (int i = 0; i < 3; ++i) {
        sq[i] = 0.0;
        for (int j = 0; j < 3; ++j) sq[i] += Hx[i][j] * Hx[i][j];
    }
    Ht[0][0] = std::sqrt(sq[0]);
    for (int i = 0; i < 3; ++i) Ht[1][0] += Hx[0][i] * Hx[1][i];
    Ht[1][0] /= Ht[0][0];
    Ht[1][1] = std::sqrt(sq[1] - Ht[1][0] * Ht[1][0]);
    for (int i = 0; i < 3; ++i) Ht[2][0] += Hx[0][i] * Hx[2][i];
    Ht[2][0] /= Ht[0][0];
    // unit normal of the first two edges
    b[0] = Hx[0][1] * Hx[1][2] - Hx[0][2] * Hx[1][1];
    b[1] = Hx[0][2] * Hx[1][0] - Hx[0][0] * Hx[1][2];
    b[2] = Hx[0][0] * Hx[1][1] - Hx[0][1] * Hx[1][0];
    const double babs = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    for (int i = 0; i < 3; ++i) b[i] /= babs;
    for (int i = 0; i < 3; ++i) Ht[2][2] += Hx[2][i] * b[i];
    Ht[2][1] = std::sqrt(sq[2] - Ht[2][2] * Ht[2][2] - Ht[2][0] * Ht[2][0]);
    std::memcpy(out, Ht, sizeof(Ht));
}

}  // namespace cfgdetail

// Returns "" or an error message.
inline std::string read_cfg_file(const char* path, CfgData& out) {
    using namespace cfgdetail;
    FILE* fp = std::fopen(path, "r");
    if (!fp) return std::string("file ") + path + " not found";
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{fp};
    double H0[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, T[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double A = 1.0;
    int entry_count = 6, ext = 0, tot = -1;
    bool vel = true;
    std::vector<char> buf(1 << 16);
    std::vector<std::string> pending;  // the first line after the header
    bool have_pending = false;
    // header (ReadCFG::read_header, src/read_cfg.cpp:228-350)
    while (std::fgets(buf.data(), (int)buf.size(), fp)) {
        std::vector<std::string> t = split_tokens(buf.data());
        const size_t n = t.size();
        if (n == 0) continue;
        double v;
        int i, j;
        if (n == 1) {
            if (t[0] == ".NO_VELOCITY.") {
                vel = false;
                ext = 1;
                entry_count -= 3;
            } else if (ext) {
                pending = t;
                have_pending = true;
                break;
            } else {
                return "unknown command: " + t[0];
            }
        } else if (n == 3 && t[1] == "=") {
            if (parse_index2(t[0], "Transform", &i, &j) && number(t[2], &v) && i >= 1 && i <= 3 && j >= 1 && j <= 3)
                T[i - 1][j - 1] = v;
            else if (parse_index2(t[0], "eta", &i, &j) && number(t[2], &v)) {
                if (v != 0.0) return "eta(i,j) != 0 is not supported";
            } else if (t[0] == "entry_count" && number(t[2], &v)) {
                entry_count = (int)v;
                ext = 1;
                if (entry_count < 3 + 3 * (vel ? 1 : 0)) return "wrong number of entry_count";
            } else {
                return "unknown command: " + t[0];
            }
        } else if (n == 4 && t[1] == "=") {
            if (parse_index2(t[0], "H0", &i, &j) && number(t[2], &v) && i >= 1 && i <= 3 && j >= 1 && j <= 3)
                H0[i - 1][j - 1] = v;
            else if (t[0] == "R" || t[0].compare(0, 10, "auxiliary[") == 0)
                ;  // velocity scale / column names: not needed here
            else
                return "unknown command: " + t[0];
        } else if (n == 5 && t[0] == "Number" && t[1] == "of" && t[2] == "particles" && number(t[4], &v)) {
            tot = (int)v;
        } else if (n == 5 && t[0].compare(0, 10, "auxiliary[") == 0) {
            ;
        } else if (n == 6 && t[0] == "A" && number(t[2], &v)) {
            A = v;
        } else {
            return "unknown command: " + t[0];
        }
    }
    if (!ext) return "DMD mode can only read extended cfg";
    if ((entry_count - 3) % 2 != 0) return "wrong number of entry for DMD in cfg file";
    const int nt = (entry_count - 3) / 2;
    out.nt = nt;
    set_box(H0, T, A, out.H);
    out.x.clear();
    out.c.clear();
    out.mass.clear();
    out.names.clear();
    // atoms (ReadCFG::read_atom_dmd + add_atom_read_x, src/read_cfg.cpp:562-657)
    int last_type = -1;
    std::vector<std::string> t = pending;
    bool use_pending = have_pending;
    while (true) {
        if (!use_pending) {
            if (!std::fgets(buf.data(), (int)buf.size(), fp)) break;
            t = split_tokens(buf.data());
        }
        use_pending = false;
        if (t.empty()) continue;
        if (t.size() == 1) {
            double m;
            if (!number(t[0], &m)) return "error cfg: " + t[0];
            if (m <= 0.0) return "mass cannot be zero";
            if (!std::fgets(buf.data(), (int)buf.size(), fp)) return "missing element name";
            std::vector<std::string> nm = split_tokens(buf.data());
            if (nm.size() != 1) return "error cfg: element line";
            int k = -1;
            for (size_t q = 0; q < out.names.size(); ++q)
                if (out.names[q] == nm[0]) k = (int)q;
            if (k < 0) {
                out.names.push_back(nm[0]);
                out.mass.push_back(m);
                k = (int)out.names.size() - 1;
            }
            last_type = k;
            if (last_type >= nt) return "exceeded the number of types in DMD";
        } else if ((int)t.size() == entry_count) {
            double b[64];
            if (entry_count > 64) return "entry_count too large";
            for (int k = 0; k < entry_count; ++k)
                if (!number(t[k], &b[k])) return "bad number: " + t[k];
            for (int d = 0; d < 3; ++d) {
                while (b[d] >= 1) b[d]--;
                while (b[d] < 0) b[d]++;
            }
            for (int k = 0; k < 3 + nt; ++k) out.x.push_back(b[k]);
            for (int k = 0; k < nt; ++k) out.c.push_back(b[3 + nt + k]);
        } else {
            return "wrong extended format";
        }
    }
    out.natms = (int)(out.x.size() / (3 + nt));
    if (tot >= 0 && tot != out.natms)
        return "the number of atoms dont match: " + std::to_string(tot) + " " + std::to_string(out.natms);
    return "";
}

// WriteCFG::write_file_dmd layout; x holds reduced positions and alpha.
inline std::string write_cfg_file(const char* path, const CfgData& d) {
    FILE* fp = std::fopen(path, "w");
    if (!fp) return std::string("cannot open ") + path;
    const int nt = d.nt;
    std::fprintf(fp, "Number of particles = %d\n", d.natms);
    std::fprintf(fp, "A = 1 Angstrom (basic length-scale)\n");
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) std::fprintf(fp, "H0(%d,%d) = %.17g A\n", i + 1, j + 1, d.H[i][j]);
    std::fprintf(fp, ".NO_VELOCITY.\n");
    std::fprintf(fp, "entry_count = %d\n", 3 + 2 * nt);
    int icomp = 0;
    if (nt == 1) std::fprintf(fp, "auxiliary[%d] = alpha [reduced unit]\n", icomp++);
    else
        for (int k = 0; k < nt; ++k) std::fprintf(fp, "auxiliary[%d] = alpha_%d [reduced unit]\n", icomp++, k);
    if (nt == 1) std::fprintf(fp, "auxiliary[%d] = c [reduced unit]\n", icomp++);
    else
        for (int k = 0; k < nt; ++k) std::fprintf(fp, "auxiliary[%d] = c_%d [reduced unit]\n", icomp++, k);
    for (size_t k = 0; k < d.names.size(); ++k) std::fprintf(fp, "%.17g \n%s \n", d.mass[k], d.names[k].c_str());
    for (int i = 0; i < d.natms; ++i) {
        for (int k = 0; k < 3 + nt; ++k) std::fprintf(fp, "%.17g ", d.x[(size_t)i * (3 + nt) + k]);
        for (int k = 0; k < nt; ++k) std::fprintf(fp, "%.17g ", d.c[(size_t)i * nt + k]);
        std::fprintf(fp, "\n");
    }
    const bool ok = std::fclose(fp) == 0;
    return ok ? "" : std::string("write failed: ") + path;
}

}  // namespace mappdmd

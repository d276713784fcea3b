// Host-side table construction of the DMD force field:
//   cubic Hermite rows from a 5-point derivative   ForceField_DMD_I::interpolate   src/ff_dmd_I.cpp:1760-1790
//   quartic-corrected rows for F(rho)              ForceField_DMD_I::interpolate_m src/ff_dmd_I.cpp:1794-1836
//   Finnis-Sinclair file reader                    ForceField_DMD_II::set_fs       src/ff_dmd_II.cpp:1253-1480
#pragma once
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace mappdmd {

// rows of `stride` doubles; column 0 holds the sampled function on entry.
// quartic == false: c1..c3 cubic, c4..c6 derivative coefficients / delta.
// quartic == true:  c4 quartic correction, c5..c8 derivative coefficients / delta.
inline void spline_rows(int n, double delta, double* s, int stride, bool quartic) {
    auto S = [&](int i, int k) -> double& { return s[(size_t)i * stride + k]; };
    S(0, 1) = S(1, 0) - S(0, 0);
    S(1, 1) = 0.5 * (S(2, 0) - S(0, 0));
    S(n - 2, 1) = 0.5 * (S(n - 1, 0) - S(n - 3, 0));
    S(n - 1, 1) = S(n - 1, 0) - S(n - 2, 0);
    for (int i = 2; i < n - 2; ++i)
        S(i, 1) = ((S(i - 2, 0) - S(i + 2, 0)) + 8.0 * (S(i + 1, 0) - S(i - 1, 0))) / 12.0;
    for (int i = 0; i < n - 1; ++i) {
        S(i, 2) = 3.0 * (S(i + 1, 0) - S(i, 0)) - 2.0 * S(i, 1) - S(i + 1, 1);
        S(i, 3) = S(i, 1) + S(i + 1, 1) - 2.0 * (S(i + 1, 0) - S(i, 0));
    }
    S(n - 1, 2) = 0.0;
    S(n - 1, 3) = 0.0;
    if (!quartic) {
        for (int i = 0; i < n; ++i) {
            S(i, 4) = S(i, 1) / delta;
            S(i, 5) = 2.0 * S(i, 2) / delta;
            S(i, 6) = 3.0 * S(i, 3) / delta;
        }
        return;
    }
    for (int i = 0; i < n - 1; ++i) S(i, 4) = (S(i + 1, 2) - S(i, 2)) / 6.0 - 0.5 * S(i, 3);
    S(n - 1, 4) = 0.0;
    for (int i = 0; i < n; ++i) {
        S(i, 5) = S(i, 1) / delta;
        S(i, 6) = 2.0 * S(i, 2) / delta;
        S(i, 7) = 3.0 * S(i, 3) / delta;
        S(i, 8) = 4.0 * S(i, 4) / delta;
    }
}

// Tokens of one line up to a '#' (MAPP::parse_line, src/MAPP.cpp:468-530).
inline std::vector<std::string> split_tokens(const char* line) {
    std::vector<std::string> out;
    const char* p = line;
    while (*p && *p != '#') {
        if (std::isspace((unsigned char)*p)) {
            ++p;
            continue;
        }
        const char* q = p;
        while (*q && !std::isspace((unsigned char)*q) && *q != '#') ++q;
        out.emplace_back(p, q - p);
        p = q;
    }
    return out;
}

struct FsData {
    int nrho = 0, nr = 0;
    double drho = 0, dr = 0;
    std::vector<double> F, rho, rphi;  // type order: F[a][nrho], rho[a][b][nr], rphi[i][j<=i][nr]
};

// Reads an FS-format potential: three comment lines, the element line
// ("N name_1 .. name_N"), "nrho drho nr dr cut", one element header line, then
// a flat token stream of N*nrho F values, N*N*nr rho values and N(N+1)/2*nr
// r*phi values.  `names` gives the element of each type.  Returns "" or an error.
inline std::string read_fs_file(const char* path, const std::vector<std::string>& names, FsData& out) {
    FILE* fp = std::fopen(path, "r");
    if (!fp) return std::string("file ") + path + " not found";
    std::vector<char> buf(1 << 16);
    auto next_line = [&](std::string& line) -> bool {
        line.clear();
        if (!std::fgets(buf.data(), (int)buf.size(), fp)) return false;
        line = buf.data();
        while (!line.empty() && line.back() != '\n') {  // long line
            if (!std::fgets(buf.data(), (int)buf.size(), fp)) break;
            line += buf.data();
        }
        return true;
    };
    std::string line;
    for (int i = 0; i < 4; ++i)
        if (!next_line(line)) {
            std::fclose(fp);
            return "eam potential file ended immaturely";
        }
    std::vector<std::string> el = split_tokens(line.c_str());
    const int nt = (int)names.size();
    if (el.size() < 2 || (int)el.size() - 1 < nt) {
        std::fclose(fp);
        return "wrong format in eam potential file";
    }
    const int tot = (int)el.size() - 1;
    std::vector<int> file_of_type(nt, -1);
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < tot; ++k)
            if (el[k + 1] == names[t]) {
                file_of_type[t] = k;
                break;
            }
    for (int t = 0; t < nt; ++t)
        if (file_of_type[t] < 0) {
            std::fclose(fp);
            return "atom type not found in eam file";
        }
    if (!next_line(line)) {
        std::fclose(fp);
        return "eam potential file ended immaturely";
    }
    std::vector<std::string> hdr = split_tokens(line.c_str());
    if (hdr.size() != 5) {
        std::fclose(fp);
        return "wrong format in eam potential file";
    }
    out.nrho = std::atoi(hdr[0].c_str());
    out.drho = std::atof(hdr[1].c_str());
    out.nr = std::atoi(hdr[2].c_str());
    out.dr = std::atof(hdr[3].c_str());
    if (!next_line(line)) {
        std::fclose(fp);
        return "eam potential file ended immaturely";
    }
    const long need = (long)tot * out.nrho + (long)tot * tot * out.nr + (long)tot * (tot + 1) / 2 * out.nr;
    std::vector<double> v;
    v.reserve(need);
    while ((long)v.size() < need) {
        if (!next_line(line)) {
            std::fclose(fp);
            return "eam potential file ended immaturely";
        }
        for (auto& tok : split_tokens(line.c_str())) v.push_back(std::atof(tok.c_str()));
        if ((long)v.size() > need) {
            std::fclose(fp);
            return "eam potential file ended immaturely";
        }
    }
    std::fclose(fp);
    const int nrho = out.nrho, nr = out.nr;
    out.F.assign((size_t)nt * nrho, 0.0);
    out.rho.assign((size_t)nt * nt * nr, 0.0);
    out.rphi.assign((size_t)nt * (nt + 1) / 2 * nr, 0.0);
    // file layout: element k: F (nrho), rho_{k,l} for every file element l (nr each)
    auto elem_base = [&](int k) { return (long)k * (nrho + (long)tot * nr); };
    for (int a = 0; a < nt; ++a) {
        long b0 = elem_base(file_of_type[a]);
        for (int i = 0; i < nrho; ++i) out.F[(size_t)a * nrho + i] = v[b0 + i];
        for (int b = 0; b < nt; ++b) {
            long r0 = b0 + nrho + (long)file_of_type[b] * nr;
            for (int i = 0; i < nr; ++i) out.rho[((size_t)a * nt + b) * nr + i] = v[r0 + i];
        }
    }
    long phi0 = (long)tot * (nrho + (long)tot * nr);
    auto phi_base = [&](int k, int l) {  // file pair (k, l), l <= k
        return phi0 + ((long)k * (k + 1) / 2 + l) * nr;
    };
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j <= i; ++j) {
            int k = file_of_type[i], l = file_of_type[j];
            if (l > k) std::swap(k, l);
            long b0 = phi_base(k, l);
            size_t dst = ((size_t)i * (i + 1) / 2 + j) * nr;
            for (int r = 0; r < nr; ++r) out.rphi[dst + r] = v[b0 + r];
        }
    return "";
}

// Finnis-Sinclair parameter file of ForceField_fs::coef (src/ff_fs.cpp:113-335):
// the element line, one A per element, nt*nt "t1 t2" lines (row element,
// column element), nt(nt+1)/2 "cut_rho cut_phi k1 k2 k3" lines (i <= j), in
// the file's element order; blank lines and '#' comments are skipped.  names
// gives the type order of the outputs (A[nt], t1/t2[nt*nt], per-pair arrays
// in COMP order).  Returns "" or an error message.
struct FsParams {
    std::vector<double> A, t1, t2, cut_rho, cut_phi, k1, k2, k3;
};
inline std::string read_fs_params(const char* path, const std::vector<std::string>& names, FsParams& out) {
    FILE* fp = std::fopen(path, "r");
    if (!fp) return std::string("FS file ") + path + " not found";
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{fp};
    std::vector<char> buf(1 << 16);
    auto next = [&](std::vector<std::string>& t) -> bool {
        while (std::fgets(buf.data(), (int)buf.size(), fp)) {
            t = split_tokens(buf.data());
            if (!t.empty()) return true;
        }
        return false;
    };
    const int nt = (int)names.size();
    std::vector<std::string> t;
    if (!next(t)) return "FS file: missing element line";
    if ((int)t.size() != nt) return "number of types represented and forcefield do not match";
    std::vector<int> type(nt, -1);
    for (int i = 0; i < nt; ++i) {
        for (int k = 0; k < nt; ++k)
            if (names[k] == t[i]) type[i] = k;
        if (type[i] < 0) return "Atom type " + t[i] + " not found";
        for (int k = 0; k < i; ++k)
            if (type[k] == type[i]) return "Atom type " + t[i] + " duplicated in the force field file";
    }
    const int nc = nt * (nt + 1) / 2;
    out.A.assign(nt, 0.0);
    out.t1.assign((size_t)nt * nt, 0.0);
    out.t2.assign((size_t)nt * nt, 0.0);
    for (auto* v : {&out.cut_rho, &out.cut_phi, &out.k1, &out.k2, &out.k3}) v->assign(nc, 0.0);
    auto num = [](const std::string& s, double* v) {
        char* end = nullptr;
        *v = std::strtod(s.c_str(), &end);
        return end && end != s.c_str();
    };
    for (int i = 0; i < nt; ++i) {
        if (!next(t) || t.size() != 1 || !num(t[0], &out.A[type[i]])) return "Wrong format in the force filed file";
    }
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j) {
            double a, b;
            if (!next(t) || t.size() != 2 || !num(t[0], &a) || !num(t[1], &b))
                return "Wrong format in the force filed file";
            out.t1[(size_t)type[i] * nt + type[j]] = a;
            out.t2[(size_t)type[i] * nt + type[j]] = b;
        }
    for (int i = 0; i < nt; ++i)
        for (int j = i; j < nt; ++j) {
            double v[5];
            if (!next(t) || t.size() != 5) return "Wrong format in the force filed file";
            for (int k = 0; k < 5; ++k)
                if (!num(t[k], &v[k])) return "Wrong format in the force filed file";
            const int c = md_comp(type[i], type[j]);
            out.cut_rho[c] = v[0];
            out.cut_phi[c] = v[1];
            out.k1[c] = v[2];
            out.k2[c] = v[3];
            out.k3[c] = v[4];
        }
    return "";
}

}  // namespace mappdmd

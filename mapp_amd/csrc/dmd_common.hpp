// Shared definitions of the MI355X DMD force field: kernel parameter blocks,
// device buffers and the context behind the C ABI (include/mapp_dmd.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mapp_dmd.h"

namespace mappdmd {

constexpr int kMaxTypes = 4;
constexpr int kMaxPoints = 14;
constexpr double kPiInSq = 0.564189583547756286948079451561;  // 1/sqrt(pi)

// Scalars every force-field kernel needs (passed by value).
template <typename T>
struct FFParams {
    int natms, ntot, nr, nrho, no_i;
    T dr_inv, drho_inv, rho_max, rc, cut_sq, cut_sq_mod, alpha_min, alpha_max, inv_amin, inv_amax;
    T kbT, beta;
    T xw;  // energy weight of cross-type pair terms: 1 (reference) or 2 (consistent energy)
    T xi[kMaxPoints], w0[kMaxPoints], w1[kMaxPoints], w2[kMaxPoints];
    T c0[kMaxTypes];
};

constexpr int kMaxComp = kMaxTypes * (kMaxTypes + 1) / 2;

// COMP(i, j) of the reference (src/init.h:114): the unordered type pair
__host__ __device__ constexpr int md_comp(int i, int j) { return i < j ? (j + 1) * j / 2 + i : (i + 1) * i / 2 + j; }

// Geometry of the single-rank periodic domain (reduced coordinates in [0,1)).
// pairwise: the list cutoff is cut_sk_sq_tp[COMP(type_i, type_j)] (MD-mode
// potentials with a cutoff per type pair), else cut_sk_sq.
struct BoxParams {
    double H[3][3], B[3][3];
    double cut_ph_s[3], s_ph_lo[3], s_ph_hi[3];
    double bin_size[3];
    int grid[3], denom[3], tot_bin;
    double cut_sk_sq;
    int pairwise;
    int full;  // every j != i (MEAM's full list) instead of j > i
    double cut_sk_sq_tp[kMaxComp];
};

// Coefficients of the MD-mode pair potentials per type pair (COMP order):
// Lennard-Jones (src/ff_lj.cpp) and Finnis-Sinclair (src/ff_fs.cpp).
struct MdParams {
    int nt;
    int chk[kMaxComp];         // the pair has coefficients
    double cut_sq[kMaxComp];   // force cutoff squared
    double eps[kMaxComp], sig[kMaxComp], offset[kMaxComp];  // LJ
    double A[kMaxTypes], t1[kMaxTypes][kMaxTypes], t2[kMaxTypes][kMaxTypes];  // FS
    double cut_rho[kMaxComp], cut_phi[kMaxComp], k1[kMaxComp], k2[kMaxComp], k3[kMaxComp];
    double cut_sq_rho[kMaxComp], cut_sq_phi[kMaxComp];  // FSM: the cutoffs squared, tested as rsq < cut^2
    double lj_cut[kMaxComp];   // LJ cutoff of the coef call (list cutoff of pairs without coefficients)
    int set[kMaxComp];         // the pair was given coefficients
};

#define HIP_OK(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) throw ::mappdmd::HipError(e_, #expr, __LINE__); \
    } while (0)

struct HipError {
    hipError_t code;
    const char* what;
    int line;
    HipError(hipError_t c, const char* w, int l) : code(c), what(w), line(l) {}
};

struct ApiError {
    int code;
    std::string msg;
    ApiError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// Minimal owning device buffer (capacity grows, contents not preserved unless asked).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    void reserve(size_t n, bool keep = false, hipStream_t s = nullptr) {
        if (n <= cap) return;
        size_t nc = n + n / 8 + 64;
        T* q = nullptr;
        HIP_OK(hipMalloc(&q, nc * sizeof(T)));
        if (keep && p && cap) HIP_OK(hipMemcpyAsync(q, p, cap * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (keep) HIP_OK(hipStreamSynchronize(s));
        if (p) HIP_OK(hipFree(p));
        p = q;
        cap = nc;
    }
};

}  // namespace mappdmd

// Geometry kernels: periodic ghosts in reduced coordinates, cell binning and
// the half neighbor list, restating the single-rank reference algorithms
//   Atoms::set_ph / ph_setup / ph_xchng   src/atoms.cpp:249-372, :464-497
//   Atoms::update (x and plain vectors)   src/atoms.cpp:1088-1220
//   Atoms::update_0 / x2s / s2x           src/atoms.cpp:1780-1941
//   Neighbor::create_bin_list / find_bin_no / bin_atoms_s / s2bin / create_list
//                                         src/neighbor.cpp:137-525
// The library is compiled with -ffp-contract=off, so the s->x transform and
// the squared distances are the same separately rounded binary64 operations
// as the reference's; the list is therefore bit-identical, including order.
#pragma once
#include "dmd_common.hpp"

namespace mappdmd {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ double& comp3(double4& p, int d) {
    return d == 0 ? p.x : (d == 1 ? p.y : p.z);
}
__device__ __forceinline__ double comp3c(const double4& p, int d) {
    return d == 0 ? p.x : (d == 1 ? p.y : p.z);
}

// ---- total of n non-negative int counts in 64 bits (the int scan below wraps
// past 2^31; integer sums do not depend on the order of the additions) ----
constexpr int kCountThreads = 256;
__global__ __launch_bounds__(kCountThreads) void k_count_total(const int* __restrict__ in, long n,
                                                        unsigned long long* __restrict__ total) {
    unsigned long long s = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += (unsigned)in[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(total, s);
}

// ---- exclusive scan over int (tile scan + recursive scan of tile sums) ----
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const int* __restrict__ in, int* __restrict__ out,
                                                            int* __restrict__ sums, long n) {
    __shared__ int sh[kScanThreads];
    const int tid = threadIdx.x;
    const long base = (long)blockIdx.x * kScanTile + (long)tid * kScanItems;
    int v[kScanItems];
    int tsum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        long idx = base + k;
        v[k] = idx < n ? in[idx] : 0;
        tsum += v[k];
    }
    sh[tid] = tsum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int t = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    int excl = sh[tid] - tsum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        long idx = base + k;
        if (idx < n) out[idx] = excl;
        excl += v[k];
    }
    if (tid == kScanThreads - 1) sums[blockIdx.x] = sh[tid];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int* __restrict__ out, const int* __restrict__ add, long n) {
    const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    const int a = add[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        long idx = base + k;
        if (idx < n) out[idx] += a;
    }
}

// ---- ghosts ----------------------------------------------------------------
// flag[i-lo] = 1 when atom i in [lo,hi) lies within the ghost shell of the
// low (dir 0: s <= s_ph_lo) or high (dir 1: s_ph_hi <= s) face of dimension idim
__global__ void k_ghost_flag(const double4* __restrict__ pos, int lo, int hi, int idim, int dir, double thr,
                             int* __restrict__ flag) {
    int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    double s = comp3c(pos[i], idim);
    flag[i - lo] = dir == 0 ? (s <= thr ? 1 : 0) : (thr <= s ? 1 : 0);
}

// copies flagged atoms to ghost rows gbase + offs[i-lo], shifted by +-1 in s
template <int NT>
__global__ void k_ghost_emit(double4* __restrict__ pos, double* __restrict__ alpha, double* __restrict__ conc,
                             int lo, int hi, int idim, int dir, const int* __restrict__ flag,
                             const int* __restrict__ offs, int gbase, int natms, int* __restrict__ ghost_src,
                             int* __restrict__ ghost_owner) {
    int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi || !flag[i - lo]) return;
    int g = gbase + offs[i - lo];
    double4 p = pos[i];
    if (dir == 0) comp3(p, idim) += 1.0;
    else comp3(p, idim) -= 1.0;
    pos[g] = p;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        alpha[(long)g * NT + t] = alpha[(long)i * NT + t];
        conc[(long)g * NT + t] = conc[(long)i * NT + t];
    }
    ghost_src[g - natms] = i;
    ghost_owner[g - natms] = i < natms ? i : ghost_owner[i - natms];
}

// Atoms::update of x between rebuilds: ghost = source +- H[idim][0..idim]
template <int NT>
__global__ void k_ghost_update_x(double4* __restrict__ pos, double* __restrict__ alpha, int natms, int gbeg,
                                 int gend, const int* __restrict__ ghost_src, int idim, int dir, double h0,
                                 double h1, double h2) {
    int g = gbeg + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gend) return;
    int src = ghost_src[g];
    double4 p = pos[src];
    if (dir == 0) {
        p.x += h0;
        if (idim >= 1) p.y += h1;
        if (idim >= 2) p.z += h2;
    } else {
        p.x -= h0;
        if (idim >= 1) p.y -= h1;
        if (idim >= 2) p.z -= h2;
    }
    pos[natms + g] = p;
#pragma unroll
    for (int t = 0; t < NT; ++t) alpha[(long)(natms + g) * NT + t] = alpha[(long)src * NT + t];
}

// forward copy of a per-atom vector to its ghosts (Atoms::update of E, dE, n ...)
template <typename T>
__global__ void k_ghost_copy(T* __restrict__ v, int d, int natms, int nghost, const int* __restrict__ owner) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nghost) return;
    int o = owner[g];
    for (int k = 0; k < d; ++k) v[(long)(natms + g) * d + k] = v[(long)o * d + k];
}

// ---- box transforms --------------------------------------------------------
// Atoms::s2x: x_j = s_j H_jj + sum_{k>j} H_kj s_k (lower-triangular H)
__global__ void k_s2x(double4* __restrict__ pos, int n, BoxParams bp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 s = pos[i];
    double x0 = s.x * bp.H[0][0];
    x0 += bp.H[1][0] * s.y;
    x0 += bp.H[2][0] * s.z;
    double x1 = s.y * bp.H[1][1];
    x1 += bp.H[2][1] * s.z;
    double x2 = s.z * bp.H[2][2];
    pos[i] = make_double4(x0, x1, x2, 0.0);
}

// Atoms::x2s with the wrap into [0,1)
__global__ void k_x2s(double4* __restrict__ pos, int n, BoxParams bp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 x = pos[i];
    double s0 = x.x * bp.B[0][0];
    s0 += bp.B[1][0] * x.y;
    s0 += bp.B[2][0] * x.z;
    double s1 = x.y * bp.B[1][1];
    s1 += bp.B[2][1] * x.z;
    double s2 = x.z * bp.B[2][2];
    double s[3] = {s0, s1, s2};
    for (int j = 0; j < 3; ++j) {
        // bounded loops: a site more than 64 boxes away is an input error
        for (int it = 0; it < 64 && s[j] < 0.0; ++it) s[j]++;
        for (int it = 0; it < 64 && s[j] >= 1.0; ++it) s[j]--;
    }
    pos[i] = make_double4(s[0], s[1], s[2], 0.0);
}

// ---- cell binning (reduced coordinates) -----------------------------------
__global__ void k_s2bin(const double4* __restrict__ pos, int n, BoxParams bp, int* __restrict__ bin_of,
                        int* __restrict__ bin_cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 s = pos[i];
    int no = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int t = (int)((comp3c(s, d) + bp.cut_ph_s[d] - 0.0) / bp.bin_size[d]);
        t = t < 0 ? 0 : (t > bp.grid[d] - 1 ? bp.grid[d] - 1 : t);  // in range for valid input
        no += t * bp.denom[d];
    }
    bin_of[i] = no;
    atomicAdd(&bin_cnt[no], 1);
}

__global__ void k_bin_fill(const int* __restrict__ bin_of, int n, const int* __restrict__ bin_start,
                           int* __restrict__ bin_cur, int* __restrict__ bin_atoms) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = bin_of[i];
    int k = atomicAdd(&bin_cur[b], 1);
    bin_atoms[bin_start[b] + k] = i;
}

// Ascending order inside every segment (atom indices inside each bin, as the
// reference's linked bins built by pushing atoms from the highest index down,
// src/neighbor.cpp:428-434; pair indices inside each reverse-list segment).
// One thread per value: value x of segment s = key[x] goes to
// out[start[s] + r], r = the number of the segment's values below x.  The
// values of a segment are distinct, so the ranks are a permutation of the
// segment and every write stays inside it; the segment is read through the
// caches instead of being shifted entry by entry in memory.  With pos_of, also
// pos_of[x] = the position x lands at (the reverse position of each pair).
__global__ void k_rank_place(const int* __restrict__ v, long n, const int* __restrict__ key,
                             const int* __restrict__ start, int* __restrict__ out, int* __restrict__ pos_of) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int x = v[q];
    const int s = key[x];
    const int lo = start[s], hi = start[s + 1];
    int r = 0;
    for (int t = lo; t < hi; ++t) r += v[t] < x ? 1 : 0;
    out[lo + r] = x;
    if (pos_of) pos_of[x] = lo + r;
}

// ---- half neighbor list ----------------------------------------------------
// stencil order of find_bin_no: offset of dimension 0 outermost, 2 innermost;
// inside a bin ascending index; j > i (every j != i with bp.full, the full list
// of src/neighbor.cpp:240-284); rsq = ((d0*d0) + d1*d1) + d2*d2 < cut.
// One wavefront per site: lane s < 27 takes stencil bin s (s = 9(a+1) + 3(b+1)
// + c+1 for offsets a, b, c), a prefix sum of the bin sizes concatenates the 27
// bins in stencil order, and 64 consecutive candidates are tested per round;
// ballot ranks write the accepted ones in candidate order, which is the list
// order of a single thread walking the stencil.
template <bool FILL>
__global__ __launch_bounds__(256) void k_half_list(const double4* __restrict__ pos, int natms,
                                                   const int* __restrict__ bin_of,
                                                   const int* __restrict__ bin_start,
                                                   const int* __restrict__ bin_atoms, BoxParams bp,
                                                   const double* __restrict__ type,
                                                   int* __restrict__ cnt_or_off, int* __restrict__ nl) {
    const int lane = threadIdx.x & 63;
    const long site = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (site >= natms) return;  // the whole wavefront
    const int i = (int)site;
    const int ti = bp.pairwise ? (int)type[i] : 0;
    int tmp = bin_of[i];
    int loc[3];
    for (int d = 2; d > -1; --d) {
        loc[d] = tmp / bp.denom[d];
        tmp -= loc[d] * bp.denom[d];
    }
    int bstart = 0, bsize = 0;
    if (lane < 27) {
        const int t0 = loc[0] + lane / 9 - 1, t1 = loc[1] + (lane / 3) % 3 - 1, t2 = loc[2] + lane % 3 - 1;
        if (t0 >= 0 && t0 < bp.grid[0] && t1 >= 0 && t1 < bp.grid[1] && t2 >= 0 && t2 < bp.grid[2]) {
            const int no = t0 * bp.denom[0] + t1 * bp.denom[1] + t2 * bp.denom[2];
            bstart = bin_start[no];
            bsize = bin_start[no + 1] - bstart;
        }
    }
    int incl = bsize;  // inclusive prefix over the stencil bins
    for (int d = 1; d < 32; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const int excl = incl - bsize;
    const int total = __shfl(incl, 26, 64);
    const double4 xi = pos[i];
    int w = FILL ? cnt_or_off[i] : 0;
    for (int base = 0; base < total; base += 64) {
        const int q = base + lane;
        // stencil bin of candidate q: the number of bins whose inclusive prefix is <= q
        int s = 0;
#pragma unroll
        for (int step = 16; step >= 1; step >>= 1) {
            const int probe = s + step - 1;
            const int v = __shfl(incl, probe < 27 ? probe : 26, 64);
            if (probe < 27 && v <= q) s += step;
        }
        s = s < 26 ? s : 26;
        const int bs = __shfl(bstart, s, 64), ex = __shfl(excl, s, 64);
        bool take = false;
        int j = 0;
        if (q < total) {
            j = bin_atoms[bs + (q - ex)];
            if (bp.full ? j != i : j > i) {
                const double4 xj = pos[j];
                double d0 = xi.x - xj.x, d1 = xi.y - xj.y, d2 = xi.z - xj.z;
                double rsq = 0.0;
                rsq += d0 * d0;
                rsq += d1 * d1;
                rsq += d2 * d2;
                const double cut = bp.pairwise ? bp.cut_sk_sq_tp[md_comp(ti, (int)type[j])] : bp.cut_sk_sq;
                take = rsq < cut;
            }
        }
        const unsigned long long m = __ballot(take);
        if (FILL && take)
            nl[w + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = j;
        w += __popcll(m);
    }
    if (!FILL && lane == 0) cnt_or_off[i] = w;
}

// ---- reverse CSR: for owned j, the list positions p whose partner is j ----
// kRevG lanes per site over its list entries; the fill order inside a segment
// is arbitrary (atomics) and is put in ascending pair order by k_rank_place
constexpr int kRevG = 16;
__global__ void k_rev_count(const int* __restrict__ off, const int* __restrict__ nl, int natms,
                            int* __restrict__ rcnt) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(t / kRevG), l = (int)(t % kRevG);
    if (i >= natms) return;
    for (int p = off[i] + l; p < off[i + 1]; p += kRevG) {
        const int j = nl[p];
        if (j < natms) atomicAdd(&rcnt[j], 1);
    }
}

__global__ void k_rev_fill(const int* __restrict__ off, const int* __restrict__ nl, int natms,
                           const int* __restrict__ rstart, int* __restrict__ rcur, int* __restrict__ rev) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(t / kRevG), l = (int)(t % kRevG);
    if (i >= natms) return;
    for (int p = off[i] + l; p < off[i + 1]; p += kRevG) {
        const int j = nl[p];
        if (j < natms) {
            const int k = atomicAdd(&rcur[j], 1);
            rev[rstart[j] + k] = p;
        }
    }
}

// sub-list of the half list with rsq < rsq_crd (create_2nd_neigh_lst); one
// wavefront per site, 64 list entries per round, ballot ranks keep the list order
template <bool FILL>
__global__ __launch_bounds__(256) void k_list2(const double4* __restrict__ pos, int natms,
                                               const int* __restrict__ off, const int* __restrict__ nl,
                                               double rsq_crd, int* __restrict__ cnt_or_off,
                                               int* __restrict__ nl2) {
    const int lane = threadIdx.x & 63;
    const long site = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (site >= natms) return;  // the whole wavefront
    const int i = (int)site;
    const double4 xi = pos[i];
    int w = FILL ? cnt_or_off[i] : 0;
    const int p1 = off[i + 1];
    for (int p0 = off[i]; p0 < p1; p0 += 64) {
        const int p = p0 + lane;
        bool take = false;
        int j = 0;
        if (p < p1) {
            j = nl[p];
            const double4 xj = pos[j];
            double d0 = xi.x - xj.x, d1 = xi.y - xj.y, d2 = xi.z - xj.z;
            double rsq = d0 * d0 + d1 * d1 + d2 * d2;
            take = rsq < rsq_crd;
        }
        const unsigned long long m = __ballot(take);
        if (FILL && take)
            nl2[w + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = j;
        w += __popcll(m);
    }
    if (!FILL && lane == 0) cnt_or_off[i] = w;
}

// coordination: own 2nd-list size + number of lower-index owners listing i
__global__ void k_crd(const int* __restrict__ off2, const int* __restrict__ roff2, int natms,
                      int* __restrict__ crd) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms) return;
    crd[i] = (off2[i + 1] - off2[i]) + (roff2[i + 1] - roff2[i]);
}

// Atoms::update_0 skin test: any owned site moved more than skin since the build
__global__ void k_skin_check(const double4* __restrict__ pos, const double4* __restrict__ x0, int natms,
                             double sq_skin, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms) return;
    double4 a = pos[i], b = x0[i];
    double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
    double sq = 0.0;
    sq += d0 * d0;
    sq += d1 * d1;
    sq += d2 * d2;
    if (sq > sq_skin) atomicOr(flag, 1);
}

}  // namespace mappdmd

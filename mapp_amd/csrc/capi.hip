// C ABI of the MI355X DMD force field (declarations and reference citations in
// include/mapp_dmd.h).  Host orchestration: tables, parameters, the GPU build
// of ghosts and lists, and the kernel sequences of force_calc / energy_calc /
// c_d_calc / calc_g.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <type_traits>

#include "context.hpp"
#include "geometry.hpp"
#include "force.hpp"
#include "mobility.hpp"
#include "eam.hpp"
#include "meam.hpp"
#include "surrogate.hpp"
#include "lbfgs.hpp"
#include "cfg.hpp"
#include "bdf.hpp"
#include <chrono>
#include <numeric>
#include <random>
#include "gh_table.h"
#include "tables.hpp"

using namespace mappdmd;

mapp_dmd::~mapp_dmd() {
    for (auto* b : scan_levels) delete b;
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

inline int nblk(long n, int b = kBlock) { return (int)((n + b - 1) / b); }

template <typename F>
int guarded(mapp_dmd* ctx, F&& fn) {
    if (!ctx) return MAPP_DMD_E_ARG;
    try {
        int rc = fn();
        return rc;
    } catch (const HipError& e) {
        ctx->err = std::string("HIP error ") + hipGetErrorString(e.code) + " at " + e.what + " (line " +
                   std::to_string(e.line) + ")";
        return MAPP_DMD_E_HIP;
    } catch (const ApiError& e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return MAPP_DMD_E_HIP;
    } catch (const std::exception& e) {
        ctx->err = std::string("unexpected host error: ") + e.what();
        return MAPP_DMD_E_INTERNAL;
    } catch (...) {
        ctx->err = "unexpected host error";
        return MAPP_DMD_E_INTERNAL;
    }
}

// the same for the entry points without a context: no exception crosses the C ABI,
// the message goes to the caller's err buffer
void put_err(char* err, int cap, const std::string& m);
template <typename F>
int guarded_noctx(char* err, int err_cap, F&& fn) {
    try {
        return fn();
    } catch (const std::bad_alloc&) {
        put_err(err, err_cap, "host allocation failed");
        return MAPP_DMD_E_INTERNAL;
    } catch (const std::exception& e) {
        put_err(err, err_cap, std::string("unexpected host error: ") + e.what());
        return MAPP_DMD_E_INTERNAL;
    } catch (...) {
        put_err(err, err_cap, "unexpected host error");
        return MAPP_DMD_E_INTERNAL;
    }
}

void launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw HipError(e, what, 0);
}

// ascending order inside every segment of v[0..n) (segments start[0..nseg]),
// key[value] = the segment that holds the value; pos_of (optional) receives the
// final position of every value
void sort_segments(mapp_dmd* ctx, const int* start, const int* key, int* v, long n, int* pos_of = nullptr) {
    if (n <= 0) return;
    ctx->d_sort_tmp.reserve(n);
    hipLaunchKernelGGL(k_rank_place, dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, v, n, key, start,
                       ctx->d_sort_tmp.p, pos_of);
    launch_check("k_rank_place");
    HIP_OK(hipMemcpyAsync(v, ctx->d_sort_tmp.p, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->stream));
}

// exclusive scan of n ints; returns the total (synchronizes the stream)
long exclusive_scan(mapp_dmd* ctx, const int* in, int* out, long n, int depth = 0) {
    if (n <= 0) return 0;
    const int nb = (int)((n + kScanTile - 1) / kScanTile);
    while ((int)ctx->scan_levels.size() < 2 * (depth + 1)) ctx->scan_levels.push_back(new DevBuf<int>());
    DevBuf<int>& sums = *ctx->scan_levels[2 * depth];
    DevBuf<int>& sums_x = *ctx->scan_levels[2 * depth + 1];
    sums.reserve(nb);
    sums_x.reserve(nb);
    hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(kScanThreads), 0, ctx->stream, in, out, sums.p, n);
    launch_check("k_scan_tile");
    if (nb > 1) {
        exclusive_scan(ctx, sums.p, sums_x.p, nb, depth + 1);
        hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, ctx->stream, out, sums_x.p, n);
        launch_check("k_scan_add");
    }
    int last[2];
    HIP_OK(hipMemcpyAsync(&last[0], out + n - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(&last[1], in + n - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return (long)last[0] + last[1];
}

// CSR offsets (n+1) from counts (n); the offsets are int, so a total of 2^31
// entries or more is refused before any caller sizes its entry array by it
long counts_to_offsets(mapp_dmd* ctx, const int* cnt, DevBuf<int>& off, long n) {
    off.reserve(n + 1);
    unsigned long long total = 0;
    if (n > 0) {
        ctx->d_total.reserve(1);
        HIP_OK(hipMemsetAsync(ctx->d_total.p, 0, sizeof(unsigned long long), ctx->stream));
        hipLaunchKernelGGL(k_count_total, dim3((int)std::min<long>(nblk(n, kCountThreads), 1024)), dim3(kCountThreads),
                           0, ctx->stream,
                           cnt, n, ctx->d_total.p);
        launch_check("k_count_total");
        HIP_OK(hipMemcpyAsync(&total, ctx->d_total.p, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
    }
    long tot = exclusive_scan(ctx, cnt, off.p, n);  // synchronizes the stream: total is in
    if (total >= (1ull << 31))
        throw ApiError(MAPP_DMD_E_RANGE, "a list or table needs " + std::to_string(total) +
                                             " entries; its int offsets hold fewer than 2^31");
    int t = (int)tot;
    HIP_OK(hipMemcpyAsync(off.p + n, &t, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return tot;
}

// partials per first-stage block of reduce_partials
constexpr long kReduceChunk = 4096;
// doubles of stage memory reduce_partials needs for nparts partials
inline long reduce_stage_doubles(long nparts) { return ((nparts + kReduceChunk - 1) / kReduceChunk + 1) * 8; }

// out[0..nv) = sum over nparts partials (stride doubles each, starting at offset),
// in two fixed-shape stages through stage[reduce_stage_doubles(nparts)]
void reduce_partials_in(mapp_dmd* ctx, const double* part, long nparts, int nv, int stride, int offset, double* stage,
                        double* out) {
    const long nb1 = (nparts + kReduceChunk - 1) / kReduceChunk;
    hipLaunchKernelGGL(k_reduce_stage, dim3((int)nb1), dim3(kBlock), 0, ctx->stream, part, nparts, nv, stride, offset,
                       kReduceChunk, stage, 8);
    hipLaunchKernelGGL(k_reduce_stage, dim3(1), dim3(kBlock), 0, ctx->stream, stage, nb1, nv, 8, 0, nb1, out, 1);
    launch_check("k_reduce_stage");
}

// several reduce_partials (parts[s], nparts[s], nv[s] <= 8 values, stride[s] -> out[out_off[s]..]) in two launches
void reduce_partials_segs(mapp_dmd* ctx, int nseg, const double* const* parts, const long* nparts, const int* nv,
                          const int* stride, const int* out_off, double* out) {
    ReduceSegs R{};
    R.nseg = nseg;
    int tot = 0;
    for (int k = 0; k < nseg; ++k) {
        R.part[k] = parts[k];
        R.nparts[k] = nparts[k];
        R.nv[k] = nv[k];
        R.stride[k] = stride[k];
        R.out_off[k] = out_off[k];
        R.first[k] = tot;
        R.nb[k] = (int)((nparts[k] + kReduceChunk - 1) / kReduceChunk);
        tot += R.nb[k];
    }
    ctx->d_stage.reserve((size_t)(tot + 1) * 8);
    if (tot > 0)
        hipLaunchKernelGGL(k_reduce_segs1, dim3(tot), dim3(kBlock), 0, ctx->stream, R, kReduceChunk, ctx->d_stage.p);
    hipLaunchKernelGGL(k_reduce_segs2, dim3(nseg), dim3(kBlock), 0, ctx->stream, R, ctx->d_stage.p, out);
    launch_check("k_reduce_segs");
}

// the same with the context's stage buffer
void reduce_partials(mapp_dmd* ctx, const double* part, long nparts, int nv, int stride, int offset, double* out) {
    ctx->d_stage.reserve((size_t)reduce_stage_doubles(nparts));
    reduce_partials_in(ctx, part, nparts, nv, stride, offset, ctx->d_stage.p, out);
}

void require(bool cond, int code, const char* msg) {
    if (!cond) throw ApiError(code, msg);
}

// Atoms::set_ph(TYPE0) for one rank (s_lo = 0, s_hi = 1), src/atoms.cpp:249-280
void set_ph(mapp_dmd* ctx, double cut) {
    BoxParams& bp = ctx->bp;
    const double tot_cut_ph = cut + ctx->skin;
    for (int i = 0; i < 3; ++i) {
        double tmp = 0.0;
        for (int j = i; j < 3; ++j) tmp += bp.B[j][i] * bp.B[j][i];
        bp.cut_ph_s[i] = tot_cut_ph * std::sqrt(tmp);
    }
    for (int i = 0; i < 3; ++i) {
        bp.s_ph_hi[i] = 1.0 - bp.cut_ph_s[i];
        bp.s_ph_lo[i] = 0.0 + bp.cut_ph_s[i];
    }
    // Neighbor::create_bin_list, src/neighbor.cpp:311-326
    bp.tot_bin = 1;
    for (int i = 0; i < 3; ++i) {
        bp.bin_size[i] = bp.cut_ph_s[i];
        bp.grid[i] = (int)((2.0 * bp.cut_ph_s[i] + 1.0 - 0.0) / bp.bin_size[i]) + 1;
        bp.tot_bin *= bp.grid[i];
    }
    for (int i = 0; i < 3; ++i) {
        int no = 1;
        for (int j = 0; j < i; ++j) no *= bp.grid[j];
        bp.denom[i] = no;
    }
}

template <int NT>
void build_ghosts(mapp_dmd* ctx) {
    BoxParams& bp = ctx->bp;
    const int natms = ctx->natms;
    ctx->nghost = 0;
    ctx->swap_off.assign(1, 0);
    ctx->swap_dim.clear();
    ctx->swap_dir.clear();
    int lo_dim = 0, hi_dim = natms;
    for (int idim = 0; idim < 3; ++idim) {
        const int comm_need = (int)(bp.cut_ph_s[idim] / (1.0 - 0.0)) + 1;
        for (int dir = 0; dir < 2; ++dir) {
            int lo = lo_dim, hi = hi_dim;
            for (int k = 0; k < comm_need; ++k) {
                const int tmp_r = natms + ctx->nghost;
                const int n = hi - lo;
                int cnt = 0;
                if (n > 0) {
                    ctx->d_flag.reserve(n);
                    ctx->d_offs.reserve(n);
                    double thr = dir == 0 ? bp.s_ph_lo[idim] : bp.s_ph_hi[idim];
                    hipLaunchKernelGGL(k_ghost_flag, dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, ctx->d_pos.p, lo,
                                       hi, idim, dir, thr, ctx->d_flag.p);
                    launch_check("k_ghost_flag");
                    cnt = (int)exclusive_scan(ctx, ctx->d_flag.p, ctx->d_offs.p, n);
                }
                const int need = natms + ctx->nghost + cnt;
                ctx->d_pos.reserve(need, true, ctx->stream);
                ctx->d_alpha.reserve((size_t)need * NT, true, ctx->stream);
                ctx->d_conc.reserve((size_t)need * NT, true, ctx->stream);
                ctx->d_ghost_src.reserve(ctx->nghost + cnt + 1, true, ctx->stream);
                ctx->d_ghost_owner.reserve(ctx->nghost + cnt + 1, true, ctx->stream);
                if (cnt > 0)
                    hipLaunchKernelGGL((k_ghost_emit<NT>), dim3(nblk(n)), dim3(kBlock), 0, ctx->stream,
                                       ctx->d_pos.p, ctx->d_alpha.p, ctx->d_conc.p, lo, hi, idim, dir,
                                       ctx->d_flag.p, ctx->d_offs.p, natms + ctx->nghost, natms,
                                       ctx->d_ghost_src.p, ctx->d_ghost_owner.p);
                launch_check("k_ghost_emit");
                ctx->nghost += cnt;
                ctx->swap_dim.push_back(idim);
                ctx->swap_dir.push_back(dir);
                ctx->swap_off.push_back(ctx->nghost);
                hi = natms + ctx->nghost;
                lo = tmp_r;
            }
        }
        lo_dim = 0;
        hi_dim = natms + ctx->nghost;
    }
    HIP_OK(hipStreamSynchronize(ctx->stream));
}

template <int NT>
void finish_lists(mapp_dmd* ctx);
void finish_meam(mapp_dmd* ctx);

__global__ void k_quirk_identity(int ntot, int nt, int* __restrict__ qidx) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)ntot * nt) return;
    qidx[e] = (int)(e / nt + e % nt);
}

// the same element for the sites of cell k of a batch, numbered within the cell
// (owned 0..n-1, ghosts n..n+g-1) and mapped back to batch indices
__global__ void k_quirk_cell(int n, int g, int own_base, int gh_dst, int nt, int* __restrict__ qidx) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n + g) return;
    const int gs = s < n ? own_base + s : gh_dst + (s - n);
    for (int t = 0; t < nt; ++t) {
        const int L = s + t, a = L / nt, tt = L % nt;
        const int ga = a < n ? own_base + a : gh_dst + (a - n);
        qidx[(size_t)gs * nt + t] = ga * nt + tt;
    }
}

// ghosts + bins + s->x + half list + reverse CSR; positions hold s on entry
template <int NT>
void rebuild(mapp_dmd* ctx) {
    BoxParams& bp = ctx->bp;
    const int natms = ctx->natms;
    build_ghosts<NT>(ctx);
    const int ntot = natms + ctx->nghost;
    hipStream_t st = ctx->stream;
    // bins
    ctx->d_bin_of.reserve(ntot);
    ctx->d_bin_cnt.reserve(bp.tot_bin + 1);
    ctx->d_bin_atoms.reserve(ntot);
    HIP_OK(hipMemsetAsync(ctx->d_bin_cnt.p, 0, sizeof(int) * (bp.tot_bin + 1), st));
    hipLaunchKernelGGL(k_s2bin, dim3(nblk(ntot)), dim3(kBlock), 0, st, ctx->d_pos.p, ntot, bp, ctx->d_bin_of.p,
                       ctx->d_bin_cnt.p);
    launch_check("k_s2bin");
    counts_to_offsets(ctx, ctx->d_bin_cnt.p, ctx->d_bin_start, bp.tot_bin);
    HIP_OK(hipMemsetAsync(ctx->d_bin_cnt.p, 0, sizeof(int) * (bp.tot_bin + 1), st));
    hipLaunchKernelGGL(k_bin_fill, dim3(nblk(ntot)), dim3(kBlock), 0, st, ctx->d_bin_of.p, ntot,
                       ctx->d_bin_start.p, ctx->d_bin_cnt.p, ctx->d_bin_atoms.p);
    launch_check("k_bin_fill");
    sort_segments(ctx, ctx->d_bin_start.p, ctx->d_bin_of.p, ctx->d_bin_atoms.p, ntot);
    // Neighbor::bin_atoms_s ends with s2x of every atom
    hipLaunchKernelGGL(k_s2x, dim3(nblk(ntot)), dim3(kBlock), 0, st, ctx->d_pos.p, ntot, bp);
    launch_check("k_s2x");
    // half list
    ctx->d_cnt.reserve(natms + 1);
    hipLaunchKernelGGL((k_half_list<false>), dim3(nblk((long)natms * 64)), dim3(kBlock), 0, st, ctx->d_pos.p, natms,
                       ctx->d_bin_of.p, ctx->d_bin_start.p, ctx->d_bin_atoms.p, bp, ctx->d_conc.p, ctx->d_cnt.p,
                       (int*)nullptr);
    launch_check("k_half_list count");
    ctx->npairs = counts_to_offsets(ctx, ctx->d_cnt.p, ctx->d_off, natms);
    ctx->d_nl.reserve(ctx->npairs + 1);
    hipLaunchKernelGGL((k_half_list<true>), dim3(nblk((long)natms * 64)), dim3(kBlock), 0, st, ctx->d_pos.p, natms,
                       ctx->d_bin_of.p, ctx->d_bin_start.p, ctx->d_bin_atoms.p, bp, ctx->d_conc.p, ctx->d_off.p,
                       ctx->d_nl.p);
    launch_check("k_half_list fill");
    if (ctx->variant == MAPP_DMD_VARIANT_MEAM) finish_meam(ctx);
    else finish_lists<NT>(ctx);
}

// reverse CSR, pair owners, reverse positions and per-site / per-pair buffers
// for the half list in d_off / d_nl
template <int NT>
void finish_lists(mapp_dmd* ctx) {
    const int natms = ctx->natms;
    const int ntot = natms + ctx->nghost;
    hipStream_t st = ctx->stream;
    ctx->d_cnt.reserve(natms + 1);
    // reverse CSR (partner j owned)
    HIP_OK(hipMemsetAsync(ctx->d_cnt.p, 0, sizeof(int) * (natms + 1), st));
    hipLaunchKernelGGL(k_rev_count, dim3(nblk((long)natms * kRevG)), dim3(kBlock), 0, st, ctx->d_off.p, ctx->d_nl.p, natms,
                       ctx->d_cnt.p);
    launch_check("k_rev_count");
    long nrev = counts_to_offsets(ctx, ctx->d_cnt.p, ctx->d_roff, natms);
    ctx->d_rev.reserve(nrev + 1);
    HIP_OK(hipMemsetAsync(ctx->d_cnt.p, 0, sizeof(int) * (natms + 1), st));
    hipLaunchKernelGGL(k_rev_fill, dim3(nblk((long)natms * kRevG)), dim3(kBlock), 0, st, ctx->d_off.p, ctx->d_nl.p, natms,
                       ctx->d_roff.p, ctx->d_cnt.p, ctx->d_rev.p);
    launch_check("k_rev_fill");
    // reverse positions of the pairs with an owned partner (rpos) come with the ordering
    ctx->d_rpos.reserve(ctx->npairs + 1);
    sort_segments(ctx, ctx->d_roff.p, ctx->d_nl.p, ctx->d_rev.p, nrev, ctx->d_rpos.p);
    ctx->d_nlrp.reserve(ctx->npairs + 1);
    if (ctx->npairs > 0) {
        hipLaunchKernelGGL(k_pack_nlrp, dim3(nblk(ctx->npairs)), dim3(kBlock), 0, st, ctx->d_nl.p, ctx->d_rpos.p,
                           ctx->npairs, ctx->d_nlrp.p);
        launch_check("k_pack_nlrp");
    }
    ctx->d_pair_i.reserve(ctx->npairs + 1);
    hipLaunchKernelGGL(k_pair_owner, dim3(nblk(natms)), dim3(kBlock), 0, st, ctx->d_off.p, natms, ctx->d_pair_i.p);
    launch_check("k_pair_owner");
    ctx->nrev = nrev;
    // element read by the reference's dE[jatm + jtype] quirk: jatm + jtype (one cell)
    ctx->d_qidx.reserve((size_t)ntot * NT);
    hipLaunchKernelGGL(k_quirk_identity, dim3(nblk((long)ntot * NT)), dim3(kBlock), 0, st, ntot, NT, ctx->d_qidx.p);
    launch_check("k_quirk_identity");
    // per-atom vectors and pair buffers
    const size_t nv = (size_t)ntot * NT;
    for (auto* b : {&ctx->d_E, &ctx->d_dE, &ctx->d_ddE, &ctx->d_n, &ctx->d_s, &ctx->d_t, &ctx->d_v})
        b->reserve(nv);
    ctx->d_cd.reserve((size_t)natms * NT);
    ctx->d_a.reserve((size_t)natms * NT);
    ctx->d_g.reserve((size_t)natms * NT);
    ctx->d_fa.reserve((size_t)natms * NT);
    ctx->d_Eown.reserve((size_t)natms * NT);
    ctx->d_f.reserve((size_t)natms * (3 + NT));
    ctx->d_fown.reserve((size_t)natms * (3 + NT));
    ctx->d_pairE.reserve((size_t)(nrev + 1) * NT);
    ctx->d_deriv.reserve((size_t)(ctx->npairs + 1) * DerivRec<NT>::kLen);
    ctx->d_pairF.reserve((size_t)(nrev + 1) * (3 + NT));
    // every pair starts outside the quadrature cut with zero partner shares; pass 1
    // writes the shares of the pairs inside and clears those of a pair that leaves
    HIP_OK(hipMemsetAsync(ctx->d_pairE.p, 0, sizeof(double) * (size_t)(nrev + 1) * NT, st));
    HIP_OK(hipMemsetAsync(ctx->d_pairF.p, 0, sizeof(double) * (size_t)(nrev + 1) * (3 + NT), st));
    ctx->d_crd.reserve(ntot);
    ctx->d_x0.reserve(natms);
    HIP_OK(hipMemcpyAsync(ctx->d_x0.p, ctx->d_pos.p, sizeof(double4) * natms, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    ctx->have_list2 = false;
    ctx->cache_valid = false;
}

// columns of the per-site alpha / c arrays: nt for DMD, one (the type) for EAM
inline bool is_md(const mapp_dmd* ctx) { return ctx->variant >= MAPP_DMD_VARIANT_EAM; }
inline int acols(const mapp_dmd* ctx) { return is_md(ctx) ? 1 : ctx->nt; }

// f(std::integral_constant<int, NT>) for the number of DMD types nt = 1..kMaxTypes
template <typename F>
decltype(auto) with_nt(int nt, F&& f) {
    switch (nt) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    default: return f(std::integral_constant<int, 4>{});
    }
}

void rebuild_any(mapp_dmd* ctx) {
    with_nt(acols(ctx), [&](auto c) { rebuild<decltype(c)::value>(ctx); });
}

template <int NT>
void ghost_update_x(mapp_dmd* ctx) {
    for (size_t sw = 0; sw + 1 < ctx->swap_off.size(); ++sw) {
        int g0 = ctx->swap_off[sw], g1 = ctx->swap_off[sw + 1];
        if (g1 <= g0) continue;
        int idim = ctx->swap_dim[sw], dir = ctx->swap_dir[sw];
        hipLaunchKernelGGL((k_ghost_update_x<NT>), dim3(nblk(g1 - g0)), dim3(kBlock), 0, ctx->stream,
                           ctx->d_pos.p, ctx->d_alpha.p, ctx->natms, g0, g1, ctx->d_ghost_src.p, idim, dir,
                           ctx->bp.H[idim][0], ctx->bp.H[idim][1], ctx->bp.H[idim][2]);
        launch_check("k_ghost_update_x");
    }
}

void ghost_copy(mapp_dmd* ctx, double* v, int d) {
    if (ctx->nghost <= 0) return;
    hipLaunchKernelGGL((k_ghost_copy<double>), dim3(nblk(ctx->nghost)), dim3(kBlock), 0, ctx->stream, v, d,
                       ctx->natms, ctx->nghost, ctx->d_ghost_owner.p);
    launch_check("k_ghost_copy");
}

// AoS (3+nt) owned rows <-> positions / alpha
// NA = alpha columns carried in x (nt for DMD, 0 for EAM); NT = columns of the alpha array
template <int NT, int NA = NT>
__global__ void k_scatter_x(const double* __restrict__ x, int natms, double4* __restrict__ pos,
                            double* __restrict__ alpha) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms) return;
    const double* r = x + (size_t)i * (3 + NA);
    pos[i] = make_double4(r[0], r[1], r[2], 0.0);
#pragma unroll
    for (int t = 0; t < NA; ++t) alpha[(size_t)i * NT + t] = r[3 + t];
}
template <int NT, int NA = NT>
__global__ void k_gather_x(double* __restrict__ x, int n, const double4* __restrict__ pos,
                           const double* __restrict__ alpha) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* r = x + (size_t)i * (3 + NA);
    double4 p = pos[i];
    r[0] = p.x;
    r[1] = p.y;
    r[2] = p.z;
#pragma unroll
    for (int t = 0; t < NA; ++t) r[3 + t] = alpha[(size_t)i * NT + t];
}

template <int NT, int NA = NT>
void scatter_x(mapp_dmd* ctx, const double* dev_aos) {
    hipLaunchKernelGGL((k_scatter_x<NT, NA>), dim3(nblk(ctx->natms)), dim3(kBlock), 0, ctx->stream, dev_aos,
                       ctx->natms, ctx->d_pos.p, ctx->d_alpha.p);
    launch_check("k_scatter_x");
}

// x columns of the public layout: 3 + nt (DMD) or 3 (EAM)
inline int xcols(const mapp_dmd* ctx) { return is_md(ctx) ? 3 : 3 + ctx->nt; }

void gather_x_any(mapp_dmd* ctx, double* dst, int n, const double4* pos, const double* alpha) {
    if (is_md(ctx))
        hipLaunchKernelGGL((k_gather_x<1, 0>), dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, dst, n, pos, alpha);
    else
        with_nt(ctx->nt, [&](auto c) {
            hipLaunchKernelGGL((k_gather_x<decltype(c)::value>), dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, dst, n,
                               pos, alpha);
        });
    launch_check("k_gather_x");
}

// ------------------------------------------------------------------ MD-mode potentials
template <int MODEL>
void md_passes(mapp_dmd* ctx, bool energy_only, int nb, const PairBufs& L, const EamBufs& T, double2* drho,
               double* part1, double* part2) {
    hipStream_t st = ctx->stream;
    if (energy_only)
        hipLaunchKernelGGL((k_eam_pass1<MODEL, true>), dim3(nb), dim3(kBlock), 0, st, ctx->P, T, ctx->md, ctx->d_pos.p,
                           ctx->d_conc.p, L, ctx->d_rpos.p, ctx->d_Eown.p, ctx->d_fown.p, ctx->d_pairE.p,
                           ctx->d_pairF.p, drho, part1);
    else
        hipLaunchKernelGGL((k_eam_pass1<MODEL, false>), dim3(nb), dim3(kBlock), 0, st, ctx->P, T, ctx->md,
                           ctx->d_pos.p, ctx->d_conc.p, L, ctx->d_rpos.p, ctx->d_Eown.p, ctx->d_fown.p,
                           ctx->d_pairE.p, ctx->d_pairF.p, drho, part1);
    launch_check("k_eam_pass1");
    if constexpr (MODEL != kModelLj) {
        if (energy_only)
            hipLaunchKernelGGL((k_eam_embed<MODEL, true>), dim3(nb), dim3(kBlock), 0, st, ctx->P, T, ctx->md,
                               ctx->d_conc.p, L, ctx->d_Eown.p, ctx->d_pairE.p, ctx->d_E.p, ctx->d_dE.p, part2);
        else
            hipLaunchKernelGGL((k_eam_embed<MODEL, false>), dim3(nb), dim3(kBlock), 0, st, ctx->P, T, ctx->md,
                               ctx->d_conc.p, L, ctx->d_Eown.p, ctx->d_pairE.p, ctx->d_E.p, ctx->d_dE.p, part2);
        launch_check("k_eam_embed");
    } else {
        HIP_OK(hipMemsetAsync(part2, 0, sizeof(double) * nb, st));  // no embedding term
    }
}

double run_meam(mapp_dmd* ctx, bool energy_only, int st_clc, double* en_st);

double run_eam(mapp_dmd* ctx, bool energy_only, int st_clc, double* en_st) {
    if (ctx->variant == MAPP_DMD_VARIANT_MEAM) return run_meam(ctx, energy_only, st_clc, en_st);
    hipStream_t st = ctx->stream;
    const int natms = ctx->natms, nb = nblk((long)natms * kEamG);  // 16-lane site groups
    PairBufs L{ctx->d_off.p, ctx->d_nl.p, ctx->d_roff.p, ctx->d_rev.p};
    EamBufs T{ctx->d_rho7.p, ctx->d_phi7.p, ctx->d_F7.p, ctx->nt};
    ctx->d_part.reserve((size_t)nb * 14 + 16);
    ctx->d_out.reserve(16);
    ctx->d_drho.reserve(2 * ((size_t)ctx->npairs + 1));  // one double2 per pair
    ctx->d_pairF2.reserve((size_t)(ctx->nrev + 1) * 3);
    double* part1 = ctx->d_part.p;
    double* part2 = part1 + (size_t)nb * 7;
    double* part3 = part2 + (size_t)nb;
    double2* drho = reinterpret_cast<double2*>(ctx->d_drho.p);
    const bool lj = ctx->variant == MAPP_DMD_VARIANT_LJ;
    if (ctx->variant == MAPP_DMD_VARIANT_EAM) md_passes<kModelEam>(ctx, energy_only, nb, L, T, drho, part1, part2);
    else if (lj) md_passes<kModelLj>(ctx, energy_only, nb, L, T, drho, part1, part2);
    else if (ctx->variant == MAPP_DMD_VARIANT_FS) md_passes<kModelFs>(ctx, energy_only, nb, L, T, drho, part1, part2);
    else md_passes<kModelFsm>(ctx, energy_only, nb, L, T, drho, part1, part2);
    double* out = ctx->d_out.p;
    reduce_partials(ctx, part1, nb, 7, 7, 0, out);
    reduce_partials(ctx, part2, nb, 1, 1, 0, out + 7);
    if (!energy_only) {
        if (!lj) {
            ghost_copy(ctx, ctx->d_dE.p, 1);
            hipLaunchKernelGGL(k_eam_pass2, dim3(nb), dim3(kBlock), 0, st, ctx->P, ctx->d_pos.p, L, ctx->d_rpos.p,
                               drho, ctx->d_dE.p, ctx->d_fown.p, ctx->d_pairF2.p, part3);
            launch_check("k_eam_pass2");
        } else {
            HIP_OK(hipMemsetAsync(part3, 0, sizeof(double) * nb * 6, st));
        }
        hipLaunchKernelGGL(k_eam_gather, dim3(nb), dim3(kBlock), 0, st, natms, L, ctx->d_fown.p, ctx->d_pairF.p,
                           lj ? (const double*)nullptr : ctx->d_pairF2.p, ctx->d_f.p);
        launch_check("k_eam_gather");
        reduce_partials(ctx, part3, nb, 6, 6, 0, out + 8);
    }
    double h[14];
    HIP_OK(hipMemcpyAsync(h, out, sizeof(double) * (energy_only ? 8 : 14), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const double energy = h[0] + h[7];
    if (en_st) {
        en_st[0] = energy;
        if (st_clc)
            for (int k = 0; k < 6; ++k) en_st[1 + k] = h[1 + k] + h[8 + k];
    }
    return energy;
}

// ------------------------------------------------------------------ MEAM
// buffers of the full list: per-entry screening and dU/dS, per-site records
void finish_meam(mapp_dmd* ctx) {
    const int natms = ctx->natms, ntot = natms + ctx->nghost;
    hipStream_t st = ctx->stream;
    ctx->d_iflag.reserve(4);
    HIP_OK(hipMemsetAsync(ctx->d_iflag.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_meam_maxlist, dim3(nblk(natms)), dim3(kBlock), 0, st, ctx->d_off.p, natms, ctx->d_iflag.p);
    launch_check("k_meam_maxlist");
    int mx = 0;
    HIP_OK(hipMemcpyAsync(&mx, ctx->d_iflag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    require(mx <= kMeamMaxNb, MAPP_DMD_E_RANGE, "MEAM: a site has more list entries than the kernels stage (192)");
    ctx->d_mpairs.reserve((size_t)ctx->npairs + 1);
    ctx->d_mrec.reserve((size_t)ntot * kMeamRec);
    ctx->d_mactive.reserve((size_t)natms + 1);
    ctx->d_mact.reserve((size_t)ctx->npairs + 1);
    ctx->d_f.reserve((size_t)natms * 3);
}

// force_calc (src/ff_meam.cpp:699-2142) or energy_calc (:2148-2589)
double run_meam(mapp_dmd* ctx, bool energy_only, int st_clc, double* en_st) {
    hipStream_t st = ctx->stream;
    const int natms = ctx->natms;
    const int nb = (natms + kMeamWaves - 1) / kMeamWaves;
    const MeamDev M = ctx->meam.device(ctx->meam_fix);
    ctx->d_part.reserve((size_t)nb * 14 + 16);
    ctx->d_out.reserve(16);
    double* part1 = ctx->d_part.p;
    double* part2 = part1 + nb;
    double* part3 = part2 + (size_t)nb * 7;
    double* out = ctx->d_out.p;
    if (energy_only)
        hipLaunchKernelGGL((k_meam_density<true>), dim3(nb), dim3(kBlock), 0, st, M, ctx->d_pos.p, ctx->d_conc.p,
                           ctx->d_off.p, ctx->d_nl.p, ctx->d_mtab.p, natms, ctx->d_mpairs.p, ctx->d_mrec.p, part1);
    else
        hipLaunchKernelGGL((k_meam_density<false>), dim3(nb), dim3(kBlock), 0, st, M, ctx->d_pos.p, ctx->d_conc.p,
                           ctx->d_off.p, ctx->d_nl.p, ctx->d_mtab.p, natms, ctx->d_mpairs.p, ctx->d_mrec.p, part1);
    launch_check("k_meam_density");
    reduce_partials(ctx, part1, nb, 1, 1, 0, out);
    if (!energy_only) {
        ghost_copy(ctx, ctx->d_mrec.p, kMeamRec);
        hipLaunchKernelGGL(k_meam_force, dim3(nb), dim3(kBlock), 0, st, M, ctx->d_pos.p, ctx->d_conc.p, ctx->d_off.p,
                           ctx->d_nl.p, ctx->d_mtab.p, natms, ctx->d_mpairs.p, ctx->d_mrec.p, ctx->d_f.p,
                           ctx->d_mactive.p, ctx->d_mact.p, part2);
        launch_check("k_meam_force");
        hipLaunchKernelGGL(k_meam_kforce, dim3(nb), dim3(kBlock), 0, st, M, ctx->d_pos.p, ctx->d_conc.p, ctx->d_off.p,
                           ctx->d_nl.p, natms, ctx->d_mpairs.p, ctx->d_mactive.p, ctx->d_mact.p, ctx->d_f.p, part3);
        launch_check("k_meam_kforce");
        reduce_partials(ctx, part2, nb, 7, 7, 0, out + 1);
        reduce_partials(ctx, part3, nb, 6, 6, 0, out + 8);
    }
    double h[14];
    HIP_OK(hipMemcpyAsync(h, out, sizeof(double) * (energy_only ? 1 : 14), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const double energy = energy_only ? h[0] : h[0] + h[1];
    if (en_st) {
        en_st[0] = energy;
        if (st_clc)
            for (int k = 0; k < 6; ++k) en_st[1 + k] = h[2 + k] + h[8 + k];
    }
    return energy;
}

// ------------------------------------------------------------ force kernels
template <int NT>
void pack_atoms(mapp_dmd* ctx) {
    const long n = (long)(ctx->natms + ctx->nghost) * NT;
    ctx->d_arec.reserve(n);
    hipLaunchKernelGGL(k_pack_atoms, dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, ctx->d_alpha.p, ctx->d_conc.p, n,
                       ctx->d_arec.p);
    launch_check("k_pack_atoms");
}

// blocks of a force-pass launch: one wave per site unless the shape caps the waves
inline int shape_blocks(const mapp_dmd::Shape& sh, int natms) {
    const long w = sh.waves > 0 ? std::min<long>(sh.waves, natms) : natms;
    return (int)std::max<long>(1, (w + sh.wpb - 1) / sh.wpb);
}

// the DMD force shares inside d_pairF: nrev + 1 slots of 4 doubles, then nrev + 1 of nt - 1
template <int NT>
ForceShares force_shares(mapp_dmd* ctx) {
    return ForceShares{ctx->d_pairF.p, ctx->d_pairF.p + (size_t)(ctx->nrev + 1) * 4};
}

// pass 1 with the schedule's lanes per (pair, type pair)
template <int NT, typename Q, int NQ, bool ENERGY_ONLY, bool FULL>
void launch_pass1(mapp_dmd* ctx, dim3 grid, const Q* quad, const FFParams<Q>& PQ, const PairBufs& L, Q* deriv,
                  double* part) {
    const dim3 blk(64 * ctx->sh1.wpb);
    const size_t lds = (size_t)ctx->sh1.wpb * Pass1Stage<NT>::kBytes;  // one stage per wavefront
    hipStream_t st = ctx->stream;
    const CutState C{ctx->d_nlrp.p, force_shares<NT>(ctx)};
    if (ctx->sh1.parts == 1)
        hipLaunchKernelGGL((k_pass1<NT, Q, NQ, ENERGY_ONLY, FULL, 1>), grid, blk, lds, st, ctx->P, PQ, ctx->d_pos.p,
                           ctx->d_arec.p, L, ctx->d_nlrp.p, quad, ctx->d_rho7.p, ctx->d_phi7.p, ctx->d_Eown.p,
                           ctx->d_pairE.p, deriv, part, C);
    else if (ctx->sh1.parts == 2)
        hipLaunchKernelGGL((k_pass1<NT, Q, NQ, ENERGY_ONLY, FULL, 2>), grid, blk, lds, st, ctx->P, PQ, ctx->d_pos.p,
                           ctx->d_arec.p, L, ctx->d_nlrp.p, quad, ctx->d_rho7.p, ctx->d_phi7.p, ctx->d_Eown.p,
                           ctx->d_pairE.p, deriv, part, C);
    else
        hipLaunchKernelGGL((k_pass1<NT, Q, NQ, ENERGY_ONLY, FULL, 4>), grid, blk, lds, st, ctx->P, PQ, ctx->d_pos.p,
                           ctx->d_arec.p, L, ctx->d_nlrp.p, quad, ctx->d_rho7.p, ctx->d_phi7.p, ctx->d_Eown.p,
                           ctx->d_pairE.p, deriv, part, C);
    launch_check("k_pass1");
}

template <int NT, typename Q, int NQ>
void run_force(mapp_dmd* ctx, const Q* quad, const FFParams<Q>& PQ, int st_clc, double* en_st, bool sur = false) {
    hipStream_t st = ctx->stream;
    const int natms = ctx->natms;
    const int nb = nblk((long)natms * kGatherG);             // force gather
    const int nbe = nblk((long)natms * kEmbedG);             // embedding
    const int nbw = shape_blocks(ctx->sh2, natms);           // pass 2
    // the surrogate's pass 1: one block of kSurWaves wavefronts per CU, persistent
    const int nb1 = sur ? std::min(natms, kSurBlocks) : shape_blocks(ctx->sh1, natms);
    PairBufs L{ctx->d_off.p, ctx->d_nl.p, ctx->d_roff.p, ctx->d_rev.p};
    // partials: one per wavefront in pass 1, the embedding and pass 2 (one per block
    // in the surrogate's pass 1)
    const long np1 = sur ? (long)nb1 : (long)nb1 * ctx->sh1.wpb;
    const long np2 = (long)nbe * (kBlock / 64), np3 = (long)nbw * ctx->sh2.wpb;
    ctx->d_part.reserve((size_t)np1 * 2 + (size_t)np3 * 6 + (size_t)np2 + 16);
    ctx->d_out.reserve(16);
    double* part1 = ctx->d_part.p;
    double* part2 = part1 + (size_t)np1 * 2;
    double* part3 = part2 + (size_t)np2;
    Q* deriv = reinterpret_cast<Q*>(ctx->d_deriv.p);
    if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[0], st));
    pack_atoms<NT>(ctx);
    if constexpr (NT == 2 && std::is_same<Q, double>::value) {
        if (sur) {
            SurDev S{ctx->d_sp.p, reinterpret_cast<const __bf16*>(ctx->d_sfrag.p), ctx->d_smask.p, ctx->d_senv.p,
                     ctx->snorm};
            ctx->d_nfall.reserve(natms);
            HIP_OK(hipMemsetAsync(ctx->d_nfall.p, 0, sizeof(int) * natms, st));
            const int q = ctx->Pf.no_i;
            auto launch = [&](auto kern) {
                HIP_OK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kSurLds));
                hipLaunchKernelGGL(kern, dim3(nb1), dim3(kSurBlock), kSurLds, st, ctx->P, ctx->Pf, S, ctx->d_pos.p,
                                   ctx->d_arec.p, L, ctx->d_nlrp.p, ctx->d_quad32.p, ctx->d_rho7.p, ctx->d_phi7.p,
                                   ctx->d_Eown.p, ctx->d_pairE.p, (double*)deriv, part1, ctx->d_nfall.p,
                                   CutState{ctx->d_nlrp.p, force_shares<NT>(ctx)});
            };
            if (q == 5) launch(k_pass1_sur<5>);
            else if (q == 3) launch(k_pass1_sur<3>);
            else launch(k_pass1_sur<0>);
            launch_check("k_pass1_sur");
        }
    }
    if (!sur) {
        launch_pass1<NT, Q, NQ, false, false>(ctx, dim3(nb1), quad, PQ, L, deriv, part1);
        launch_check("k_pass1");
    }
    if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[1], st));
    hipLaunchKernelGGL((k_embed<NT, false>), dim3(nbe), dim3(kBlock), 0, st, ctx->P, ctx->d_alpha.p,
                       ctx->d_conc.p, L, ctx->d_Eown.p, ctx->d_pairE.p, ctx->d_F9.p, ctx->d_E.p, ctx->d_dE.p,
                       ctx->d_fa.p, ctx->d_arec.p, part2);
    launch_check("k_embed");
    ghost_copy(ctx, ctx->d_dE.p, NT);
    if (ctx->nghost > 0) {
        hipLaunchKernelGGL((k_ghost_copy<double4>), dim3(nblk(ctx->nghost)), dim3(kBlock), 0, st, ctx->d_arec.p, NT,
                           natms, ctx->nghost, ctx->d_ghost_owner.p);
        launch_check("k_ghost_copy arec");
    }
    if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[2], st));
    hipLaunchKernelGGL((k_pass2<NT, Q>), dim3(nbw), dim3(64 * ctx->sh2.wpb), 0, st, ctx->P, ctx->d_pos.p,
                       ctx->d_arec.p, L, ctx->d_nlrp.p, deriv, ctx->d_fown.p, force_shares<NT>(ctx), part3);
    launch_check("k_pass2");
    if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[3], st));
    hipLaunchKernelGGL((k_fgather<NT>), dim3(nb), dim3(kBlock), 0, st, natms, L, ctx->d_fown.p, force_shares<NT>(ctx),
                       ctx->d_fa.p, ctx->d_f.p);
    launch_check("k_fgather");
    double* out = ctx->d_out.p;
    {
        const double* parts[3] = {part1, part2, part3};
        const long nparts[3] = {np1, np2, np3};
        const int nv[3] = {2, 1, 6}, stride[3] = {2, 1, 6}, off[3] = {0, 2, 3};
        reduce_partials_segs(ctx, 3, parts, nparts, nv, stride, off, out);
    }
    if (ctx->timing) HIP_OK(hipEventRecord(ctx->ev[4], st));
    double h[9];
    HIP_OK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    ctx->n_inrange = (long)h[1];
    en_st[0] = h[0] + h[2];
    if (st_clc)
        for (int k = 0; k < 6; ++k) en_st[1 + k] = h[3 + k];
    if (ctx->timing) {
        float ms;
        for (int k = 0; k < 4; ++k) {
            HIP_OK(hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]));
            ctx->last_ms[k] = ms;
        }
        HIP_OK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]));
        ctx->last_ms[4] = ms;
    }
}

template <int NT, typename Q, int NQ>
double run_energy(mapp_dmd* ctx, const Q* quad, const FFParams<Q>& PQ) {
    hipStream_t st = ctx->stream;
    const int nb = nblk((long)ctx->natms * kEmbedG);
    const int nbw = shape_blocks(ctx->sh1, ctx->natms);
    PairBufs L{ctx->d_off.p, ctx->d_nl.p, ctx->d_roff.p, ctx->d_rev.p};
    const long np1 = (long)nbw * ctx->sh1.wpb, np2 = (long)nb * (kBlock / 64);  // one partial per wavefront
    ctx->d_part.reserve((size_t)np1 * 2 + (size_t)np2 + 16);
    ctx->d_out.reserve(16);
    double* part1 = ctx->d_part.p;
    double* part2 = part1 + (size_t)np1 * 2;
    pack_atoms<NT>(ctx);
    if (ctx->energy_full) launch_pass1<NT, Q, NQ, true, true>(ctx, dim3(nbw), quad, PQ, L, (Q*)nullptr, part1);
    else launch_pass1<NT, Q, NQ, true, false>(ctx, dim3(nbw), quad, PQ, L, (Q*)nullptr, part1);
    launch_check("k_pass1 energy");
    hipLaunchKernelGGL((k_embed<NT, true>), dim3(nb), dim3(kBlock), 0, st, ctx->P, ctx->d_alpha.p, ctx->d_conc.p,
                       L, ctx->d_Eown.p, ctx->d_pairE.p, ctx->d_F9.p, ctx->d_E.p, ctx->d_dE.p, ctx->d_fa.p,
                       ctx->d_arec.p, part2);
    launch_check("k_embed energy");
    double* out = ctx->d_out.p;
    {
        const double* parts[2] = {part1, part2};
        const long nparts[2] = {np1, np2};
        const int nv[2] = {1, 1}, stride[2] = {2, 1}, off[2] = {0, 1};
        reduce_partials_segs(ctx, 2, parts, nparts, nv, stride, off, out);
    }
    double h[2];
    HIP_OK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return h[0] + h[1];
}

template <int NT>
void dispatch_force(mapp_dmd* ctx, int st_clc, double* en_st) {
    const int q = ctx->P.no_i;
    // the networks serve no cell at the current bound: every pair takes the FP32
    // quadrature, which the mixed path's pass 1 evaluates faster than the
    // surrogate kernel's per-lane fallback
    const bool sur = ctx->prec == MAPP_DMD_PREC_SURROGATE && ctx->sur_served > 0.0;
    ctx->sur_last_quadrature = ctx->prec == MAPP_DMD_PREC_SURROGATE && !sur;
    if (sur) {
        run_force<NT, double, 0>(ctx, ctx->d_quad64.p, ctx->P, st_clc, en_st, true);
        return;
    }
    if (ctx->prec == MAPP_DMD_PREC_MIXED || ctx->prec == MAPP_DMD_PREC_SURROGATE) {
        if (q == 3) run_force<NT, float, 3>(ctx, ctx->d_quad32.p, ctx->Pf, st_clc, en_st);
        else if (q == 5) run_force<NT, float, 5>(ctx, ctx->d_quad32.p, ctx->Pf, st_clc, en_st);
        else run_force<NT, float, 0>(ctx, ctx->d_quad32.p, ctx->Pf, st_clc, en_st);
    } else {
        if (q == 3) run_force<NT, double, 3>(ctx, ctx->d_quad64.p, ctx->P, st_clc, en_st);
        else if (q == 5) run_force<NT, double, 5>(ctx, ctx->d_quad64.p, ctx->P, st_clc, en_st);
        else run_force<NT, double, 0>(ctx, ctx->d_quad64.p, ctx->P, st_clc, en_st);
    }
}

template <int NT>
double dispatch_energy(mapp_dmd* ctx) {
    const int q = ctx->P.no_i;
    // the surrogate's energy_calc runs its fallback, the FP32 quadrature
    if (ctx->prec == MAPP_DMD_PREC_MIXED || ctx->prec == MAPP_DMD_PREC_SURROGATE) {
        if (q == 3) return run_energy<NT, float, 3>(ctx, ctx->d_quad32.p, ctx->Pf);
        if (q == 5) return run_energy<NT, float, 5>(ctx, ctx->d_quad32.p, ctx->Pf);
        return run_energy<NT, float, 0>(ctx, ctx->d_quad32.p, ctx->Pf);
    }
    if (q == 3) return run_energy<NT, double, 3>(ctx, ctx->d_quad64.p, ctx->P);
    if (q == 5) return run_energy<NT, double, 5>(ctx, ctx->d_quad64.p, ctx->P);
    return run_energy<NT, double, 0>(ctx, ctx->d_quad64.p, ctx->P);
}

void upload_tables(mapp_dmd* ctx) {
    const int nt = ctx->nt, nr = ctx->nr;
    std::vector<double> q64((size_t)nt * nt * nr * kQuadRow, 0.0);
    std::vector<float> q32(q64.size(), 0.0f);
    for (int it = 0; it < nt; ++it)
        for (int jt = 0; jt < nt; ++jt) {
            const double* rs = ctx->rho7.data() + (size_t)type2rho(nt, it, jt) * nr * 7;
            const double* ps = ctx->phi7.data() + (size_t)type2phi(it, jt) * nr * 7;
            double* d = q64.data() + (size_t)(it * nt + jt) * kQuadRow;
            for (int m = 0; m < nr; ++m)
                for (int k = 0; k < 4; ++k) {
                    d[(size_t)m * quad_stride(nt) + k] = rs[(size_t)m * 7 + k];
                    d[(size_t)m * quad_stride(nt) + 4 + k] = ps[(size_t)m * 7 + k];
                }
        }
    for (size_t k = 0; k < q64.size(); ++k) q32[k] = (float)q64[k];
    ctx->d_quad64.reserve(q64.size());
    ctx->d_quad32.reserve(q32.size());
    ctx->d_rho7.reserve(ctx->rho7.size());
    ctx->d_phi7.reserve(ctx->phi7.size());
    ctx->d_F9.reserve(ctx->F9.size());
    HIP_OK(hipMemcpy(ctx->d_quad64.p, q64.data(), q64.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_quad32.p, q32.data(), q32.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_rho7.p, ctx->rho7.data(), ctx->rho7.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_phi7.p, ctx->phi7.data(), ctx->phi7.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_F9.p, ctx->F9.data(), ctx->F9.size() * sizeof(double), hipMemcpyHostToDevice));
    ctx->d_F7.reserve(ctx->F7.size());
    HIP_OK(hipMemcpy(ctx->d_F7.p, ctx->F7.data(), ctx->F7.size() * sizeof(double), hipMemcpyHostToDevice));
}

void fill_float_params(mapp_dmd* ctx) {
    const FFParams<double>& P = ctx->P;
    FFParams<float>& F = ctx->Pf;
    F.natms = P.natms;
    F.ntot = P.ntot;
    F.nr = P.nr;
    F.nrho = P.nrho;
    F.no_i = P.no_i;
    F.dr_inv = (float)P.dr_inv;
    F.drho_inv = (float)P.drho_inv;
    F.rho_max = (float)P.rho_max;
    F.rc = (float)P.rc;
    F.cut_sq = (float)P.cut_sq;
    F.cut_sq_mod = (float)P.cut_sq_mod;
    F.alpha_min = (float)P.alpha_min;
    F.alpha_max = (float)P.alpha_max;
    F.inv_amin = (float)P.inv_amin;
    F.inv_amax = (float)P.inv_amax;
    F.kbT = (float)P.kbT;
    F.beta = (float)P.beta;
    F.xw = (float)P.xw;
    for (int k = 0; k < kMaxPoints; ++k) {
        F.xi[k] = (float)P.xi[k];
        F.w0[k] = (float)P.w0[k];
        F.w1[k] = (float)P.w1[k];
        F.w2[k] = (float)P.w2[k];
    }
    for (int t = 0; t < kMaxTypes; ++t) F.c0[t] = (float)P.c0[t];
}

void sync_sizes(mapp_dmd* ctx) {
    ctx->P.natms = ctx->natms;
    ctx->P.ntot = ctx->natms + ctx->nghost;
    ctx->Pf.natms = ctx->P.natms;
    ctx->Pf.ntot = ctx->P.ntot;
}


void ensure_pair_cache(mapp_dmd* ctx);

template <int NT, typename Q, int NQ>
void run_pair_values(mapp_dmd* ctx, const Q* quad, const FFParams<Q>& PQ) {
    const size_t np = (size_t)ctx->npairs + 1;
    ctx->d_rhoc.reserve(np * NT * NT);
    ctx->d_phic.reserve(np * NT * (NT + 1) / 2);
    hipLaunchKernelGGL((k_pair_values<NT, Q, NQ>), dim3(nblk(ctx->npairs)), dim3(kBlock), 0, ctx->stream, ctx->P, PQ,
                       ctx->d_pos.p, ctx->d_alpha.p, ctx->d_nl.p, ctx->d_pair_i.p, (long)ctx->npairs, quad,
                       ctx->d_rho7.p, ctx->d_phi7.p, ctx->d_rhoc.p, ctx->d_phic.p);
    launch_check("k_pair_values");
}

template <int NT>
void pair_values_dispatch(mapp_dmd* ctx) {
    const int q = ctx->P.no_i;
    // the surrogate's c_d_calc / calc_g pair values come from its fallback, the FP32 quadrature
    if (ctx->prec == MAPP_DMD_PREC_MIXED || ctx->prec == MAPP_DMD_PREC_SURROGATE) {
        if (q == 3) run_pair_values<NT, float, 3>(ctx, ctx->d_quad32.p, ctx->Pf);
        else if (q == 5) run_pair_values<NT, float, 5>(ctx, ctx->d_quad32.p, ctx->Pf);
        else run_pair_values<NT, float, 0>(ctx, ctx->d_quad32.p, ctx->Pf);
    } else {
        if (q == 3) run_pair_values<NT, double, 3>(ctx, ctx->d_quad64.p, ctx->P);
        else if (q == 5) run_pair_values<NT, double, 5>(ctx, ctx->d_quad64.p, ctx->P);
        else run_pair_values<NT, double, 0>(ctx, ctx->d_quad64.p, ctx->P);
    }
}

RevBufs rev_half(mapp_dmd* ctx) { return RevBufs{ctx->d_off.p, ctx->d_nl.p, ctx->d_roff.p, ctx->d_rev.p, ctx->d_pair_i.p}; }
RevBufs rev_second(mapp_dmd* ctx) {
    return RevBufs{ctx->d_off2.p, ctx->d_nl2.p, ctx->d_roff2.p, ctx->d_rev2.p, ctx->d_pair_i2.p};
}

// site table of the pair values (both ends of every contributing pair)
template <int NT>
void build_site_table(mapp_dmd* ctx) {
    hipStream_t st = ctx->stream;
    const int natms = ctx->natms;
    const RevBufs L = rev_half(ctx);
    ctx->d_cnt.reserve(natms + 1);
    hipLaunchKernelGGL((k_table_count<NT>), dim3(nblk((long)natms * kTableG)), dim3(kBlock), 0, st, L, natms, ctx->d_rhoc.p,
                       ctx->d_phic.p, ctx->d_cnt.p);
    launch_check("k_table_count");
    const long ne = counts_to_offsets(ctx, ctx->d_cnt.p, ctx->d_toff, natms);
    ctx->d_tidx.reserve(ne + 1);
    ctx->d_trho.reserve((size_t)(ne + 1) * NT * NT);
    ctx->d_tphi.reserve((size_t)(ne + 1) * NT * (NT + 1) / 2);
    hipLaunchKernelGGL((k_table_fill<NT>), dim3(nblk((long)natms * kTableG)), dim3(kBlock), 0, st, L, natms, ctx->d_rhoc.p,
                       ctx->d_phic.p, ctx->d_toff.p, ctx->d_tidx.p, ctx->d_trho.p, ctx->d_tphi.p);
    launch_check("k_table_fill");
    ctx->ntable = ne;
}

SiteTable site_table(mapp_dmd* ctx) { return SiteTable{ctx->d_toff.p, ctx->d_tidx.p, ctx->d_trho.p, ctx->d_tphi.p}; }

void ensure_pair_cache(mapp_dmd* ctx) {
    if (ctx->cache_valid) return;
    with_nt(ctx->nt, [&](auto c) {
        pair_values_dispatch<decltype(c)::value>(ctx);
        build_site_table<decltype(c)::value>(ctx);
    });
    ctx->cache_valid = true;
}

template <int NT>
void build_list2(mapp_dmd* ctx) {
    hipStream_t st = ctx->stream;
    const int natms = ctx->natms;
    ctx->d_cnt.reserve(natms + 1);
    hipLaunchKernelGGL((k_list2<false>), dim3(nblk((long)natms * 64)), dim3(kBlock), 0, st, ctx->d_pos.p, natms, ctx->d_off.p,
                       ctx->d_nl.p, ctx->rsq_crd, ctx->d_cnt.p, (int*)nullptr);
    launch_check("k_list2 count");
    ctx->npairs2 = counts_to_offsets(ctx, ctx->d_cnt.p, ctx->d_off2, natms);
    ctx->d_nl2.reserve(ctx->npairs2 + 1);
    hipLaunchKernelGGL((k_list2<true>), dim3(nblk((long)natms * 64)), dim3(kBlock), 0, st, ctx->d_pos.p, natms, ctx->d_off.p,
                       ctx->d_nl.p, ctx->rsq_crd, ctx->d_off2.p, ctx->d_nl2.p);
    launch_check("k_list2 fill");
    HIP_OK(hipMemsetAsync(ctx->d_cnt.p, 0, sizeof(int) * (natms + 1), st));
    hipLaunchKernelGGL(k_rev_count, dim3(nblk((long)natms * kRevG)), dim3(kBlock), 0, st, ctx->d_off2.p, ctx->d_nl2.p, natms,
                       ctx->d_cnt.p);
    launch_check("k_rev_count 2");
    long nrev = counts_to_offsets(ctx, ctx->d_cnt.p, ctx->d_roff2, natms);
    ctx->d_rev2.reserve(nrev + 1);
    HIP_OK(hipMemsetAsync(ctx->d_cnt.p, 0, sizeof(int) * (natms + 1), st));
    hipLaunchKernelGGL(k_rev_fill, dim3(nblk((long)natms * kRevG)), dim3(kBlock), 0, st, ctx->d_off2.p, ctx->d_nl2.p, natms,
                       ctx->d_roff2.p, ctx->d_cnt.p, ctx->d_rev2.p);
    launch_check("k_rev_fill 2");
    sort_segments(ctx, ctx->d_roff2.p, ctx->d_nl2.p, ctx->d_rev2.p, nrev);
    ctx->d_pair_i2.reserve(ctx->npairs2 + 1);
    hipLaunchKernelGGL(k_pair_owner, dim3(nblk(natms)), dim3(kBlock), 0, st, ctx->d_off2.p, natms, ctx->d_pair_i2.p);
    launch_check("k_pair_owner 2");
    // both ends of every 2nd-list pair per site
    const RevBufs L2 = rev_second(ctx);
    hipLaunchKernelGGL(k_list2_count, dim3(nblk(natms)), dim3(kBlock), 0, st, L2, natms, ctx->d_cnt.p);
    launch_check("k_list2_count");
    const long nf = counts_to_offsets(ctx, ctx->d_cnt.p, ctx->d_foff2, natms);
    ctx->d_fidx2.reserve(nf + 1);
    hipLaunchKernelGGL(k_list2_fill, dim3(nblk(natms)), dim3(kBlock), 0, st, L2, natms, ctx->d_foff2.p, ctx->d_fidx2.p);
    launch_check("k_list2_fill");
    const int ntot = natms + ctx->nghost;
    ctx->d_crd.reserve(ntot);
    hipLaunchKernelGGL(k_crd, dim3(nblk(natms)), dim3(kBlock), 0, st, ctx->d_off2.p, ctx->d_roff2.p, natms, ctx->d_crd.p);
    launch_check("k_crd");
    if (ctx->nghost > 0) {
        hipLaunchKernelGGL((k_ghost_copy<int>), dim3(nblk(ctx->nghost)), dim3(kBlock), 0, st, ctx->d_crd.p, 1, natms,
                           ctx->nghost, ctx->d_ghost_owner.p);
        launch_check("k_ghost_copy crd");
    }
    HIP_OK(hipStreamSynchronize(st));
    ctx->have_list2 = true;
}

// c_d_calc: density and n from the site table, then rates over the 2nd list
template <int NT, int VAR>
void run_c_d(mapp_dmd* ctx) {
    hipStream_t st = ctx->stream;
    ensure_pair_cache(ctx);
    const int nbt = nblk((long)ctx->natms * kTableG), nb2 = nblk((long)ctx->natms * kList2G);
    const SiteTable T = site_table(ctx);
    hipLaunchKernelGGL((k_density_table<NT>), dim3(nbt), dim3(kBlock), 0, st, ctx->P, ctx->d_conc.p, T, ctx->d_F9.p,
                       ctx->d_E.p, ctx->d_n.p, ctx->d_dE.p, (double*)nullptr);
    launch_check("k_density_table");
    ghost_copy(ctx, ctx->d_dE.p, NT);
    constexpr int CLOSE = VAR == 2 ? 1 : 0;
    hipLaunchKernelGGL((k_chem_pot<NT, CLOSE>), dim3(nbt), dim3(kBlock), 0, st, ctx->P, ctx->d_alpha.p, ctx->d_conc.p,
                       T, ctx->d_dE.p, ctx->d_qidx.p, ctx->d_n.p);
    launch_check("k_chem_pot");
    ghost_copy(ctx, ctx->d_n.p, NT);
    hipLaunchKernelGGL((k_c_d<NT, VAR>), dim3(nb2), dim3(kBlock), 0, st, ctx->P, ctx->d_conc.p, ctx->d_foff2.p,
                       ctx->d_fidx2.p, ctx->d_n.p, ctx->d_cd.p);
    launch_check("k_c_d");
    HIP_OK(hipStreamSynchronize(st));
}

// device memory calc_g needs beyond the per-atom vectors: the per-block partials
// of the cost, the reduction's stage and the reduced cost
struct GScratch {
    double* part;
    double* stage;
    double* out;
};

// blocks of the 2nd-list sweeps of calc_g (one partial each)
inline long calc_g_blocks(const mapp_dmd* ctx) { return nblk((long)ctx->natms * kList2G); }

// fixed: caller-provided scratch; the site table must then be current, since
// building it allocates
template <int NT, int VAR>
double run_calc_g(mapp_dmd* ctx, int chk, double alpha_t, const double* a, double* g,
                  const GScratch* fixed = nullptr) {
    hipStream_t st = ctx->stream;
    if (!fixed) ensure_pair_cache(ctx);
    const int nbt = nblk((long)ctx->natms * kTableG), nb2 = (int)calc_g_blocks(ctx);
    const SiteTable T = site_table(ctx);
    hipLaunchKernelGGL((k_density_table<NT>), dim3(nbt), dim3(kBlock), 0, st, ctx->P, ctx->d_conc.p, T, ctx->d_F9.p,
                       ctx->d_E.p, ctx->d_n.p, ctx->d_dE.p, ctx->d_ddE.p);
    launch_check("k_density_table g");
    ghost_copy(ctx, ctx->d_dE.p, NT);
    constexpr int CLOSE = VAR == 2 ? 3 : 2;
    hipLaunchKernelGGL((k_chem_pot<NT, CLOSE>), dim3(nbt), dim3(kBlock), 0, st, ctx->P, ctx->d_alpha.p, ctx->d_conc.p,
                       T, ctx->d_dE.p, ctx->d_qidx.p, ctx->d_n.p);
    launch_check("k_chem_pot g");
    ghost_copy(ctx, ctx->d_n.p, NT);
    GScratch S;
    if (fixed) {
        S = *fixed;
    } else {
        ctx->d_part.reserve((size_t)nb2 + 16);
        ctx->d_out.reserve(16);
        ctx->d_stage.reserve((size_t)reduce_stage_doubles(nb2));
        S = GScratch{ctx->d_part.p, ctx->d_stage.p, ctx->d_out.p};
    }
    hipLaunchKernelGGL((k_calc_g_s<NT, VAR>), dim3(nb2), dim3(kBlock), 0, st, ctx->P, alpha_t, ctx->d_conc.p, a,
                       ctx->d_foff2.p, ctx->d_fidx2.p, ctx->d_n.p, ctx->d_s.p, ctx->d_cd.p, S.part);
    launch_check("k_calc_g_s");
    reduce_partials_in(ctx, S.part, nb2, 1, 1, 0, S.stage, S.out);
    double inner = 0.0;
    HIP_OK(hipMemcpyAsync(&inner, S.out, sizeof(double), hipMemcpyDeviceToHost, st));
    if (!chk) {
        ghost_copy(ctx, ctx->d_s.p, NT);
        hipLaunchKernelGGL((k_calc_g_t<NT, VAR>), dim3(nb2), dim3(kBlock), 0, st, ctx->P, alpha_t, ctx->d_conc.p,
                           ctx->d_foff2.p, ctx->d_fidx2.p, ctx->d_n.p, ctx->d_s.p, ctx->d_t.p, g);
        launch_check("k_calc_g_t");
        ghost_copy(ctx, ctx->d_t.p, NT);
        hipLaunchKernelGGL((k_calc_g_v<NT>), dim3(nbt), dim3(kBlock), 0, st, ctx->P, alpha_t, T, ctx->d_dE.p,
                           ctx->d_t.p, ctx->d_v.p, g);
        launch_check("k_calc_g_v");
        ghost_copy(ctx, ctx->d_v.p, NT);
        ghost_copy(ctx, ctx->d_ddE.p, NT);
        hipLaunchKernelGGL((k_calc_g_w<NT>), dim3(nbt), dim3(kBlock), 0, st, ctx->P, alpha_t, ctx->d_conc.p, T,
                           ctx->d_ddE.p, ctx->d_v.p, g);
        launch_check("k_calc_g_w");
    }
    HIP_OK(hipStreamSynchronize(st));
    return inner * 0.25;
}

template <int NT>
void c_d_variant(mapp_dmd* ctx) {
    if (ctx->variant == 0) run_c_d<NT, 0>(ctx);
    else if (ctx->variant == 1) run_c_d<NT, 1>(ctx);
    else run_c_d<NT, 2>(ctx);
}
template <int NT>
double calc_g_variant(mapp_dmd* ctx, int chk, double alpha_t, const double* a, double* g,
                      const GScratch* fixed = nullptr) {
    if (ctx->variant == 0) return run_calc_g<NT, 0>(ctx, chk, alpha_t, a, g, fixed);
    if (ctx->variant == 1) return run_calc_g<NT, 1>(ctx, chk, alpha_t, a, g, fixed);
    return run_calc_g<NT, 2>(ctx, chk, alpha_t, a, g, fixed);
}


// ------------------------------------------------------------------ surrogate
constexpr int kNormFloats = 8 + 2 * 4 * kOut;

void norm_to_flat(const SurNorm& N, float* f) {
    const float hdr[8] = {N.r_lo, N.r_hi, N.a_lo, N.a_hi, N.r_c, N.r_s, N.a_c, N.a_s};
    std::memcpy(f, hdr, sizeof(hdr));
    std::memcpy(f + 8, N.y_mean, sizeof(N.y_mean));
    std::memcpy(f + 8 + 4 * kOut, N.y_std, sizeof(N.y_std));
}
void norm_from_flat(SurNorm& N, const float* f) {
    N.r_lo = f[0]; N.r_hi = f[1]; N.a_lo = f[2]; N.a_hi = f[3];
    N.r_c = f[4]; N.r_s = f[5]; N.a_c = f[6]; N.a_s = f[7];
    std::memcpy(N.y_mean, f + 8, sizeof(N.y_mean));
    std::memcpy(N.y_std, f + 8 + 4 * kOut, sizeof(N.y_std));
}

void sur_samples(mapp_dmd* ctx, const SurNorm& N, long n_per_tp, unsigned long long seed, DevBuf<float>& X,
                 DevBuf<float>& T, DevBuf<float>& B) {
    const long n = n_per_tp * 4;
    X.reserve(n * 3);
    T.reserve(n * kOut);
    B.reserve(n * kOut);
    const int q = ctx->Pf.no_i;
    const dim3 g(nblk(n)), b(kBlock);
    if (q == 3)
        hipLaunchKernelGGL((k_sur_samples<2, 3>), g, b, 0, ctx->stream, ctx->Pf, N, ctx->d_quad32.p, n_per_tp, seed,
                           X.p, T.p, B.p);
    else if (q == 5)
        hipLaunchKernelGGL((k_sur_samples<2, 5>), g, b, 0, ctx->stream, ctx->Pf, N, ctx->d_quad32.p, n_per_tp, seed,
                           X.p, T.p, B.p);
    else
        hipLaunchKernelGGL((k_sur_samples<2, 0>), g, b, 0, ctx->stream, ctx->Pf, N, ctx->d_quad32.p, n_per_tp, seed,
                           X.p, T.p, B.p);
    launch_check("k_sur_samples");
}

// one gradient evaluation: grad of each network's mean squared error over its n_per_tp samples
void sur_gradient(mapp_dmd* ctx, const SurNorm& N, const float* params, const float* X, const float* T,
                  const int* idx, long n_per_tp, float inv_count, float* gpart, float* lpart) {
    const int bpt = (int)(n_per_tp / kMB);
    HIP_OK(hipFuncSetAttribute((const void*)k_sur_train, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kSurTrainLds));
    hipLaunchKernelGGL(k_sur_train, dim3(4 * bpt), dim3(256), kSurTrainLds, ctx->stream, params, N, X, T, idx, bpt,
                       inv_count, gpart, lpart);
    launch_check("k_sur_train");
}

// Atoms::update_0 (src/atoms.cpp:1780-1865) for owned x already in d_xaos
// (cartesian + alpha rows): full-skin test against the positions of the last
// build, then a rebuild (x2s, ghosts, bins, lists) or a forward ghost update.
int update_x_device(mapp_dmd* ctx) {
    const int nt = acols(ctx), natms = ctx->natms;
    if (is_md(ctx)) scatter_x<1, 0>(ctx, ctx->d_xaos.p);
    else with_nt(nt, [&](auto c) { scatter_x<decltype(c)::value>(ctx, ctx->d_xaos.p); });
    HIP_OK(hipMemsetAsync(ctx->d_iflag.p, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_skin_check, dim3(nblk(natms)), dim3(kBlock), 0, ctx->stream, ctx->d_pos.p, ctx->d_x0.p, natms,
                       ctx->skin * ctx->skin, ctx->d_iflag.p);
    launch_check("k_skin_check");
    int flag = 0;
    HIP_OK(hipMemcpyAsync(&flag, ctx->d_iflag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (flag) {
        hipLaunchKernelGGL(k_x2s, dim3(nblk(natms)), dim3(kBlock), 0, ctx->stream, ctx->d_pos.p, natms, ctx->bp);
        launch_check("k_x2s");
        rebuild_any(ctx);
        sync_sizes(ctx);
    } else {
        with_nt(nt, [&](auto c) { ghost_update_x<decltype(c)::value>(ctx); });  // EAM: one column
        HIP_OK(hipStreamSynchronize(ctx->stream));
    }
    ctx->cache_valid = false;
    return flag;
}

// ------------------------------------------------------------------ minimiser
void force_any(mapp_dmd* ctx, int st_clc, double* en) {
    if (is_md(ctx)) run_eam(ctx, false, st_clc, en);
    else with_nt(ctx->nt, [&](auto c) { dispatch_force<decltype(c)::value>(ctx, st_clc, en); });
}
double energy_any(mapp_dmd* ctx) {
    if (is_md(ctx)) return run_eam(ctx, true, 0, nullptr);
    return with_nt(ctx->nt, [&](auto c) { return dispatch_energy<decltype(c)::value>(ctx); });
}

constexpr int kVecBlocks = 1024;  // grid of the grid-stride vector kernels (fixed: deterministic sums)

template <int ND>
void dots(mapp_dmd* ctx, long n, const DotArgs& D, double* res) {
    ctx->d_part.reserve((size_t)kVecBlocks * 8 + 16);
    ctx->d_out.reserve(16);
    hipLaunchKernelGGL((k_dots<ND>), dim3(kVecBlocks), dim3(kBlock), 0, ctx->stream, D, n, ctx->d_part.p);
    launch_check("k_dots");
    reduce_partials(ctx, ctx->d_part.p, kVecBlocks, ND, ND, 0, ctx->d_out.p);
    HIP_OK(hipMemcpyAsync(res, ctx->d_out.p, sizeof(double) * ND, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
}
double dot1(mapp_dmd* ctx, long n, const double* a, const double* b) {
    DotArgs D{{a, nullptr, nullptr, nullptr}, {b, nullptr, nullptr, nullptr}};
    double r;
    dots<1>(ctx, n, D, &r);
    return r;
}

// LineSearch_BackTrack::line_min, fixed box (src/line_search.cpp:244-341)
int line_min(mapp_dmd* ctx, long n, double* nrgy, double* alph, long* n_energy) {
    const double c = 0.4, rho = 0.5, a_max = 1.0, d_max = 1.0;
    hipStream_t st = ctx->stream;
    const int xd = xcols(ctx);
    double* h = ctx->d_lh.p;
    const double* x0 = ctx->d_lx0.p;
    // min alpha of the current x (x0 at this point) and max |h| (after normalising)
    double a_min = 0.0;
    double hh = dot1(ctx, n, h, h);
    hh = d_max / std::sqrt(hh);
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(kBlock), 0, st, n, 0.0, h, hh, h);
    launch_check("k_axpby");
    const double inner = dot1(ctx, n, ctx->d_f.p, h);
    ctx->d_part.reserve((size_t)kVecBlocks * 8 + 16);
    hipLaunchKernelGGL(k_extrema, dim3(kVecBlocks), dim3(kBlock), 0, st, n, xd, h, x0, ctx->d_part.p);
    launch_check("k_extrema");
    std::vector<double> ex((size_t)kVecBlocks * 2);
    HIP_OK(hipMemcpyAsync(ex.data(), ctx->d_part.p, ex.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double max_h = 0.0, min_a = INFINITY;
    for (int b = 0; b < kVecBlocks; ++b) {
        max_h = std::max(max_h, ex[b * 2]);
        min_a = std::min(min_a, ex[b * 2 + 1]);
    }
    if (!is_md(ctx)) a_min = std::min(a_min, 0.99 * min_a);
    *alph = 0.0;
    if (inner <= 0) return 1;
    if (max_h == 0) return 2;
    double a = std::min(a_max, max_h);
    if (a <= a_min) return 3;
    auto energy_at = [&](double am) {
        hipLaunchKernelGGL(k_step, dim3(nblk(n)), dim3(kBlock), 0, st, n, x0, am, h, ctx->d_xaos.p);
        launch_check("k_step");
        update_x_device(ctx);
        ++*n_energy;
        return energy_any(ctx);
    };
    while (true) {
        const double ideal = *nrgy - a * c * inner;
        const double cur = energy_at(a);
        if (cur <= ideal) {
            *nrgy = cur;
            *alph = a;
            return 0;
        }
        a *= rho;
        if (a <= a_min) {
            *nrgy = energy_at(0.0);
            return 3;
        }
    }
}

// ------------------------------------------------------------------ BDF clock
struct BdfState {
    mapp_dmd* ctx;
    long n;
    int nt;
    double t[kBdfMaxOrder + 1] = {0};
    double* y[kBdfMaxOrder] = {nullptr};
    double *dy, *a, *y0, *g, *g0, *h, *c0, *e_n;
    double beta = 0.0;
    long n_calc_g = 0;
    bool fix = false;  // correct the sign of the c1 c2 term of dalpha_y[0]

    double calc_g(int chk) {
        ++n_calc_g;
        return with_nt(nt, [&](auto c) { return calc_g_variant<decltype(c)::value>(ctx, chk, beta, a, g); });
    }
    void set_c_from(const double* src) {
        if (src != ctx->d_conc.p)
            HIP_OK(hipMemcpyAsync(ctx->d_conc.p, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        ghost_copy(ctx, ctx->d_conc.p, nt);
    }
    double min_blocks(const double* part, int nb, double init) {
        std::vector<double> h_(nb);
        HIP_OK(hipMemcpyAsync(h_.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        double r = init;
        for (double v : h_) r = v < r ? v : r;
        return r;
    }
    double bound(const double* base, const double* dir, double scale, double init) {
        ctx->d_part.reserve((size_t)kVecBlocks * 8 + 16);
        hipLaunchKernelGGL(k_bdf_bound, dim3(kVecBlocks), dim3(kBlock), 0, ctx->stream, n, base, dir, scale,
                           ctx->d_part.p);
        launch_check("k_bdf_bound");
        return min_blocks(ctx->d_part.p, kVecBlocks, init);
    }
    double dot(const double* u, const double* v) { return dot1(ctx, n, u, v); }

    // Clock_BDF::interpolate (src/clock_bdf.cpp:295-369)
    int interpolate(int q, double del_t) {
        double alpha_y[kBdfMaxOrder], dalpha_y[kBdfMaxOrder];
        const double t_new = t[0] + del_t;
        double c0_ = 1.0, c1 = 0.0, c2 = 0.0;
        for (int i = 1; i < q; ++i) {
            double tmp0 = (t_new - t[0]) / (t[i] - t[0]);
            double tmp1 = tmp0 * tmp0;
            for (int j = 1; j < q; ++j)
                if (i != j) tmp1 *= (t_new - t[j]) / (t[i] - t[j]);
            alpha_y[i] = tmp1;
            c0_ *= (t_new - t[i]) / (t[0] - t[i]);
            c1 += 1.0 / (t[0] - t[i]);
            c2 += 1.0 / (t_new - t[i]);
        }
        alpha_y[0] = (1.0 - (t_new - t[0]) * c1) * c0_;
        const double alpha_dy_0 = c0_ * (t_new - t[0]);
        const double c3 = 2.0 / (t_new - t[0]);
        const double dalpha_dy_0 = alpha_dy_0 * (c2 + 0.5 * c3);
        // the reference adds the (t_new - t[0]) c1 c2 term (src/clock_bdf.cpp:327); the
        // derivative of alpha_y[0] subtracts it
        dalpha_y[0] = fix ? c0_ * (c2 - c1 - (t_new - t[0]) * c1 * c2) : c0_ * (c2 - c1 + (t_new - t[0]) * c1 * c2);
        for (int i = 1; i < q; ++i) dalpha_y[i] = alpha_y[i] * (c2 + c3 - 1.0 / (t_new - t[i]));
        double tmp0 = 0.0;
        for (int i = 0; i < q; ++i) tmp0 += 1.0 / (double)(i + 1);
        beta = del_t / tmp0;
        BdfInterp C{};
        C.b_ady0 = beta * dalpha_dy_0;
        C.ady0 = alpha_dy_0;
        BdfHist Y{};
        for (int j = 0; j < q; ++j) {
            C.b_ay[j] = beta * dalpha_y[j];
            C.ay[j] = alpha_y[j];
            Y.y[j] = y[j];
        }
        HIP_OK(hipMemsetAsync(ctx->d_iflag.p, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(k_bdf_interp, dim3(nblk(n)), dim3(kBlock), 0, ctx->stream, n, q, C, Y, dy, a, y0,
                           ctx->d_iflag.p);
        launch_check("k_bdf_interp");
        int flag = 0;
        HIP_OK(hipMemcpyAsync(&flag, ctx->d_iflag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return flag ? -1 : 1;
    }

    // Clock_BDF::solve (src/clock_bdf.cpp:677-883), curr_t = t[0] + del_t; -1 if c left [0,1]
    double solve(int q, double del_t, double* err, int* iters) {
        const double min_gamma = 1.0e-10, gamma_red = 0.8, slope = 0.4, m_tol = 1.0e-10, a_tol = 1.0e-6;
        const int max_iter = 50;
        hipStream_t st = ctx->stream;
        double* c = ctx->d_conc.p;
        const double ratio0 = bound(y[0], dy, del_t, 1.0);
        hipLaunchKernelGGL(k_bdf_pred, dim3(nblk(n)), dim3(kBlock), 0, st, n, y[0], dy, del_t, ratio0, c);
        launch_check("k_bdf_pred");
        set_c_from(c);
        double curr_cost = calc_g(0);
        HIP_OK(hipMemcpyAsync(h, g, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        double g0_g0 = dot(g, g), g_h = g0_g0, max_gamma = 1.0;
        int iter = 0;
        while (curr_cost > m_tol * (double)n && iter < max_iter && max_gamma > min_gamma) {
            HIP_OK(hipMemcpyAsync(g0, g, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIP_OK(hipMemcpyAsync(c0, c, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            max_gamma = bound(c0, h, 1.0, 0.99) * 0.9;
            int chk = 1;
            const double cost = curr_cost;
            while (chk && max_gamma > min_gamma) {
                hipLaunchKernelGGL(k_step, dim3(nblk(n)), dim3(kBlock), 0, st, n, c0, max_gamma, h, c);
                launch_check("k_step");
                set_c_from(c);
                curr_cost = calc_g(0);
                const double ideal = cost - slope * max_gamma * g_h;
                if (curr_cost < ideal) chk = 0;
                max_gamma *= gamma_red;
                if (max_gamma <= min_gamma) {
                    set_c_from(c0);
                    curr_cost = calc_g(1);
                }
            }
            if (chk == 0) {
                double r2[2];
                DotArgs D{{g, g, nullptr, nullptr}, {g0, g, nullptr, nullptr}};
                dots<2>(ctx, n, D, r2);
                const double g_g0 = r2[0], g_g = r2[1];
                const double r = (g_g - g_g0) / g0_g0;
                g0_g0 = g_g;
                hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(kBlock), 0, st, n, 1.0, g, r, h);
                launch_check("k_axpby");
                g_h = dot(h, g);
                if (g_h < 0.0) {
                    HIP_OK(hipMemcpyAsync(h, g, n * sizeof(double), hipMemcpyDeviceToDevice, st));
                    g_h = g_g;
                }
            }
            ++iter;
        }
        *iters = iter;
        const double curr_t = t[0] + del_t;
        double a_1 = 1.0, l1_q = 0.0;
        for (int i = 0; i < q; ++i) {
            l1_q += 1.0 / (curr_t - t[i]);
            if (i != 0) a_1 *= (curr_t - t[i]) / (t[0] - t[i]);
        }
        const double err_pre_fac = -1.0 / (l1_q * (1.0 + a_1));
        ctx->d_part.reserve((size_t)kVecBlocks * 8 + 16);
        hipLaunchKernelGGL(k_bdf_err, dim3(kVecBlocks), dim3(kBlock), 0, st, n, y0, c, ctx->d_cd.p, ctx->d_part.p);
        launch_check("k_bdf_err");
        reduce_partials(ctx, ctx->d_part.p, kVecBlocks, 2, 2, 0, ctx->d_out.p);
        double e2[2];
        HIP_OK(hipMemcpyAsync(e2, ctx->d_out.p, sizeof(e2), hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_minmax, dim3(kVecBlocks), dim3(kBlock), 0, st, n, c, ctx->d_part.p);
        launch_check("k_minmax");
        std::vector<double> mm((size_t)kVecBlocks * 2);
        HIP_OK(hipMemcpyAsync(mm.data(), ctx->d_part.p, mm.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        *err = std::sqrt(e2[0] / ((double)n * a_tol)) * std::fabs(err_pre_fac);
        for (int b = 0; b < kVecBlocks; ++b)
            if (mm[b * 2] < 0.0 || mm[b * 2 + 1] > 1.0) return -1.0;
        return curr_cost / (m_tol * (double)n);
    }

    // Clock_BDF::hi_lo_err_est (src/clock_bdf.cpp:372-478)
    void hi_lo(int q, double del_t, bool hi_avail, bool lo_avail, double* lo_err, double* hi_err) {
        const double a_tol = 1.0e-6;
        const double curr_t = t[0] + del_t;
        double a_0 = 1.0, a_1 = 1.0, l1_q = 0.0;
        for (int i = 0; i < q; ++i) {
            l1_q += 1.0 / (curr_t - t[i]);
            a_0 *= curr_t - t[i];
            if (i != 0) a_1 *= (curr_t - t[i]) / (t[0] - t[i]);
        }
        BdfHiLo H{};
        H.hi = hi_avail;
        H.lo = lo_avail;
        H.q = q;
        H.Q_n = 1.0;
        H.hi_pre = 1.0;
        if (hi_avail) {
            double hi_a_0 = 1.0, hi_a_1 = 1.0;
            for (int i = 0; i < q; ++i) {
                hi_a_0 *= t[0] - t[i + 1];
                if (i != 0) hi_a_1 *= (t[0] - t[i + 1]) / (t[1] - t[i + 1]);
            }
            H.Q_n = (a_0 / hi_a_0) * (1.0 + a_1) / (1.0 + hi_a_1) * (del_t / (t[0] - t[1]));
            const double hi_l1_q = l1_q + 1.0 / (curr_t - t[q]);
            H.hi_pre = -(curr_t - t[q]) / (q + 2);
            H.hi_pre *= 1.0 / (hi_l1_q * (1.0 + a_1) * del_t * del_t);
        }
        BdfHist Y{};
        if (lo_avail) {
            const double lo_pre = -del_t * a_0 / (curr_t - t[q - 1]);
            double c0_ = 0.0, c1 = 1.0;
            for (int i = 0; i < q - 1; ++i) {
                const double tmp0 = 1.0 / (curr_t - t[i]);
                c0_ += tmp0;
                c1 *= tmp0;
                H.la[i] = tmp0 * tmp0;
                for (int j = 0; j < q - 1; ++j)
                    if (i != j) H.la[i] *= 1.0 / (t[i] - t[j]);
                H.la[i] *= lo_pre;
                Y.y[i] = y[i];
            }
            H.la_dy = c1 * lo_pre;
            H.la_y = -c0_ * c1 * lo_pre;
            H.c1c0 = c1 / c0_;
        }
        ctx->d_part.reserve((size_t)kVecBlocks * 8 + 16);
        hipLaunchKernelGGL(k_bdf_hilo, dim3(kVecBlocks), dim3(kBlock), 0, ctx->stream, n, H, Y, ctx->d_conc.p,
                           ctx->d_cd.p, y0, e_n, ctx->d_part.p);
        launch_check("k_bdf_hilo");
        reduce_partials(ctx, ctx->d_part.p, kVecBlocks, 2, 2, 0, ctx->d_out.p);
        double r[2];
        HIP_OK(hipMemcpyAsync(r, ctx->d_out.p, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        *lo_err = std::sqrt(r[0] / ((double)n * a_tol));
        *hi_err = std::sqrt(r[1] / ((double)n * a_tol));
    }
};

// ------------------------------------------------------------------ batched cells
// Cells are built one by one in scratch contexts with the single-cell pipeline,
// then concatenated: the owned sites of all cells first (cell order), then all
// ghosts (cell order).  Site j is owned iff j < natms, as for one cell, so every
// force, mobility and BDF kernel runs unchanged in one launch over all cells;
// lists keep each cell's order, so per-site sums match a cell run alone.
template <int NT>
__global__ void k_merge_sites(int n_own, int n_gh, const double4* __restrict__ pos, const double* __restrict__ alpha,
                              const double* __restrict__ conc, const int* __restrict__ owner, int own_base,
                              int gh_dst, int gh_index_base, double4* __restrict__ Pos, double* __restrict__ Alpha,
                              double* __restrict__ Conc, int* __restrict__ Owner) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_own + n_gh) return;
    const int dst = s < n_own ? own_base + s : gh_dst + (s - n_own);
    Pos[dst] = pos[s];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        Alpha[(size_t)dst * NT + t] = alpha[(size_t)s * NT + t];
        Conc[(size_t)dst * NT + t] = conc[(size_t)s * NT + t];
    }
    if (s >= n_own) Owner[gh_index_base + (s - n_own)] = own_base + owner[s - n_own];
}

__global__ void k_merge_list(int n_own, const int* __restrict__ off, const int* __restrict__ nl, int own_base,
                             int pair_base, int gh_dst, int* __restrict__ Off, int* __restrict__ Nl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_own) return;
    Off[own_base + i] = pair_base + off[i];
    for (int p = off[i]; p < off[i + 1]; ++p) {
        const int j = nl[p];
        Nl[pair_base + p] = j < n_own ? own_base + j : gh_dst + (j - n_own);
    }
}

template <int NT>
void init_batch(mapp_dmd* ctx, double skin) {
    const int ncell = (int)ctx->cell_natms.size();
    std::vector<mapp_dmd*> sub(ncell, nullptr);
    struct Guard {
        std::vector<mapp_dmd*>& v;
        ~Guard() {
            for (auto* p : v)
                if (p) mapp_dmd_destroy(p);
        }
    } guard{sub};
    std::vector<int> own_base(ncell + 1, 0), gh_base(ncell + 1, 0);
    std::vector<long> pair_base(ncell + 1, 0);
    const int xc = xcols(ctx);
    size_t xo = 0;
    for (int k = 0; k < ncell; ++k) {
        mapp_dmd_t* sp = nullptr;
        if (mapp_dmd_create(&sp, ctx->nt, ctx->variant, ctx->device) != 0 || !sp)
            throw ApiError(MAPP_DMD_E_HIP, "cannot create the context of a cell");
        sub[k] = sp;
        sp->P = ctx->P;
        sp->Pf = ctx->Pf;
        sp->mod_rc = ctx->mod_rc;
        sp->rsq_crd = ctx->rsq_crd;
        sp->have_params = true;
        const int n = ctx->cell_natms[k];
        if (mapp_dmd_set_system(sp, ctx->cell_H.data() + (size_t)k * 9, n, ctx->cell_x.data() + xo * xc,
                                ctx->cell_c.data() + xo * NT) != 0 ||
            mapp_dmd_init(sp, skin) != 0)
            throw ApiError(MAPP_DMD_E_ARG, "cell " + std::to_string(k) + ": " + sp->err);
        xo += n;
        own_base[k + 1] = own_base[k] + sp->natms;
        gh_base[k + 1] = gh_base[k] + sp->nghost;
        pair_base[k + 1] = pair_base[k] + sp->npairs;
    }
    const int natms = own_base[ncell], nghost = gh_base[ncell];
    require(pair_base[ncell] < (1L << 31), MAPP_DMD_E_RANGE, "batch exceeds 2^31 pairs");
    hipStream_t st = ctx->stream;
    ctx->natms = natms;
    ctx->nghost = nghost;
    ctx->npairs = pair_base[ncell];
    const int ntot = natms + nghost;
    ctx->d_pos.reserve(ntot);
    ctx->d_alpha.reserve((size_t)ntot * NT);
    ctx->d_conc.reserve((size_t)ntot * NT);
    ctx->d_ghost_owner.reserve(nghost + 1);
    ctx->d_off.reserve(natms + 1);
    ctx->d_nl.reserve(ctx->npairs + 1);
    for (int k = 0; k < ncell; ++k) {
        mapp_dmd* c = sub[k];
        const int n = c->natms, g = c->nghost;
        hipLaunchKernelGGL((k_merge_sites<NT>), dim3(nblk(n + g)), dim3(kBlock), 0, st, n, g, c->d_pos.p,
                           c->d_alpha.p, c->d_conc.p, c->d_ghost_owner.p, own_base[k], natms + gh_base[k], gh_base[k],
                           ctx->d_pos.p, ctx->d_alpha.p, ctx->d_conc.p, ctx->d_ghost_owner.p);
        launch_check("k_merge_sites");
        hipLaunchKernelGGL(k_merge_list, dim3(nblk(n)), dim3(kBlock), 0, st, n, c->d_off.p, c->d_nl.p, own_base[k],
                           (int)pair_base[k], natms + gh_base[k], ctx->d_off.p, ctx->d_nl.p);
        launch_check("k_merge_list");
    }
    const int last = (int)ctx->npairs;
    HIP_OK(hipMemcpyAsync(ctx->d_off.p + natms, &last, sizeof(int), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    ctx->swap_off.assign(1, 0);  // no box-wide ghost swaps: positions stay fixed in a batch
    ctx->swap_dim.clear();
    ctx->swap_dir.clear();
    finish_lists<NT>(ctx);
    for (int k = 0; k < ncell; ++k) {
        const int n = sub[k]->natms, g = sub[k]->nghost;
        hipLaunchKernelGGL(k_quirk_cell, dim3(nblk(n + g)), dim3(kBlock), 0, st, n, g, own_base[k], natms + gh_base[k],
                           NT, ctx->d_qidx.p);
        launch_check("k_quirk_cell");
    }
    HIP_OK(hipStreamSynchronize(st));
    ctx->cell_nghost.assign(ncell, 0);
    for (int k = 0; k < ncell; ++k) ctx->cell_nghost[k] = gh_base[k + 1] - gh_base[k];
}
// ---------------------------------------------------------- no-alloc calc_g
// Scratch of mapp_dmd_calc_g_noalloc: calc_g's per-block partials of the cost
// sweep over the 2nd list, the two-stage reduction's stage and its output, for
// systems of up to kNoAllocMaxSites owned sites (about 2.1 MB; each GPU the
// library runs on holds its own copy until the process ends).  Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU race.  Each call
// writes every element it reads (the partials of its own blocks, then the
// stages), so nothing has to start at zero.
constexpr long kNoAllocMaxSites = 1L << 24;
constexpr long kNoAllocParts = (kNoAllocMaxSites * kList2G + kBlock - 1) / kBlock;
constexpr long kNoAllocDoubles = kNoAllocParts + ((kNoAllocParts + kReduceChunk - 1) / kReduceChunk + 1) * 8 + 16;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

// the device address of scratch on each GPU, looked up once per device at
// context creation (the array's host-side name is not a device address)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Nothing else in the library uses the array, so a failed lookup does not fail
// context creation: the no-alloc entry point then reports it.
hipError_t g_scratch_err[kMaxDevices] = {};
void scratch_setup(int device) {  // device current
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_ptr[device] = static_cast<double*>(p);
    else (void)hipGetLastError();  // clear the sticky error for the context's other calls
}

// Scales of the current system that turn a relative force bound into per-pair
// tolerances (sur_mask).  One FP32-quadrature force_calc gives F = max|f| over
// all components, max|F'| and the energy; the per-site densities give the
// range over which max|F''| is taken from the embedding table.  A sweep over
// the half list (both ends of every pair inside the quadrature cut) gives N,
// the largest number of neighbours of a site inside the cut, and P_r, P_a,
// the largest per-site sums of |d rho_bar/dr| and |d rho_bar/dalpha| over
// those neighbours, each pair taking the largest value any type pair and
// width reaches in its r bin (prof_r, prof_a: over the training samples, 64
// bins of the trained box, widened to the neighbouring bins).
void sur_scales(mapp_dmd* ctx, const SurNorm& Nm, const std::vector<double>& prof_r,
                const std::vector<double>& prof_a) {
    const int nt = ctx->nt, natms = ctx->natms;
    const int prec0 = ctx->prec;
    ctx->prec = MAPP_DMD_PREC_MIXED;
    ctx->cache_valid = false;
    double en[7];
    force_any(ctx, 1, en);
    ctx->prec = prec0;
    ctx->cache_valid = false;
    const int xd = 3 + nt;
    const long ntot = (long)natms + ctx->nghost;
    std::vector<double> f((size_t)natms * xd), dE((size_t)natms * nt), Eown((size_t)natms * nt);
    std::vector<int> off(natms + 1), roff(natms + 1), nl((size_t)std::max(ctx->npairs, 1L));
    std::vector<double4> pos((size_t)ntot);
    HIP_OK(hipMemcpy(f.data(), ctx->d_f.p, f.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(dE.data(), ctx->d_dE.p, dE.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(Eown.data(), ctx->d_Eown.p, Eown.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(off.data(), ctx->d_off.p, off.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(roff.data(), ctx->d_roff.p, roff.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(pos.data(), ctx->d_pos.p, pos.size() * sizeof(double4), hipMemcpyDeviceToHost));
    if (ctx->npairs > 0)
        HIP_OK(hipMemcpy(nl.data(), ctx->d_nl.p, (size_t)ctx->npairs * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double> pairE((size_t)std::max(roff[natms], 1) * nt);
    if (roff[natms] > 0)
        HIP_OK(hipMemcpy(pairE.data(), ctx->d_pairE.p, (size_t)roff[natms] * nt * sizeof(double),
                         hipMemcpyDeviceToHost));
    double fmax = 0.0, dfmax = 0.0, rho_lo = 1e300, rho_hi = -1e300;
    for (double v : f) fmax = std::max(fmax, std::fabs(v));
    for (double v : dE) dfmax = std::max(dfmax, std::fabs(v));
    for (int i = 0; i < natms; ++i)
        for (int t = 0; t < nt; ++t) {
            double rho = Eown[(size_t)i * nt + t];
            for (int q = roff[i]; q < roff[i + 1]; ++q) rho += pairE[(size_t)q * nt + t];
            rho_lo = std::min(rho_lo, rho);
            rho_hi = std::max(rho_hi, rho);
        }
    // neighbour counts and per-site sums of |rho_bar'| inside the quadrature cut
    const int nb = (int)prof_r.size();
    const double pr_all = *std::max_element(prof_r.begin(), prof_r.end());
    const double pa_all = *std::max_element(prof_a.begin(), prof_a.end());
    std::vector<int> cnt(natms, 0);
    std::vector<double> sr(natms, 0.0), sa(natms, 0.0);
    for (int i = 0; i < natms; ++i)
        for (int q = off[i]; q < off[i + 1]; ++q) {
            const int j = nl[q];
            const double dx = pos[i].x - pos[j].x, dy = pos[i].y - pos[j].y, dz = pos[i].z - pos[j].z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < ctx->P.cut_sq_mod * (1.0 + 1e-12))) continue;  // an ulp of margin: an upper bound
            const double r = std::sqrt(r2);
            double vr = pr_all, va = pa_all;  // below the trained box: the largest values of any bin
            if (r >= Nm.r_lo) {
                int b = (int)((r - Nm.r_lo) / (Nm.r_hi - Nm.r_lo) * nb);
                b = std::min(b, nb - 1);
                vr = prof_r[b];
                va = prof_a[b];
            }
            ++cnt[i];
            sr[i] += vr;
            sa[i] += va;
            if (j < natms) {
                ++cnt[j];
                sr[j] += vr;
                sa[j] += va;
            }
        }
    int nmax = 0;
    double pr = 0.0, pa = 0.0;
    for (int i = 0; i < natms; ++i) {
        nmax = std::max(nmax, cnt[i]);
        pr = std::max(pr, sr[i]);
        pa = std::max(pa, sa[i]);
    }
    // |F''| = |c6 + 2 c7 p + 3 c8 p^2| / drho over the rows covering the densities
    // present (5% margin), at both row ends and the interior extremum
    const FFParams<double>& P = ctx->P;
    const double margin = 0.05 * (rho_hi - rho_lo) + 1e-12;
    const int m_lo = std::max(0, (int)((rho_lo - margin) * P.drho_inv));
    const int m_hi = std::min(P.nrho - 1, (int)((rho_hi + margin) * P.drho_inv) + 1);
    double ddf = 0.0;
    for (int a = 0; a < nt; ++a)
        for (int m = m_lo; m <= m_hi; ++m) {
            const double* c = ctx->F9.data() + ((size_t)a * (P.nrho + 1) + m) * 9;
            auto d2 = [&](double p) { return std::fabs(c[6] + 2.0 * c[7] * p + 3.0 * c[8] * p * p); };
            double v = std::max(d2(0.0), d2(1.0));
            if (c[8] != 0.0) {
                const double ps = -c[7] / (3.0 * c[8]);
                if (ps > 0.0 && ps < 1.0) v = std::max(v, d2(ps));
            }
            ddf = std::max(ddf, v * P.drho_inv);
        }
    double* S = ctx->sur_scale;
    S[0] = fmax;
    S[1] = dfmax;
    S[2] = ddf;
    S[3] = (double)nmax;
    S[4] = pr;
    S[5] = pa;
    S[6] = std::fabs(en[0]);
    S[7] = (double)natms;
}

// The per-pair tolerances of a bound tau, eps[o] = tau * unit[o] (sur_mask).
void sur_eps_unit(const mapp_dmd* ctx, double* e) {
    const double* S = ctx->sur_scale;
    const double F = S[0], dF = std::max(S[1], 1e-300), ddF = std::max(S[2], 1e-300), N = std::max(S[3], 1.0);
    const double Pd = std::max(std::max(S[4], S[5]), 1e-300), U = S[6], na = std::max(S[7], 1.0);
    e[0] = std::min(F / (6.0 * N * ddF * Pd), U / (2.0 * na * N * dF));
    e[1] = F / (6.0 * N * dF);
    e[2] = F / (6.0 * N * dF);
    e[3] = U / (na * N);
    e[4] = F / (3.0 * N);
    e[5] = F / (3.0 * N);
}

// Smallest bound at which each cell ([tp][cell]) is served: every validation
// sample of the cell within half of each tolerance, i.e. tau >= 2 max err/unit;
// infinity for a cell without samples.
std::vector<double> sur_cell_bounds(const mapp_dmd* ctx) {
    const int cells = kCellR * kCellA * kCellA;
    double unit[6];
    sur_eps_unit(ctx, unit);
    std::vector<double> tc((size_t)4 * cells, 0.0);
    std::vector<int> count((size_t)4 * cells, 0);
    const long nv = ctx->sur_nval;
    for (long k = 0; k < 4 * nv; ++k) {
        const int tp = (int)(k / nv);
        const size_t c = (size_t)tp * cells + ctx->h_scell[k];
        for (int o = 0; o < kOut; ++o)
            tc[c] = std::max(tc[c], 2.0 * (double)ctx->h_sjudge[(size_t)k * kOut + o] / unit[o]);
        ++count[c];
    }
    for (size_t c = 0; c < tc.size(); ++c)
        if (count[c] == 0) tc[c] = std::numeric_limits<double>::infinity();
    return tc;
}

// Fallback mask for a relative force bound tau (max |f_net - f| / max |f|):
// first-order propagation of per-pair output errors through the force
// (pass 2: every ordered type pair weighted by c_i c_j, sum 1; the density
// terms by F' at both ends; a density error of every neighbour moves F' by F''
// at a site, and F' multiplies the site's density derivatives at both ends of
// its pairs) with the triangle inequality, the budget split in thirds, gives
//   eps(phi_r) = tau F / (3N),   eps(rho_r) = tau F / (6 N F'),
//   eps(phi_a) = tau F / (3N),   eps(rho_a) = tau F / (6 N F')   (d alpha_pair / d alpha_i <= 1),
//   eps(rho)   = tau F / (6 N F'' max(P_r, P_a)),
//   eps(phi)   = tau |U| / (natms N)   (energy, N/2 pairs per site),
// and eps(rho) is also held to tau |U| / (2 natms N F') for the energy
// (scales: sur_scales).  A cell is served when it holds validation samples and
// every sample's error is within half of each tolerance; tau < 0 masks every
// cell, tau >= 1e30 serves every cell.
void sur_mask(mapp_dmd* ctx, double tau) {
    const int cells = kCellR * kCellA * kCellA;
    double* e = ctx->sur_eps;
    sur_eps_unit(ctx, e);
    for (int o = 0; o < 6; ++o) e[o] *= tau;
    std::vector<unsigned> mask((size_t)4 * cells / 32, 0u);
    long served = 0;
    if (tau >= 1e30) {
        served = 4L * cells;
    } else {
        const std::vector<double> tc = sur_cell_bounds(ctx);
        for (size_t c = 0; c < tc.size(); ++c)
            if (tau >= 0.0 && tc[c] <= tau) ++served;
            else mask[c / 32] |= 1u << (c % 32);
    }
    ctx->d_smask.reserve(mask.size());
    HIP_OK(hipMemcpy(ctx->d_smask.p, mask.data(), mask.size() * sizeof(unsigned), hipMemcpyHostToDevice));
    ctx->sur_bound = tau;
    ctx->sur_served = (double)served / (4.0 * cells);
}

}  // namespace

// =================================================================== C ABI
extern "C" {

int mapp_dmd_create(mapp_dmd_t** out, int no_types, int variant, int device) {
    if (!out) return MAPP_DMD_E_ARG;
    *out = nullptr;
    if (variant < 0 || variant > MAPP_DMD_VARIANT_MEAM) return MAPP_DMD_E_ARG;
    // DMD variants and the MD potentials other than plain EAM take up to kMaxTypes types
    if (no_types < 1 || no_types > (variant == MAPP_DMD_VARIANT_EAM ? 2 : kMaxTypes)) return MAPP_DMD_E_RANGE;
    std::unique_ptr<mapp_dmd> ctx(new mapp_dmd());
    ctx->nt = no_types;
    ctx->variant = variant;
    ctx->device = device;
    int rc = guarded(ctx.get(), [&]() {
        int ndev = 0;
        HIP_OK(hipGetDeviceCount(&ndev));
        require(device >= 0 && device < ndev, MAPP_DMD_E_ARG, "no such HIP device");
        HIP_OK(hipSetDevice(device));
        scratch_setup(device);
        HIP_OK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        for (auto& e : ctx->ev) HIP_OK(hipEventCreate(&e));
        ctx->d_iflag.reserve(4);
        return 0;
    });
    if (rc != 0) return rc;
    *out = ctx.release();
    return 0;
}

int mapp_dmd_destroy(mapp_dmd_t* ctx) {
    if (!ctx) return MAPP_DMD_E_ARG;
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return 0;
}

const char* mapp_dmd_last_error(const mapp_dmd_t* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mapp_dmd_set_fs(mapp_dmd_t* ctx, int nrho, double drho, int nr, double dr, const double* F, const double* rho,
                    const double* rphi) {
    return guarded(ctx, [&]() {
        require(F && rho && rphi, MAPP_DMD_E_ARG, "null table");
        require(nr >= 5 && nrho >= 5, MAPP_DMD_E_RANGE, "tables need at least 5 rows");
        require(nr == nrho, MAPP_DMD_E_RANGE, "nr must equal nrho (F rows are clamped with nr-2)");
        require(dr > 0 && drho > 0, MAPP_DMD_E_ARG, "non-positive grid spacing");
        HIP_OK(hipSetDevice(ctx->device));
        const int nt = ctx->nt, nphi = nt * (nt + 1) / 2;
        ctx->nr = nr;
        ctx->nrho = nrho;
        ctx->dr = dr;
        ctx->drho = drho;
        ctx->F9.assign((size_t)nt * (nrho + 1) * 9, 0.0);
        ctx->rho7.assign((size_t)nt * nt * nr * 7, 0.0);
        ctx->phi7.assign((size_t)nphi * nr * 7, 0.0);
        for (int a = 0; a < nt; ++a) {
            for (int i = 0; i < nrho; ++i) ctx->F9[((size_t)a * (nrho + 1) + i) * 9] = F[(size_t)a * nrho + i];
            for (int b = 0; b < nt; ++b) {
                int slot = type2rho(nt, a, b);
                for (int i = 0; i < nr; ++i)
                    ctx->rho7[((size_t)slot * nr + i) * 7] = rho[((size_t)a * nt + b) * nr + i];
            }
        }
        size_t k = 0;
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j <= i; ++j, ++k) {
                int slot = type2phi(i, j);
                for (int r = 0; r < nr; ++r) ctx->phi7[((size_t)slot * nr + r) * 7] = rphi[k * nr + r];
            }
        for (int a = 0; a < nt; ++a) spline_rows(nrho, drho, ctx->F9.data() + (size_t)a * (nrho + 1) * 9, 9, true);
        // ForceField_eam::set_arrays uses the cubic rows for F (src/ff_eam.cpp:1297-1321)
        ctx->F7.assign((size_t)nt * (nrho + 1) * 7, 0.0);
        for (int a = 0; a < nt; ++a) {
            for (int i = 0; i < nrho; ++i) ctx->F7[((size_t)a * (nrho + 1) + i) * 7] = F[(size_t)a * nrho + i];
            spline_rows(nrho, drho, ctx->F7.data() + (size_t)a * (nrho + 1) * 7, 7, false);
        }
        for (int s = 0; s < nt * nt; ++s) spline_rows(nr, dr, ctx->rho7.data() + (size_t)s * nr * 7, 7, false);
        for (int s = 0; s < nphi; ++s) spline_rows(nr, dr, ctx->phi7.data() + (size_t)s * nr * 7, 7, false);
        ctx->P.nr = nr;
        ctx->P.nrho = nrho;
        ctx->P.dr_inv = 1.0 / dr;
        ctx->P.drho_inv = 1.0 / drho;
        upload_tables(ctx);  // after F7 is built
        ctx->have_tables = true;
        ctx->have_params = false;
        return 0;
    });
}

int mapp_dmd_read_fs(mapp_dmd_t* ctx, const char* path, const char* const* elements) {
    return guarded(ctx, [&]() {
        require(path && elements, MAPP_DMD_E_ARG, "null path or element names");
        std::vector<std::string> names;
        for (int t = 0; t < ctx->nt; ++t) {
            require(elements[t] != nullptr, MAPP_DMD_E_ARG, "null element name");
            names.emplace_back(elements[t]);
        }
        FsData fs;
        std::string e = read_fs_file(path, names, fs);
        if (!e.empty()) throw ApiError(MAPP_DMD_E_IO, e);
        return mapp_dmd_set_fs(ctx, fs.nrho, fs.drho, fs.nr, fs.dr, fs.F.data(), fs.rho.data(), fs.rphi.data());
    });
}

int mapp_dmd_lj_coef(mapp_dmd_t* ctx, int itype, int jtype, double epsilon, double sigma, double cut) {
    return guarded(ctx, [&]() {
        // ForceField_lj::coef, src/ff_lj.cpp:72-125
        require(ctx->variant == MAPP_DMD_VARIANT_LJ, MAPP_DMD_E_STATE, "lj_coef needs the LJ variant");
        require(itype >= 0 && itype < ctx->nt && jtype >= 0 && jtype < ctx->nt, MAPP_DMD_E_ARG, "no such type");
        require(epsilon >= 0.0, MAPP_DMD_E_ARG, "LJ Epsilon cannot be smaller than zero");
        require(sigma >= 0.0, MAPP_DMD_E_ARG, "LJ Sigma cannot be smaller than zero");
        require(cut >= 0.0, MAPP_DMD_E_ARG, "LJ cutoff cannot be smaller than zero");
        MdParams& M = ctx->md;
        const int k = md_comp(itype, jtype);
        M.nt = ctx->nt;
        M.eps[k] = epsilon;
        M.sig[k] = sigma;
        M.lj_cut[k] = cut;
        M.offset[k] = 0.0;
        M.cut_sq[k] = cut * cut;
        M.chk[k] = (cut == 0.0 || sigma == 0.0 || epsilon == 0.0) ? 0 : 1;
        M.set[k] = 1;
        ctx->have_params = true;
        ctx->inited = false;
        return 0;
    });
}

int mapp_dmd_fs_coef(mapp_dmd_t* ctx, const double* A, const double* t1, const double* t2, const double* cut_rho,
                     const double* cut_phi, const double* k1, const double* k2, const double* k3) {
    return guarded(ctx, [&]() {
        // ForceField_fs::coef (src/ff_fs.cpp:113-335); cut_sq at init (:346).
        // ForceField_fsm::coef (src/ff_fsm.cpp:281-282) keeps the cutoffs
        // squared and takes sqrt(cut^2) as the cutoff distance (:424-426)
        const bool fsm = ctx->variant == MAPP_DMD_VARIANT_FSM;
        require(ctx->variant == MAPP_DMD_VARIANT_FS || fsm, MAPP_DMD_E_STATE, "fs_coef needs the FS or FSM variant");
        require(A && t1 && t2 && cut_rho && cut_phi && k1 && k2 && k3, MAPP_DMD_E_ARG, "null coefficients");
        MdParams& M = ctx->md;
        const int nt = ctx->nt, nc = nt * (nt + 1) / 2;
        M.nt = nt;
        for (int i = 0; i < nt; ++i) {
            M.A[i] = A[i];
            for (int j = 0; j < nt; ++j) {
                M.t1[i][j] = t1[i * nt + j];
                M.t2[i][j] = t2[i * nt + j];
            }
        }
        for (int k = 0; k < nc; ++k) {
            M.cut_sq_rho[k] = cut_rho[k] * cut_rho[k];
            M.cut_sq_phi[k] = cut_phi[k] * cut_phi[k];
            M.cut_rho[k] = fsm ? std::sqrt(M.cut_sq_rho[k]) : cut_rho[k];
            M.cut_phi[k] = fsm ? std::sqrt(M.cut_sq_phi[k]) : cut_phi[k];
            M.k1[k] = k1[k];
            M.k2[k] = k2[k];
            M.k3[k] = k3[k];
            M.cut_sq[k] = std::max(M.cut_sq_phi[k], M.cut_sq_rho[k]);
            M.chk[k] = 1;
            M.set[k] = 1;
        }
        ctx->have_params = true;
        ctx->inited = false;
        return 0;
    });
}

int mapp_dmd_fs_params_read(const char* path, int nt, const char* const* elements, double* A, double* t1, double* t2,
                            double* cut_rho, double* cut_phi, double* k1, double* k2, double* k3, char* err,
                            int err_cap) {
    return guarded_noctx(err, err_cap, [&]() -> int {
        auto fail = [&](const std::string& m, int code) {
            if (err && err_cap > 0) {
                std::strncpy(err, m.c_str(), (size_t)err_cap - 1);
                err[err_cap - 1] = 0;
            }
            return code;
        };
        if (!path || !elements || nt < 1 || nt > kMaxTypes) return fail("bad arguments", MAPP_DMD_E_ARG);
        std::vector<std::string> names(elements, elements + nt);
        FsParams fp;
        const std::string e = read_fs_params(path, names, fp);
        if (!e.empty()) return fail(e, MAPP_DMD_E_IO);
        const int nc = nt * (nt + 1) / 2;
        std::copy(fp.A.begin(), fp.A.end(), A);
        std::copy(fp.t1.begin(), fp.t1.end(), t1);
        std::copy(fp.t2.begin(), fp.t2.end(), t2);
        for (int k = 0; k < nc; ++k) {
            cut_rho[k] = fp.cut_rho[k];
            cut_phi[k] = fp.cut_phi[k];
            k1[k] = fp.k1[k];
            k2[k] = fp.k2[k];
            k3[k] = fp.k3[k];
        }
        return 0;
    });
}

int mapp_dmd_read_fs_params(mapp_dmd_t* ctx, const char* path, const char* const* elements) {
    return guarded(ctx, [&]() {
        require(path && elements, MAPP_DMD_E_ARG, "null path or elements");
        std::vector<std::string> names(elements, elements + ctx->nt);
        FsParams fp;
        const std::string e = read_fs_params(path, names, fp);
        if (!e.empty()) throw ApiError(MAPP_DMD_E_IO, e);
        return mapp_dmd_fs_coef(ctx, fp.A.data(), fp.t1.data(), fp.t2.data(), fp.cut_rho.data(), fp.cut_phi.data(),
                                fp.k1.data(), fp.k2.data(), fp.k3.data());
    });
}

int mapp_dmd_meam_params_read(const char* library, const char* local, int nt, const char* const* elements,
                              mapp_meam_params* p, char* err, int err_cap) {
    return guarded_noctx(err, err_cap, [&]() -> int {
        auto fail = [&](const std::string& m, int code) {
            if (err && err_cap > 0) {
                std::strncpy(err, m.c_str(), (size_t)err_cap - 1);
                err[err_cap - 1] = 0;
            }
            return code;
        };
        if (!library || !elements || !p || nt < 1 || nt > kMeamT) return fail("bad arguments", MAPP_DMD_E_ARG);
        std::vector<int> type_ref;
        std::string e = meam_host::read_library(library, nt, elements, p, &type_ref);
        if (e.empty() && local) e = meam_host::read_local(local, type_ref, p);
        if (!e.empty()) return fail(e, MAPP_DMD_E_IO);
        return 0;
    });
}

int mapp_dmd_meam_coef(mapp_dmd_t* ctx, const mapp_meam_params* p) {
    return guarded(ctx, [&]() {
        // ForceField_meam::coef after its readers (src/ff_meam.cpp:2593-2627)
        require(ctx->variant == MAPP_DMD_VARIANT_MEAM, MAPP_DMD_E_STATE, "meam_coef needs the MEAM variant");
        require(p != nullptr && p->nt == ctx->nt, MAPP_DMD_E_ARG, "parameters for another number of types");
        meam_host::Setup su;
        const std::string e = su.build(*p);
        if (!e.empty()) throw ApiError(MAPP_DMD_E_ARG, e);
        HIP_OK(hipSetDevice(ctx->device));
        ctx->d_mtab.reserve(su.table.size());
        HIP_OK(hipMemcpy(ctx->d_mtab.p, su.table.data(), sizeof(double) * su.table.size(), hipMemcpyHostToDevice));
        ctx->meam = std::move(su);
        ctx->have_params = true;
        ctx->inited = false;
        return 0;
    });
}

int mapp_dmd_read_meam_params(mapp_dmd_t* ctx, const char* library, const char* local,
                              const char* const* elements) {
    return guarded(ctx, [&]() {
        require(library && elements, MAPP_DMD_E_ARG, "null library or elements");
        mapp_meam_params p;
        char err[256] = {0};
        const int rc = mapp_dmd_meam_params_read(library, local, ctx->nt, elements, &p, err, sizeof err);
        if (rc != 0) throw ApiError(rc, err);
        return mapp_dmd_meam_coef(ctx, &p);
    });
}

int mapp_dmd_meam_set_fix(mapp_dmd_t* ctx, int fix) {
    return guarded(ctx, [&]() {
        require(ctx->variant == MAPP_DMD_VARIANT_MEAM, MAPP_DMD_E_STATE, "meam_set_fix needs the MEAM variant");
        ctx->meam_fix = fix ? 1 : 0;
        return 0;
    });
}

int mapp_dmd_coef(mapp_dmd_t* ctx, double r_crd, int no_i, double alpha_min, double alpha_max, double T,
                  const double* mass) {
    return guarded(ctx, [&]() {
        require(ctx->variant < MAPP_DMD_VARIANT_LJ, MAPP_DMD_E_STATE,
                "LJ, FS and FSM take mapp_dmd_lj_coef / mapp_dmd_fs_coef / mapp_dmd_read_fs_params");
        require(ctx->have_tables, MAPP_DMD_E_STATE, "tables must be set before coef");
        if (ctx->variant == MAPP_DMD_VARIANT_EAM) {
            // ForceField_eam::coef, src/ff_eam.cpp:424-425
            ctx->P.cut_sq = (ctx->nr - 1.0) * (ctx->nr - 1.0) * ctx->dr * ctx->dr;
            ctx->P.rho_max = (ctx->nrho - 1.0) * ctx->drho;
            ctx->have_params = true;
            return 0;
        }
        require(no_i >= 1 && no_i <= 14, MAPP_DMD_E_RANGE, "n should be between 1 and 14");
        require(alpha_min != 0.0, MAPP_DMD_E_ARG, "minimum alpha cannot be zero");
        require(mass != nullptr, MAPP_DMD_E_ARG, "null masses");
        FFParams<double>& P = ctx->P;
        ctx->rsq_crd = r_crd;
        ctx->rsq_crd *= ctx->rsq_crd;
        P.no_i = no_i;
        for (int k = 0; k < kMaxPoints; ++k) P.xi[k] = P.w0[k] = P.w1[k] = P.w2[k] = 0.0;
        for (int k = 0; k < no_i; ++k) {
            P.xi[k] = GH_NODES[no_i - 1][k];
            P.w0[k] = GH_WEIGHTS[no_i - 1][k];
        }
        for (int k = 0; k < no_i; ++k) {
            P.w1[k] = P.w0[k] * P.xi[k];
            P.w2[k] = P.w0[k] * P.xi[k] * P.xi[k];
        }
        P.alpha_min = alpha_min;
        P.alpha_max = alpha_max;
        P.inv_amin = 1.0 / alpha_min;
        P.inv_amax = 1.0 / alpha_max;
        const double kb = 8.617332478e-5, hbar = 6.5821192815e-16;
        for (int t = 0; t < kMaxTypes; ++t) P.c0[t] = 0.0;
        for (int t = 0; t < ctx->nt; ++t) {
            double m = mass[t];
            m *= 1.0364269184093291236e-28;
            double deb_l = hbar * hbar * 2.0 / (m * kb * T);
            P.c0[t] = 1.5 * kb * T * (std::log(deb_l) - 1.0);
        }
        P.kbT = kb * T;
        P.xw = ctx->energy_full ? 2.0 : 1.0;
        P.beta = 1.0 / P.kbT;
        P.rc = ((double)ctx->nr - 1.0) * ctx->dr;
        P.rho_max = ((double)ctx->nrho - 1.0) * ctx->drho;
        P.cut_sq = P.rc * P.rc;
        ctx->mod_rc = P.rc + P.xi[no_i - 1] / std::sqrt(alpha_min);
        P.cut_sq_mod = ctx->mod_rc * ctx->mod_rc;
        fill_float_params(ctx);
        ctx->have_params = true;
        ctx->cache_valid = false;
        return 0;
    });
}

int mapp_dmd_set_precision(mapp_dmd_t* ctx, int mode) {
    return guarded(ctx, [&]() {
        require(mode == MAPP_DMD_PREC_FP64 || mode == MAPP_DMD_PREC_MIXED || mode == MAPP_DMD_PREC_SURROGATE,
                MAPP_DMD_E_ARG, "unknown precision mode");
        require(mode != MAPP_DMD_PREC_SURROGATE || (ctx->sur_trained && ctx->nt == 2), MAPP_DMD_E_STATE,
                "surrogate mode needs a trained network (mapp_dmd_surrogate_train) and two types");
        ctx->prec = mode;
        ctx->cache_valid = false;
        return 0;
    });
}

int mapp_dmd_set_system(mapp_dmd_t* ctx, const double* H, int natms, const double* x, const double* c) {
    return guarded(ctx, [&]() {
        ctx->cell_natms.clear();
        require(H && x && c && natms > 0, MAPP_DMD_E_ARG, "null input or empty system");
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 3; ++j)
                require(H[i * 3 + j] == 0.0, MAPP_DMD_E_ARG, "H must be lower triangular");
        const int nt = acols(ctx), xc = xcols(ctx);
        BoxParams& bp = ctx->bp;
        std::memcpy(bp.H, H, sizeof(bp.H));
        // M3INV, src/init.h:72-85
        double(*A)[3] = bp.H;
        double(*B)[3] = bp.B;
        B[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        B[1][1] = A[2][2] * A[0][0] - A[2][0] * A[0][2];
        B[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        B[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
        B[2][1] = A[2][0] * A[0][1] - A[2][1] * A[0][0];
        B[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
        B[2][0] = A[1][0] * A[2][1] - A[2][0] * A[1][1];
        B[0][1] = A[2][1] * A[0][2] - A[0][1] * A[2][2];
        B[1][2] = A[0][2] * A[1][0] - A[1][2] * A[0][0];
        double det = A[0][0] * B[0][0] + A[0][1] * B[1][0] + A[0][2] * B[2][0];
        require(det != 0.0, MAPP_DMD_E_ARG, "singular box");
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) B[i][j] /= det;
        ctx->natms = natms;
        ctx->h_s.resize((size_t)natms * 4);
        ctx->h_alpha.resize((size_t)natms * nt);
        ctx->h_c.assign(c, c + (size_t)natms * nt);
        for (int i = 0; i < natms; ++i) {
            const double* r = x + (size_t)i * xc;
            for (int d = 0; d < 3; ++d) {
                require(r[d] >= 0.0 && r[d] < 1.0, MAPP_DMD_E_RANGE, "reduced coordinates must lie in [0,1)");
                ctx->h_s[(size_t)i * 4 + d] = r[d];
            }
            ctx->h_s[(size_t)i * 4 + 3] = 0.0;
            for (int t = 0; t < nt; ++t) ctx->h_alpha[(size_t)i * nt + t] = xc > 3 ? r[3 + t] : 1.0;
        }
        if (is_md(ctx))
            for (int i = 0; i < natms; ++i)
                require(c[i] >= 0.0 && c[i] < ctx->nt && c[i] == (double)(int)c[i], MAPP_DMD_E_ARG,
                        "MD-mode types must be integers in [0, nt)");
        ctx->have_system = true;
        ctx->inited = false;
        return 0;
    });
}

int mapp_dmd_init(mapp_dmd_t* ctx, double skin) {
    return guarded(ctx, [&]() {
        require(ctx->have_params && ctx->have_system, MAPP_DMD_E_STATE, "coef and set_system must precede init");
        require(skin >= 0.0, MAPP_DMD_E_ARG, "negative skin");
        HIP_OK(hipSetDevice(ctx->device));
        const int nt = acols(ctx), natms = ctx->natms;
        ctx->skin = skin;
        const FFParams<double>& P = ctx->P;
        ctx->bp.full = 0;
        if (ctx->variant == MAPP_DMD_VARIANT_MEAM) {
            // ForceField_meam::init (src/ff_meam.cpp:654-681): rc_meam + skin for
            // every type pair, ghosts for rc_meam, the full list (pair_wise = 0)
            const double rc = ctx->meam.p.rc;
            ctx->bp.pairwise = 0;
            ctx->bp.full = 1;
            ctx->bp.cut_sk_sq = rc * rc + skin * skin + 2 * std::sqrt(rc * rc) * skin;
            set_ph(ctx, rc);
        } else if (ctx->variant >= MAPP_DMD_VARIANT_LJ) {
            // ForceField_lj::init (src/ff_lj.cpp:129-153), ForceField_fs::init
            // (src/ff_fs.cpp:339-360): a cutoff+skin per type pair with coefficients,
            // ghosts for the largest cutoff
            MdParams& M = ctx->md;
            double ph_cut = 0.0;
            ctx->bp.pairwise = 1;
            ctx->bp.cut_sk_sq = 0.0;
            for (int k = 0; k < kMaxComp; ++k) {
                ctx->bp.cut_sk_sq_tp[k] = 0.0;
                if (M.chk[k]) {
                    ctx->bp.cut_sk_sq_tp[k] = M.cut_sq[k] + skin * skin + 2 * std::sqrt(M.cut_sq[k]) * skin;
                    ph_cut = std::max(ph_cut, std::sqrt(M.cut_sq[k]));
                } else if (M.set[k] && ctx->variant == MAPP_DMD_VARIANT_LJ) {
                    ctx->bp.cut_sk_sq_tp[k] = (M.lj_cut[k] + skin) * (M.lj_cut[k] + skin);  // coef-time value
                }
            }
            require(ph_cut > 0.0, MAPP_DMD_E_STATE, "no type pair has coefficients");
            set_ph(ctx, ph_cut);
        } else {
            // ForceField_DMD_I::init, src/ff_dmd_I.cpp:675-685; ForceField_eam::init, src/ff_eam.cpp:362-370
            const double base = ctx->variant == MAPP_DMD_VARIANT_EAM ? P.cut_sq : P.cut_sq_mod;
            ctx->bp.pairwise = 0;
            ctx->bp.cut_sk_sq = base + skin * skin + 2.0 * std::sqrt(base) * skin;
            set_ph(ctx, std::sqrt(base));
        }
        for (int d = 0; d < 3; ++d)
            require(ctx->bp.cut_ph_s[d] < 1.0, MAPP_DMD_E_RANGE,
                    "box thinner than cutoff+skin: needs more than one ghost swap per face");
        if (!ctx->cell_natms.empty()) {
            with_nt(nt, [&](auto c) { init_batch<decltype(c)::value>(ctx, skin); });
            sync_sizes(ctx);
            ctx->inited = true;
            return 0;
        }
        ctx->d_pos.reserve(natms);
        ctx->d_alpha.reserve((size_t)natms * nt);
        ctx->d_conc.reserve((size_t)natms * nt);
        HIP_OK(hipMemcpy(ctx->d_pos.p, ctx->h_s.data(), sizeof(double4) * natms, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(ctx->d_alpha.p, ctx->h_alpha.data(), sizeof(double) * natms * nt, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(ctx->d_conc.p, ctx->h_c.data(), sizeof(double) * natms * nt, hipMemcpyHostToDevice));
        rebuild_any(ctx);
        sync_sizes(ctx);
        ctx->inited = true;
        return 0;
    });
}

int mapp_dmd_fin(mapp_dmd_t* ctx) {
    return guarded(ctx, [&]() {
        ctx->inited = false;
        ctx->have_list2 = false;
        ctx->cache_valid = false;
        return 0;
    });
}

int mapp_dmd_set_x(mapp_dmd_t* ctx, const double* x_host, int* rebuilt) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede set_x");
        require(ctx->cell_natms.empty(), MAPP_DMD_E_STATE, "positions are fixed in a batch of cells");
        HIP_OK(hipSetDevice(ctx->device));
        const size_t n = (size_t)ctx->natms * xcols(ctx);
        ctx->d_xaos.reserve(n);
        if (x_host)
            HIP_OK(hipMemcpyAsync(ctx->d_xaos.p, x_host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        const int flag = update_x_device(ctx);
        if (rebuilt) *rebuilt = flag;
        return 0;
    });
}

int mapp_dmd_set_c(mapp_dmd_t* ctx, const double* c_host) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede set_c");
        HIP_OK(hipSetDevice(ctx->device));
        const size_t n = (size_t)ctx->natms * acols(ctx);
        if (c_host)
            HIP_OK(hipMemcpyAsync(ctx->d_conc.p, c_host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        ghost_copy(ctx, ctx->d_conc.p, acols(ctx));
        // the pair values of c_d_calc / calc_g depend on x and alpha only: they stay valid
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int mapp_dmd_force_calc(mapp_dmd_t* ctx, int st_clc, double* en_st) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede force_calc");
        require(en_st != nullptr, MAPP_DMD_E_ARG, "null en_st");
        HIP_OK(hipSetDevice(ctx->device));
        if (is_md(ctx)) run_eam(ctx, false, st_clc, en_st);
        else with_nt(ctx->nt, [&](auto c) { dispatch_force<decltype(c)::value>(ctx, st_clc, en_st); });
        return 0;
    });
}

int mapp_dmd_energy_calc(mapp_dmd_t* ctx, double* energy) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede energy_calc");
        require(energy != nullptr, MAPP_DMD_E_ARG, "null energy");
        HIP_OK(hipSetDevice(ctx->device));
        if (is_md(ctx)) *energy = run_eam(ctx, true, 0, nullptr);
        else *energy = with_nt(ctx->nt, [&](auto c) { return dispatch_energy<decltype(c)::value>(ctx); });
        return 0;
    });
}


int mapp_dmd_create_2nd_neigh_lst(mapp_dmd_t* ctx) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede create_2nd_neigh_lst");
        HIP_OK(hipSetDevice(ctx->device));
        with_nt(ctx->nt, [&](auto c) { build_list2<decltype(c)::value>(ctx); });
        return 0;
    });
}

int mapp_dmd_c_d_calc(mapp_dmd_t* ctx) {
    return guarded(ctx, [&]() {
        require(!is_md(ctx), MAPP_DMD_E_STATE, "c_d_calc is a DMD entry point");
        require(ctx->inited && ctx->have_list2, MAPP_DMD_E_STATE, "init and create_2nd_neigh_lst must precede c_d_calc");
        HIP_OK(hipSetDevice(ctx->device));
        with_nt(ctx->nt, [&](auto c) { c_d_variant<decltype(c)::value>(ctx); });
        return 0;
    });
}

int mapp_dmd_calc_g(mapp_dmd_t* ctx, int chk, double alpha, const double* a_dev, double* g_dev, double* cost) {
    return guarded(ctx, [&]() {
        require(!is_md(ctx), MAPP_DMD_E_STATE, "calc_g is a DMD entry point");
        require(ctx->inited && ctx->have_list2, MAPP_DMD_E_STATE, "init and create_2nd_neigh_lst must precede calc_g");
        require(cost != nullptr, MAPP_DMD_E_ARG, "null cost");
        HIP_OK(hipSetDevice(ctx->device));
        if (!a_dev) a_dev = ctx->d_a.p;  // the library's own buffers (mapp_dmd_set_vec / get_vec A, G)
        if (!g_dev) g_dev = ctx->d_g.p;
        *cost = with_nt(ctx->nt,
                        [&](auto c) { return calc_g_variant<decltype(c)::value>(ctx, chk, alpha, a_dev, g_dev); });
        return 0;
    });
}


int mapp_dmd_surrogate_train(mapp_dmd_t* ctx, long n_per_tp, int epochs, int batch_per_tp, double lr,
                             unsigned long long seed, double tol, const double* box, double* report) {
    return guarded(ctx, [&]() {
        require(ctx->have_params && ctx->nt == 2 && !is_md(ctx) && ctx->inited, MAPP_DMD_E_STATE,
                "the surrogate covers the two-type DMD force field of an initialised system");
        require(n_per_tp >= kMB && batch_per_tp >= kMB && batch_per_tp % kMB == 0 && n_per_tp >= batch_per_tp &&
                    epochs >= 1 && lr > 0,
                MAPP_DMD_E_ARG, "bad training sizes (batch a multiple of 64, samples >= batch)");
        HIP_OK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        SurNorm N{};
        N.r_lo = box ? (float)box[0] : 1.8f;
        N.r_hi = box ? (float)box[1] : (float)ctx->mod_rc;
        N.a_lo = box ? (float)box[2] : 60.0f;
        N.a_hi = box ? (float)box[3] : 120.0f;
        N.r_c = 0.5f * (N.r_lo + N.r_hi);
        N.r_s = 0.5f * (N.r_hi - N.r_lo);
        N.a_c = 0.5f * (N.a_lo + N.a_hi);
        N.a_s = 0.5f * (N.a_hi - N.a_lo);
        const long n_val = std::max<long>(n_per_tp / 8, 4096);
        sur_samples(ctx, N, n_per_tp, seed, ctx->d_sX, ctx->d_sT, ctx->d_sB);
        sur_samples(ctx, N, n_val, seed ^ 0x5bd1e995ull, ctx->d_sXv, ctx->d_sTv, ctx->d_sBv);
        // residual targets: (quadrature - baseline) / envelope, standardised per type pair
        const size_t nt_s = (size_t)n_per_tp * 4;
        std::vector<float> hT(nt_s * kOut), hB(nt_s * kOut), hX(nt_s * 3);
        HIP_OK(hipMemcpyAsync(hT.data(), ctx->d_sT.p, hT.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(hB.data(), ctx->d_sB.p, hB.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(hX.data(), ctx->d_sX.p, hX.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        // largest |d rho_bar/dr|, |d rho_bar/dalpha| per r bin (for the bound's per-site sums)
        constexpr int kProfBins = 64;
        std::vector<double> prof_r(kProfBins, 0.0), prof_a(kProfBins, 0.0);
        std::vector<double> emax((size_t)4 * kOut * kEnvBins, 0.0), gmax((size_t)4 * kOut, 0.0);
        for (size_t k = 0; k < nt_s; ++k) {
            const int tp = (int)(k / n_per_tp);
            const float r = hX[k * 3];
            int b = (int)((r - N.r_lo) / (N.r_hi - N.r_lo) * kEnvBins);
            b = std::min(std::max(b, 0), kEnvBins - 1);
            int pb = (int)((r - N.r_lo) / (N.r_hi - N.r_lo) * kProfBins);
            pb = std::min(std::max(pb, 0), kProfBins - 1);
            prof_r[pb] = std::max(prof_r[pb], (double)std::fabs(hT[k * kOut + 1]));
            prof_a[pb] = std::max(prof_a[pb], (double)std::fabs(hT[k * kOut + 2]));
            for (int o = 0; o < kOut; ++o) {
                const double res = std::fabs((double)hT[k * kOut + o] - (double)hB[k * kOut + o]);
                double& m = emax[((size_t)tp * kOut + o) * kEnvBins + b];
                m = std::max(m, res);
                gmax[(size_t)tp * kOut + o] = std::max(gmax[(size_t)tp * kOut + o], res);
            }
        }
        // envelope: the largest residual of the bin and its two neighbours, floored
        // at 1e-9 of the largest residual of the output
        std::vector<float> lenv((size_t)4 * kOut * kEnvBins);
        for (int tp = 0; tp < 4; ++tp)
            for (int o = 0; o < kOut; ++o) {
                const double* m = emax.data() + ((size_t)tp * kOut + o) * kEnvBins;
                const double floor_v = std::max(1e-9 * gmax[(size_t)tp * kOut + o], 1e-30);
                for (int b = 0; b < kEnvBins; ++b) {
                    double v = m[b];
                    if (b > 0) v = std::max(v, m[b - 1]);
                    if (b + 1 < kEnvBins) v = std::max(v, m[b + 1]);
                    lenv[((size_t)tp * kOut + o) * kEnvBins + b] = (float)std::log2(std::max(v, floor_v));
                }
            }
        ctx->d_senv.reserve(lenv.size());
        HIP_OK(hipMemcpy(ctx->d_senv.p, lenv.data(), lenv.size() * sizeof(float), hipMemcpyHostToDevice));
        for (size_t k = 0; k < nt_s; ++k) {
            const int tp = (int)(k / n_per_tp);
            for (int o = 0; o < kOut; ++o) {
                const float sc = std::exp2(sur_env_log2(lenv.data(), N, tp, o, hX[k * 3]));
                hT[k * kOut + o] = (hT[k * kOut + o] - hB[k * kOut + o]) / sc;
            }
        }
        for (int tp = 0; tp < 4; ++tp)
            for (int o = 0; o < kOut; ++o) {
                double m = 0, m2 = 0;
                for (long s2 = 0; s2 < n_per_tp; ++s2) {
                    const double v = hT[((size_t)tp * n_per_tp + s2) * kOut + o];
                    m += v;
                    m2 += v * v;
                }
                m /= n_per_tp;
                const double var = m2 / n_per_tp - m * m;
                N.y_mean[tp][o] = (float)m;
                N.y_std[tp][o] = var > 1e-30 ? (float)std::sqrt(var) : 1.0f;
            }
        HIP_OK(hipMemcpy(ctx->d_sT.p, hT.data(), hT.size() * sizeof(float), hipMemcpyHostToDevice));
        // Glorot-uniform weights, zero biases
        std::mt19937_64 rng(seed);
        std::vector<float> hp((size_t)4 * kNParam, 0.0f);
        auto fill = [&](float* w, int fin, int fout) {
            std::uniform_real_distribution<float> U(-std::sqrt(6.0f / (fin + fout)), std::sqrt(6.0f / (fin + fout)));
            for (int e = 0; e < fin * fout; ++e) w[e] = U(rng);
        };
        for (int tp = 0; tp < 4; ++tp) {
            float* P = hp.data() + (size_t)tp * kNParam;
            fill(P + kW1, kIn, kHid);
            fill(P + kW2, kHid, kHid);
            fill(P + kW3, kHid, kHid);
            fill(P + kW4, kHid, kOut);
        }
        ctx->d_sp.reserve(hp.size());
        ctx->d_sm.reserve(hp.size());
        ctx->d_sv.reserve(hp.size());
        HIP_OK(hipMemcpy(ctx->d_sp.p, hp.data(), hp.size() * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemset(ctx->d_sm.p, 0, hp.size() * sizeof(float)));
        HIP_OK(hipMemset(ctx->d_sv.p, 0, hp.size() * sizeof(float)));
        const int bpt = batch_per_tp / kMB;
        const long steps_per_epoch = n_per_tp / batch_per_tp;
        ctx->d_sgpart.reserve((size_t)4 * bpt * kNParam);
        ctx->d_slpart.reserve((size_t)4 * bpt);
        ctx->d_sidx.reserve((size_t)n_per_tp * 4);
        std::vector<int> perm(n_per_tp), idx((size_t)n_per_tp * 4);
        const long total = steps_per_epoch * epochs;
        long step = 0;
        double last_loss = 0.0;
        const auto t0 = std::chrono::steady_clock::now();
        for (int ep = 0; ep < epochs; ++ep) {
            // per type pair shuffle, laid out step-major: [step][tp][batch]
            for (int tp = 0; tp < 4; ++tp) {
                std::iota(perm.begin(), perm.end(), 0);
                std::shuffle(perm.begin(), perm.end(), rng);
                for (long s2 = 0; s2 < steps_per_epoch; ++s2)
                    for (int b = 0; b < batch_per_tp; ++b)
                        idx[((size_t)s2 * 4 + tp) * batch_per_tp + b] =
                            (int)((long)tp * n_per_tp + perm[(size_t)s2 * batch_per_tp + b]);
            }
            HIP_OK(hipMemcpyAsync(ctx->d_sidx.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, st));
            for (long s2 = 0; s2 < steps_per_epoch; ++s2, ++step) {
                sur_gradient(ctx, N, ctx->d_sp.p, ctx->d_sX.p, ctx->d_sT.p,
                             ctx->d_sidx.p + (size_t)s2 * 4 * batch_per_tp, batch_per_tp,
                             1.0f / (float)(batch_per_tp * kOut), ctx->d_sgpart.p, ctx->d_slpart.p);
                const double frac = (double)step / (double)std::max<long>(1, total - 1);
                const float lr_t = (float)(lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.14159265358979 * frac))));
                const float bc1 = 1.0f - std::pow(0.9f, (float)(step + 1));
                const float bc2 = 1.0f - std::pow(0.999f, (float)(step + 1));
                hipLaunchKernelGGL(k_sur_adam, dim3(nblk(4L * kNParam)), dim3(kBlock), 0, st, ctx->d_sp.p,
                                   ctx->d_sm.p, ctx->d_sv.p, ctx->d_sgpart.p, 4, bpt, lr_t, 0.9f, 0.999f, 1e-8f, bc1,
                                   bc2);
                launch_check("k_sur_adam");
            }
            std::vector<float> lp((size_t)4 * bpt);
            HIP_OK(hipMemcpyAsync(lp.data(), ctx->d_slpart.p, lp.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            last_loss = 0.0;
            for (float v : lp) last_loss += v;
            last_loss /= (double)(4 * batch_per_tp * kOut);
        }
        HIP_OK(hipStreamSynchronize(st));
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // the inference path judged against the quadrature on the validation samples
        ctx->d_sfrag.reserve((size_t)4 * kFragPerTp * 64 * 8);
        hipLaunchKernelGGL(k_sur_pack_frags, dim3(nblk(4L * kFragPerTp * 64)), dim3(kBlock), 0, st, ctx->d_sp.p, 4,
                           reinterpret_cast<__bf16*>(ctx->d_sfrag.p));
        launch_check("k_sur_pack_frags");
        ctx->d_serr.reserve((size_t)n_val * 4 * kOut);
        {
            SurDev S{ctx->d_sp.p, reinterpret_cast<const __bf16*>(ctx->d_sfrag.p), nullptr, ctx->d_senv.p, N};
            HIP_OK(hipFuncSetAttribute((const void*)k_sur_judge, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)kSurLds));
            hipLaunchKernelGGL(k_sur_judge, dim3(kSurBlocks), dim3(kSurBlock), kSurLds, st, ctx->Pf, S,
                               ctx->d_quad32.p, ctx->d_sXv.p, ctx->d_sTv.p, n_val, ctx->d_serr.p);
            launch_check("k_sur_judge");
        }
        ctx->h_sjudge.resize((size_t)n_val * 4 * kOut);
        std::vector<float> hXv((size_t)n_val * 4 * 3);
        HIP_OK(hipMemcpyAsync(ctx->h_sjudge.data(), ctx->d_serr.p, ctx->h_sjudge.size() * sizeof(float),
                              hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(hXv.data(), ctx->d_sXv.p, hXv.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        ctx->sur_nval = n_val;
        ctx->h_scell.resize((size_t)n_val * 4);
        // the networks' own error in standardised output units (|err| / (s(r) std))
        double emean = 0.0, emax_v = 0.0;
        for (long k = 0; k < n_val * 4; ++k) {
            const int tp = (int)(k / n_val);
            const float r = hXv[k * 3], ai = hXv[k * 3 + 1], aj = hXv[k * 3 + 2];
            int cr = (int)((r - N.r_lo) / (N.r_hi - N.r_lo) * kCellR);
            int ci = (int)((ai - N.a_lo) / (N.a_hi - N.a_lo) * kCellA);
            int cj = (int)((aj - N.a_lo) / (N.a_hi - N.a_lo) * kCellA);
            cr = std::min(std::max(cr, 0), kCellR - 1);
            ci = std::min(std::max(ci, 0), kCellA - 1);
            cj = std::min(std::max(cj, 0), kCellA - 1);
            ctx->h_scell[k] = (cr * kCellA + ci) * kCellA + cj;
            double e = 0.0;
            for (int o = 0; o < kOut; ++o) {
                const double sc = std::exp2(sur_env_log2(lenv.data(), N, tp, o, r)) * N.y_std[tp][o];
                e = std::max(e, (double)ctx->h_sjudge[(size_t)k * kOut + o] / sc);
            }
            emean += e;
            emax_v = std::max(emax_v, e);
        }
        emean /= (double)(n_val * 4);
        {
            std::vector<double> wr(prof_r), wa(prof_a);
            for (int b = 0; b < kProfBins; ++b)
                for (int d = -1; d <= 1; ++d)
                    if (b + d >= 0 && b + d < kProfBins) {
                        wr[b] = std::max(wr[b], prof_r[b + d]);
                        wa[b] = std::max(wa[b], prof_a[b + d]);
                    }
            sur_scales(ctx, N, wr, wa);
        }
        sur_mask(ctx, tol);
        ctx->snorm = N;
        ctx->sur_trained = true;
        if (report) {
            report[0] = last_loss;
            report[1] = emean;
            report[2] = emax_v;
            report[3] = 1.0 - ctx->sur_served;
            report[4] = (double)total;
            report[5] = secs;
            report[6] = (double)(n_per_tp * 4);
            report[7] = (double)(n_val * 4);
        }
        return 0;
    });
}

int mapp_dmd_surrogate_gradient(mapp_dmd_t* ctx, const float* params, const float* norm, long n_per_tp,
                                const float* X, const float* T, float* grad, float* loss) {
    return guarded(ctx, [&]() {
        require(params && norm && X && T && grad && n_per_tp >= kMB && n_per_tp % kMB == 0, MAPP_DMD_E_ARG,
                "bad gradient request (samples per type pair a multiple of 64)");
        HIP_OK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        SurNorm N{};
        norm_from_flat(N, norm);
        DevBuf<float> p, x, t, gp, lp, g;
        DevBuf<int> idx;
        const long n = n_per_tp * 4;
        const int bpt = (int)(n_per_tp / kMB);
        p.reserve((size_t)4 * kNParam);
        x.reserve(n * 3);
        t.reserve(n * kOut);
        gp.reserve((size_t)4 * bpt * kNParam);
        lp.reserve((size_t)4 * bpt);
        g.reserve((size_t)4 * kNParam);
        idx.reserve(n);
        std::vector<int> id(n);
        std::iota(id.begin(), id.end(), 0);
        HIP_OK(hipMemcpy(p.p, params, sizeof(float) * 4 * kNParam, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(x.p, X, sizeof(float) * n * 3, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(t.p, T, sizeof(float) * n * kOut, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(idx.p, id.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        sur_gradient(ctx, N, p.p, x.p, t.p, idx.p, n_per_tp, 1.0f / (float)(n_per_tp * kOut), gp.p, lp.p);
        hipLaunchKernelGGL(k_sur_sum_grad, dim3(nblk(4L * kNParam)), dim3(kBlock), 0, st, gp.p, 4, bpt, g.p);
        launch_check("k_sur_sum_grad");
        std::vector<float> hl((size_t)4 * bpt);
        HIP_OK(hipMemcpyAsync(grad, g.p, sizeof(float) * 4 * kNParam, hipMemcpyDeviceToHost, st));
        HIP_OK(hipMemcpyAsync(hl.data(), lp.p, sizeof(float) * hl.size(), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (loss)
            for (int tp = 0; tp < 4; ++tp) {
                double acc = 0.0;
                for (int b = 0; b < bpt; ++b) acc += hl[(size_t)tp * bpt + b];
                loss[tp] = (float)(acc / (double)(n_per_tp * kOut));
            }
        return 0;
    });
}

int mapp_dmd_surrogate_params(mapp_dmd_t* ctx, float* params, float* norm) {
    return guarded(ctx, [&]() {
        require(ctx->sur_trained, MAPP_DMD_E_STATE, "no trained surrogate");
        if (params) HIP_OK(hipMemcpy(params, ctx->d_sp.p, sizeof(float) * 4 * kNParam, hipMemcpyDeviceToHost));
        if (norm) norm_to_flat(ctx->snorm, norm);
        return 0;
    });
}

int mapp_dmd_surrogate_fallbacks(mapp_dmd_t* ctx, long* count) {
    return guarded(ctx, [&]() {
        require(ctx->sur_trained && count, MAPP_DMD_E_STATE, "no trained surrogate");
        if (ctx->sur_last_quadrature) {  // no cell served: the mixed pass 1 ran, every evaluation fell back
            *count = ctx->n_inrange * ctx->nt * ctx->nt;
            return 0;
        }
        std::vector<int> h(ctx->natms);
        ctx->d_nfall.reserve(ctx->natms);
        HIP_OK(hipMemcpy(h.data(), ctx->d_nfall.p, sizeof(int) * ctx->natms, hipMemcpyDeviceToHost));
        long c = 0;
        for (int v : h) c += v;
        *count = c;
        return 0;
    });
}


int mapp_dmd_calc_g_noalloc(mapp_dmd_t* ctx, int chk, double alpha, const double* a_dev, double* g_dev,
                            double* cost) {
    return guarded(ctx, [&]() {
        require(!is_md(ctx), MAPP_DMD_E_STATE, "calc_g is a DMD entry point");
        require(ctx->inited && ctx->have_list2, MAPP_DMD_E_STATE, "init and create_2nd_neigh_lst must precede calc_g");
        require(ctx->cache_valid, MAPP_DMD_E_STATE,
                "stale pair values: call mapp_dmd_c_d_calc or mapp_dmd_calc_g once after the last change of "
                "positions or widths (building them allocates)");
        require(a_dev && g_dev && cost, MAPP_DMD_E_ARG, "a and g (device pointers) and cost must be given");
        require(ctx->natms <= kNoAllocMaxSites, MAPP_DMD_E_RANGE,
                "more owned sites than the no-alloc scratch holds (16,777,216)");
        HIP_OK(hipSetDevice(ctx->device));
        require(ctx->device < kMaxDevices, MAPP_DMD_E_RANGE, "device index beyond the scratch table");
        double* base = g_scratch_ptr[ctx->device];
        if (!base)
            throw ApiError(MAPP_DMD_E_HIP, std::string("no device address of the scratch array on this device: ") +
                                               hipGetErrorString(g_scratch_err[ctx->device]));
        const long nb2 = calc_g_blocks(ctx);
        const GScratch S{base, base + nb2, base + nb2 + reduce_stage_doubles(nb2)};
        *cost = with_nt(ctx->nt,
                        [&](auto c) { return calc_g_variant<decltype(c)::value>(ctx, chk, alpha, a_dev, g_dev, &S); });
        return 0;
    });
}

int mapp_dmd_surrogate_cell_bounds(mapp_dmd_t* ctx, double* out, long cap) {
    return guarded(ctx, [&]() {
        require(ctx->sur_trained, MAPP_DMD_E_STATE, "no trained surrogate");
        const std::vector<double> tc = sur_cell_bounds(ctx);
        require(out && cap >= (long)tc.size(), MAPP_DMD_E_ARG, "out needs 4 * 64 * 4 * 4 doubles");
        std::copy(tc.begin(), tc.end(), out);
        return 0;
    });
}

int mapp_dmd_surrogate_set_bound(mapp_dmd_t* ctx, double bound, double* report) {
    return guarded(ctx, [&]() {
        require(ctx->sur_trained, MAPP_DMD_E_STATE, "no trained surrogate");
        HIP_OK(hipSetDevice(ctx->device));
        sur_mask(ctx, bound);
        if (report) {
            report[0] = ctx->sur_served;
            for (int o = 0; o < 6; ++o) report[1 + o] = ctx->sur_eps[o];
            report[7] = bound;
            for (int k = 0; k < 8; ++k) report[8 + k] = ctx->sur_scale[k];
        }
        return 0;
    });
}

int mapp_dmd_set_energy_mode(mapp_dmd_t* ctx, int full) {
    return guarded(ctx, [&]() {
        ctx->energy_full = full != 0;
        ctx->P.xw = ctx->energy_full ? 2.0 : 1.0;
        ctx->Pf.xw = (float)ctx->P.xw;
        return 0;
    });
}

int mapp_dmd_min_lbfgs(mapp_dmd_t* ctx, int max_iter, int m_it, double e_tol, double* energies, long* counts) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede min_lbfgs");
        require(ctx->cell_natms.empty(), MAPP_DMD_E_STATE, "positions are fixed in a batch of cells");
        require(max_iter >= 0 && m_it >= 0 && energies, MAPP_DMD_E_ARG, "bad min_lbfgs arguments");
        HIP_OK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const long n = (long)ctx->natms * xcols(ctx);
        const int mm = std::max(m_it, 1);
        ctx->d_lh.reserve(n);
        ctx->d_lx0.reserve(n);
        ctx->d_lf0.reserve(n);
        ctx->d_lS.reserve((size_t)n * mm);
        ctx->d_lY.reserve((size_t)n * mm);
        ctx->d_xaos.reserve(n);
        std::vector<int> sl(mm);
        std::vector<double> rh(mm, 0.0), al(mm, 0.0);
        for (int i = 0; i < mm; ++i) sl[i] = i;
        auto Sv = [&](int i) { return ctx->d_lS.p + (size_t)sl[i] * n; };
        auto Yv = [&](int i) { return ctx->d_lY.p + (size_t)sl[i] * n; };
        long n_force = 0, n_energy = 0;
        double en[7];
        force_any(ctx, 1, en);
        ++n_force;
        double curr = en[0], prev = 0.0, ratio = 1.0, a_m = 0.0;
        energies[0] = curr;
        int k_it = 0, istp = 0, err = 0;
        const dim3 g(nblk(n)), b(kBlock);
        while (err == 0) {
            // h = f, x0 = x, f0 = f
            gather_x_any(ctx, ctx->d_xaos.p, ctx->natms, ctx->d_pos.p, ctx->d_alpha.p);
            HIP_OK(hipMemcpyAsync(ctx->d_lh.p, ctx->d_f.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIP_OK(hipMemcpyAsync(ctx->d_lx0.p, ctx->d_xaos.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIP_OK(hipMemcpyAsync(ctx->d_lf0.p, ctx->d_f.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            for (int i = k_it - 1; i > -1; --i) {
                al[i] = dot1(ctx, n, ctx->d_lh.p, Sv(i)) * rh[i];
                hipLaunchKernelGGL(k_axpby, g, b, 0, st, n, -al[i], Yv(i), 1.0, ctx->d_lh.p);
            }
            hipLaunchKernelGGL(k_axpby, g, b, 0, st, n, 0.0, ctx->d_lh.p, ratio, ctx->d_lh.p);
            for (int i = 0; i < k_it; ++i) {
                const double be = dot1(ctx, n, ctx->d_lh.p, Yv(i)) * rh[i];
                hipLaunchKernelGGL(k_axpby, g, b, 0, st, n, -(be + al[i]), Sv(i), 1.0, ctx->d_lh.p);
            }
            launch_check("k_axpby");
            prev = curr;
            err = line_min(ctx, n, &curr, &a_m, &n_energy);
            if (err == 0 && prev - curr < e_tol) err = 4;
            if (err == 0 && istp == max_iter) err = 5;
            force_any(ctx, 1, en);
            ++n_force;
            curr = en[0];
            energies[istp + 1] = curr;
            if (err) break;
            double r2[2];
            if (m_it) {
                if (k_it == m_it) {
                    const int t = sl[0];
                    for (int i = 0; i < m_it - 1; ++i) {
                        sl[i] = sl[i + 1];
                        rh[i] = rh[i + 1];
                    }
                    sl[m_it - 1] = t;
                } else {
                    ++k_it;
                }
                hipLaunchKernelGGL(k_sy, g, b, 0, st, n, a_m, ctx->d_lh.p, ctx->d_f.p, ctx->d_lf0.p, Sv(k_it - 1),
                                   Yv(k_it - 1));
                launch_check("k_sy");
                DotArgs D{{Sv(k_it - 1), Yv(k_it - 1), nullptr, nullptr}, {Yv(k_it - 1), Yv(k_it - 1), nullptr, nullptr}};
                dots<2>(ctx, n, D, r2);
                rh[k_it - 1] = 1.0 / r2[0];
            } else {
                // s and y into the first slots (no history is kept)
                hipLaunchKernelGGL(k_sy, g, b, 0, st, n, a_m, ctx->d_lh.p, ctx->d_f.p, ctx->d_lf0.p, ctx->d_lS.p,
                                   ctx->d_lY.p);
                launch_check("k_sy");
                DotArgs D{{ctx->d_lS.p, ctx->d_lY.p, nullptr, nullptr}, {ctx->d_lY.p, ctx->d_lY.p, nullptr, nullptr}};
                dots<2>(ctx, n, D, r2);
            }
            ratio = -r2[0] / r2[1];
            ++istp;
        }
        if (counts) {
            counts[0] = istp;
            counts[1] = n_force;
            counts[2] = n_energy;
            counts[3] = err;
        }
        return 0;
    });
}


// ------------------------------------------------------------------ extended CFG (host only)
namespace {
void put_err(char* err, int cap, const std::string& m) {
    if (err && cap > 0) {
        std::strncpy(err, m.c_str(), (size_t)cap - 1);
        err[cap - 1] = '\0';
    }
}
}  // namespace

int mapp_dmd_cfg_read(const char* path, int cap_atoms, int* natms, int* nt, double* H, double* x, double* c,
                      double* mass, char* names, int names_cap, char* err, int err_cap) {
    return guarded_noctx(err, err_cap, [&]() -> int {
        if (!path || !natms || !nt) return MAPP_DMD_E_ARG;
        CfgData d;
        const std::string e = read_cfg_file(path, d);
        if (!e.empty()) {
            put_err(err, err_cap, e);
            return MAPP_DMD_E_IO;
        }
        *natms = d.natms;
        *nt = d.nt;
        if (!x) return MAPP_DMD_OK;  // sizes only
        if (cap_atoms < d.natms) {
            put_err(err, err_cap, "buffer holds fewer atoms than the file");
            return MAPP_DMD_E_RANGE;
        }
        if (H) std::memcpy(H, d.H, sizeof(d.H));
        std::memcpy(x, d.x.data(), d.x.size() * sizeof(double));
        if (c) std::memcpy(c, d.c.data(), d.c.size() * sizeof(double));
        if (mass)
            for (int k = 0; k < d.nt; ++k) mass[k] = k < (int)d.mass.size() ? d.mass[k] : 0.0;
        if (names && names_cap > 0) {
            std::string all;
            for (size_t k = 0; k < d.names.size(); ++k) all += (k ? "\n" : "") + d.names[k];
            put_err(names, names_cap, all);
        }
        return MAPP_DMD_OK;
    });
}

int mapp_dmd_cfg_write(const char* path, const double* H, int natms, int nt, const double* x, const double* c,
                       int ntypes, const double* mass, const char* const* names, char* err, int err_cap) {
    return guarded_noctx(err, err_cap, [&]() -> int {
        if (!path || !H || !x || !c || natms < 0 || nt < 1 || ntypes < 0 || (ntypes && (!mass || !names)))
            return MAPP_DMD_E_ARG;
        CfgData d;
        d.natms = natms;
        d.nt = nt;
        std::memcpy(d.H, H, sizeof(d.H));
        d.x.assign(x, x + (size_t)natms * (3 + nt));
        d.c.assign(c, c + (size_t)natms * nt);
        for (int k = 0; k < ntypes; ++k) {
            d.mass.push_back(mass[k]);
            d.names.emplace_back(names[k]);
        }
        const std::string e = write_cfg_file(path, d);
        if (!e.empty()) {
            put_err(err, err_cap, e);
            return MAPP_DMD_E_IO;
        }
        return MAPP_DMD_OK;
    });
}


int mapp_dmd_bdf_run(mapp_dmd_t* ctx, int nsteps, int max_order, double initial_del_t, int max_retry, int flags,
                     double* report, long* counts) {
    return guarded(ctx, [&]() {
        require(!is_md(ctx) && ctx->inited, MAPP_DMD_E_STATE,
                "bdf_run needs an initialised DMD context");
        require(nsteps >= 0 && max_order >= 1 && max_order <= kBdfMaxOrder && max_retry >= 0 && report,
                MAPP_DMD_E_ARG, "bad bdf_run arguments");
        HIP_OK(hipSetDevice(ctx->device));
        const double a_tol = 1.0e-6, min_del_t = 1.0e-12, max_del_t = 1.0e4;
        const int nt = ctx->nt;
        const long n = (long)ctx->natms * nt;
        hipStream_t st = ctx->stream;
        // Clock_BDF::init (src/clock_bdf.cpp:201-275)
        double en[7];
        force_any(ctx, 1, en);
        with_nt(nt, [&](auto c) {
            build_list2<decltype(c)::value>(ctx);
            c_d_variant<decltype(c)::value>(ctx);
        });
        ctx->d_bdf.reserve((size_t)n * (max_order + 8));
        BdfState S{ctx, n, nt};
        S.fix = (flags & 1) != 0;
        double* base = ctx->d_bdf.p;
        for (int i = 0; i < max_order; ++i) S.y[i] = base + (size_t)i * n;
        double* more = base + (size_t)max_order * n;
        S.dy = more;
        S.a = more + n;
        S.y0 = more + 2 * n;
        S.g = more + 3 * n;
        S.g0 = more + 4 * n;
        S.h = more + 5 * n;
        S.c0 = more + 6 * n;
        S.e_n = more + 7 * n;
        HIP_OK(hipMemsetAsync(S.e_n, 0, n * sizeof(double), st));
        if (initial_del_t <= 0.0) {
            const double sum = std::sqrt(S.dot(ctx->d_cd.p, ctx->d_cd.p) / (double)n);
            initial_del_t = std::max(a_tol / sum, min_del_t);
            initial_del_t = std::min(initial_del_t, max_del_t);
        }
        HIP_OK(hipMemcpyAsync(S.y[0], ctx->d_conc.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIP_OK(hipMemcpyAsync(S.dy, ctx->d_cd.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        double del_t = initial_del_t, err = 0.0, cost = 0.0;
        int ord = 1, const_stp = 0, istep = 0, rc = 0;
        while (istep < nsteps) {
            int chk = S.interpolate(ord, del_t);
            while (chk == -1) {
                del_t *= 0.8;
                del_t = std::max(del_t, min_del_t);
                chk = S.interpolate(ord, del_t);
            }
            int iters, it_sum = 0, solves = 1, retries = 0;
            cost = S.solve(ord, del_t, &err, &iters);
            it_sum += iters;
            while (cost >= 0 && (err >= 1.0 || cost >= 1.0)) {
                if (++retries > max_retry) {
                    rc = -2;
                    break;
                }
                const double err1 = std::max(err, cost);
                double ratio = std::pow(0.5 / err1, 1.0 / (double)(ord + 1));
                ratio = ratio < 0.5 ? 0.5 : (ratio > 0.9 ? 0.9 : ratio);
                const double tmp = del_t * ratio;
                del_t = tmp < min_del_t ? min_del_t : tmp;
                chk = S.interpolate(ord, del_t);
                while (chk == -1) {
                    del_t *= 0.8;
                    chk = S.interpolate(ord, del_t);
                }
                cost = S.solve(ord, del_t, &err, &iters);
                it_sum += iters;
                ++solves;
            }
            double* r = report + (size_t)istep * 6;
            r[0] = ord;
            r[1] = del_t;
            r[2] = cost;
            r[3] = err;
            r[4] = it_sum;
            r[5] = solves;
            if (cost < 0) rc = -1;
            if (rc) break;
            // history shift (src/clock_bdf.cpp:566-579)
            double* tmpy = S.y[max_order - 1];
            for (int i = max_order - 1; i > 0; --i) {
                S.t[i] = S.t[i - 1];
                S.y[i] = S.y[i - 1];
            }
            S.y[0] = tmpy;
            S.t[0] += del_t;
            HIP_OK(hipMemcpyAsync(S.y[0], ctx->d_conc.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            HIP_OK(hipMemcpyAsync(S.dy, ctx->d_cd.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            // step and order control (src/clock_bdf.cpp:582-664)
            const bool lo_avail = ord != 1 && const_stp > ord + 1;
            const bool hi_avail = ord < max_order && istep != 0 && const_stp > ord + 1;
            double lo_err, hi_err, ratio, lo_ratio, hi_ratio;
            S.hi_lo(ord, del_t, hi_avail, lo_avail, &lo_err, &hi_err);
            if (hi_avail && lo_avail) {
                lo_ratio = std::pow(0.5 / lo_err, 1.0 / (double)ord);
                hi_ratio = std::pow(0.5 / hi_err, 1.0 / (double)(ord + 2));
                ratio = std::pow(0.5 / err, 1.0 / (double)(ord + 1));
                if (hi_ratio > lo_ratio && hi_ratio > ratio) {
                    ++ord;
                    ratio = hi_ratio;
                } else if (lo_ratio > hi_ratio && lo_ratio > ratio) {
                    --ord;
                    ratio = lo_ratio;
                }
            } else if (hi_avail && !lo_avail) {
                hi_ratio = std::pow(0.5 / hi_err, 1.0 / (double)(ord + 2));
                ratio = std::pow(0.5 / err, 1.0 / (double)(ord + 1));
                if (hi_ratio > ratio) {
                    ++ord;
                    ratio = hi_ratio;
                }
            } else if (!hi_avail && lo_avail) {
                lo_ratio = std::pow(0.5 / lo_err, 1.0 / (double)ord);
                ratio = std::pow(0.5 / err, 1.0 / (double)(ord + 1));
                if (lo_ratio > ratio) {
                    --ord;
                    ratio = lo_ratio;
                }
            } else {
                ratio = std::pow(0.5 / err, 1.0 / (double)(ord + 1));
            }
            if (ratio >= 2.0) {
                ratio = 2.0;
                const_stp = 0;
            } else if (ratio <= 0.5) {
                ratio = 0.5;
                const_stp = 0;
            } else {
                ratio = 1.0;
                ++const_stp;
            }
            const double tmp = ratio * del_t;
            del_t = tmp > max_del_t ? max_del_t : (tmp < min_del_t ? min_del_t : tmp);
            ++istep;
        }
        if (counts) {
            counts[0] = rc ? rc : istep;
            counts[1] = S.n_calc_g;
        }
        return 0;
    });
}


int mapp_dmd_autotune(mapp_dmd_t* ctx, int reps, double* report, int cap, int* ncand) {
    return guarded(ctx, [&]() {
        require(ctx->inited && !is_md(ctx), MAPP_DMD_E_STATE,
                "autotune needs an initialised DMD context");
        require(reps >= 1, MAPP_DMD_E_ARG, "reps >= 1");
        HIP_OK(hipSetDevice(ctx->device));
        const bool timing0 = ctx->timing;
        ctx->timing = true;
        // candidate schedules: waves per block x total waves (0 = one wavefront per site;
        // otherwise a persistent grid of 1..16 waves per SIMD over the chip's 1024 SIMDs)
        std::vector<mapp_dmd::Shape> cand;
        for (int parts : {4, 2, 1})
            for (int wpb : {1, 2, 4})
                for (long waves : {0L, 1024L, 2048L, 4096L, 8192L, 16384L})
                    if (waves == 0 || waves < ctx->natms) cand.push_back({wpb, waves, parts});
        double en[7];
        double best1 = 1e300, best2 = 1e300;
        mapp_dmd::Shape b1 = ctx->sh1, b2 = ctx->sh2;
        int row = 0;
        for (const auto& c : cand) {
            ctx->sh1 = c;
            ctx->sh2 = c;
            double t1 = 1e300, t2 = 1e300;
            for (int r = 0; r < reps + 1; ++r) {  // the first run warms up
                with_nt(ctx->nt, [&](auto c) { dispatch_force<decltype(c)::value>(ctx, 1, en); });
                if (r == 0) continue;
                t1 = std::min(t1, ctx->last_ms[0]);
                t2 = std::min(t2, ctx->last_ms[2]);
            }
            if (report && row < cap) {
                report[row * 5 + 0] = c.parts;
                report[row * 5 + 1] = c.wpb;
                report[row * 5 + 2] = (double)c.waves;
                report[row * 5 + 3] = t1;
                report[row * 5 + 4] = t2;
            }
            ++row;
            if (t1 < best1) {
                best1 = t1;
                b1 = c;
            }
            if (t2 < best2) {
                best2 = t2;
                b2 = c;
            }
        }
        ctx->sh1 = b1;
        ctx->sh2 = b2;
        ctx->timing = timing0;
        if (ncand) *ncand = row;
        return 0;
    });
}

int mapp_dmd_get_schedule(mapp_dmd_t* ctx, int* out) {
    return guarded(ctx, [&]() {
        require(out != nullptr, MAPP_DMD_E_ARG, "null out");
        out[0] = ctx->sh1.parts;
        out[1] = ctx->sh1.wpb;
        out[2] = (int)ctx->sh1.waves;
        out[3] = ctx->sh2.wpb;
        out[4] = (int)ctx->sh2.waves;
        return 0;
    });
}

int mapp_dmd_set_schedule(mapp_dmd_t* ctx, const int* in) {
    return guarded(ctx, [&]() {
        require(in != nullptr, MAPP_DMD_E_ARG, "null schedule");
        require(in[0] == 1 || in[0] == 2 || in[0] == 4, MAPP_DMD_E_ARG, "pass-1 lanes per (pair, type pair): 1, 2 or 4");
        for (int k : {1, 3}) require(in[k] == 1 || in[k] == 2 || in[k] == 4, MAPP_DMD_E_ARG, "waves per block 1, 2 or 4");
        for (int k : {2, 4}) require(in[k] >= 0, MAPP_DMD_E_ARG, "negative wave count");
        ctx->sh1 = {in[1], (long)in[2], in[0]};
        ctx->sh2 = {in[3], (long)in[4], 4};
        return 0;
    });
}


int mapp_dmd_set_cells(mapp_dmd_t* ctx, int ncell, const double* H, const int* natms, const double* x,
                       const double* c) {
    return guarded(ctx, [&]() {
        require(ncell >= 1 && H && natms && x && c, MAPP_DMD_E_ARG, "bad cells");
        require(!is_md(ctx), MAPP_DMD_E_ARG, "batches hold DMD cells");
        const int nt = ctx->nt, xc = xcols(ctx);
        long tot = 0;
        for (int k = 0; k < ncell; ++k) {
            require(natms[k] > 0, MAPP_DMD_E_ARG, "empty cell");
            tot += natms[k];
        }
        require(tot < (1L << 30), MAPP_DMD_E_RANGE, "too many sites");
        ctx->cell_natms.assign(natms, natms + ncell);
        ctx->cell_H.assign(H, H + (size_t)ncell * 9);
        ctx->cell_x.assign(x, x + (size_t)tot * xc);
        ctx->cell_c.assign(c, c + (size_t)tot * nt);
        ctx->natms = (int)tot;
        ctx->inited = false;
        ctx->have_list2 = false;
        ctx->cache_valid = false;
        ctx->have_system = true;
        return 0;
    });
}

int mapp_dmd_cells(mapp_dmd_t* ctx, int* ncell, int* natms, int* nghost, int cap) {
    return guarded(ctx, [&]() {
        require(ncell != nullptr, MAPP_DMD_E_ARG, "null ncell");
        const int n = (int)ctx->cell_natms.size();
        *ncell = n;
        for (int k = 0; k < n && k < cap; ++k) {
            if (natms) natms[k] = ctx->cell_natms[k];
            if (nghost) nghost[k] = k < (int)ctx->cell_nghost.size() ? ctx->cell_nghost[k] : 0;
        }
        return 0;
    });
}

int mapp_dmd_get_vec(mapp_dmd_t* ctx, int which, double* host, long count) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede get_vec");
        require(host != nullptr, MAPP_DMD_E_ARG, "null destination");
        HIP_OK(hipSetDevice(ctx->device));
        const int nt = acols(ctx), natms = ctx->natms;
        const double* src = nullptr;
        long n = (long)natms * nt;
        switch (which) {
            case MAPP_DMD_VEC_X: {
                n = (long)natms * xcols(ctx);
                ctx->d_xaos.reserve(n);
                gather_x_any(ctx, ctx->d_xaos.p, natms, ctx->d_pos.p, ctx->d_alpha.p);
                src = ctx->d_xaos.p;
                break;
            }
            case MAPP_DMD_VEC_C: src = ctx->d_conc.p; break;
            case MAPP_DMD_VEC_F: src = ctx->d_f.p; n = (long)natms * xcols(ctx); break;
            case MAPP_DMD_VEC_E: src = ctx->d_E.p; break;
            case MAPP_DMD_VEC_DE: src = ctx->d_dE.p; break;
            case MAPP_DMD_VEC_N: src = ctx->d_n.p; break;
            case MAPP_DMD_VEC_CD: src = ctx->d_cd.p; break;
            case MAPP_DMD_VEC_DDE: src = ctx->d_ddE.p; break;
            case MAPP_DMD_VEC_S: src = ctx->d_s.p; break;
            case MAPP_DMD_VEC_T: src = ctx->d_t.p; break;
            case MAPP_DMD_VEC_V: src = ctx->d_v.p; break;
            case MAPP_DMD_VEC_A: src = ctx->d_a.p; break;
            case MAPP_DMD_VEC_G: src = ctx->d_g.p; break;
            default: throw ApiError(MAPP_DMD_E_ARG, "unknown vector");
        }
        require(count >= n, MAPP_DMD_E_ARG, "destination too small");
        HIP_OK(hipMemcpyAsync(host, src, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int mapp_dmd_device_ptr(mapp_dmd_t* ctx, int which, void** dev_ptr) {
    return guarded(ctx, [&]() {
        require(ctx->inited && dev_ptr, MAPP_DMD_E_STATE, "init must precede device_ptr");
        switch (which) {
            case MAPP_DMD_VEC_X:
                ctx->d_xaos.reserve((size_t)ctx->natms * (3 + ctx->nt));
                *dev_ptr = ctx->d_xaos.p;
                break;
            case MAPP_DMD_VEC_C: *dev_ptr = ctx->d_conc.p; break;
            case MAPP_DMD_VEC_F: *dev_ptr = ctx->d_f.p; break;
            case MAPP_DMD_VEC_E: *dev_ptr = ctx->d_E.p; break;
            case MAPP_DMD_VEC_DE: *dev_ptr = ctx->d_dE.p; break;
            case MAPP_DMD_VEC_N: *dev_ptr = ctx->d_n.p; break;
            case MAPP_DMD_VEC_CD: *dev_ptr = ctx->d_cd.p; break;
            case MAPP_DMD_VEC_A: *dev_ptr = ctx->d_a.p; break;
            case MAPP_DMD_VEC_G: *dev_ptr = ctx->d_g.p; break;
            default: throw ApiError(MAPP_DMD_E_ARG, "no device pointer for this vector");
        }
        return 0;
    });
}

int mapp_dmd_set_vec(mapp_dmd_t* ctx, int which, const double* host, long count) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede set_vec");
        require(host != nullptr, MAPP_DMD_E_ARG, "null source");
        HIP_OK(hipSetDevice(ctx->device));
        const int nt = ctx->nt, natms = ctx->natms;
        double* dst = nullptr;
        long n = (long)natms * nt;
        switch (which) {
            case MAPP_DMD_VEC_X:
                n = (long)natms * xcols(ctx);
                ctx->d_xaos.reserve(n);
                dst = ctx->d_xaos.p;
                break;
            case MAPP_DMD_VEC_A: dst = ctx->d_a.p; break;
            case MAPP_DMD_VEC_G: dst = ctx->d_g.p; break;
            default: throw ApiError(MAPP_DMD_E_ARG, "set_vec accepts X, A or G");
        }
        require(count == n, MAPP_DMD_E_ARG, "count does not match the vector size");
        HIP_OK(hipMemcpyAsync(dst, host, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int mapp_dmd_sync(mapp_dmd_t* ctx) {
    return guarded(ctx, [&]() {
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int mapp_dmd_sizes(mapp_dmd_t* ctx, long* out) {
    return guarded(ctx, [&]() {
        require(out != nullptr, MAPP_DMD_E_ARG, "null output");
        out[0] = ctx->natms;
        out[1] = ctx->nghost;
        out[2] = ctx->npairs;
        out[3] = ctx->have_list2 ? ctx->npairs2 : 0;
        out[4] = ctx->swap_off.empty() ? 0 : (long)ctx->swap_off.size() - 1;
        out[5] = ctx->n_inrange;
        return 0;
    });
}

int mapp_dmd_get_cutoffs(mapp_dmd_t* ctx, double* out) {
    return guarded(ctx, [&]() {
        require(out != nullptr, MAPP_DMD_E_ARG, "null output");
        require(ctx->have_params && !is_md(ctx), MAPP_DMD_E_STATE, "the DMD cutoffs follow coef");
        out[0] = ctx->P.rc;
        out[1] = ctx->mod_rc;
        out[2] = std::sqrt(ctx->rsq_crd);
        return 0;
    });
}

int mapp_dmd_get_list(mapp_dmd_t* ctx, int* offsets, int* entries) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede get_list");
        if (offsets)
            HIP_OK(hipMemcpy(offsets, ctx->d_off.p, sizeof(int) * (ctx->natms + 1), hipMemcpyDeviceToHost));
        if (entries && ctx->npairs > 0)
            HIP_OK(hipMemcpy(entries, ctx->d_nl.p, sizeof(int) * ctx->npairs, hipMemcpyDeviceToHost));
        return 0;
    });
}

int mapp_dmd_get_list2(mapp_dmd_t* ctx, int* offsets, int* entries) {
    return guarded(ctx, [&]() {
        require(ctx->have_list2, MAPP_DMD_E_STATE, "create_2nd_neigh_lst must precede get_list2");
        if (offsets)
            HIP_OK(hipMemcpy(offsets, ctx->d_off2.p, sizeof(int) * (ctx->natms + 1), hipMemcpyDeviceToHost));
        if (entries && ctx->npairs2 > 0)
            HIP_OK(hipMemcpy(entries, ctx->d_nl2.p, sizeof(int) * ctx->npairs2, hipMemcpyDeviceToHost));
        return 0;
    });
}

int mapp_dmd_get_ghosts(mapp_dmd_t* ctx, double* x, int* src) {
    return guarded(ctx, [&]() {
        require(ctx->inited, MAPP_DMD_E_STATE, "init must precede get_ghosts");
        const int nt = acols(ctx), ng = ctx->nghost;
        if (ng == 0) return 0;
        if (x) {
            DevBuf<double> tmp;
            tmp.reserve((size_t)ng * xcols(ctx));
            gather_x_any(ctx, tmp.p, ng, ctx->d_pos.p + ctx->natms, ctx->d_alpha.p + (size_t)ctx->natms * nt);
            HIP_OK(hipMemcpyAsync(x, tmp.p, sizeof(double) * ng * xcols(ctx), hipMemcpyDeviceToHost, ctx->stream));
            HIP_OK(hipStreamSynchronize(ctx->stream));
        }
        if (src) HIP_OK(hipMemcpy(src, ctx->d_ghost_src.p, sizeof(int) * ng, hipMemcpyDeviceToHost));
        return 0;
    });
}

int mapp_dmd_get_gh(mapp_dmd_t* ctx, double* xi, double* w0, double* w1, double* w2) {
    return guarded(ctx, [&]() {
        require(ctx->have_params, MAPP_DMD_E_STATE, "coef must precede get_gh");
        for (int k = 0; k < ctx->P.no_i; ++k) {
            if (xi) xi[k] = ctx->P.xi[k];
            if (w0) w0[k] = ctx->P.w0[k];
            if (w1) w1[k] = ctx->P.w1[k];
            if (w2) w2[k] = ctx->P.w2[k];
        }
        return 0;
    });
}

int mapp_dmd_last_timing(mapp_dmd_t* ctx, double* ms) {
    return guarded(ctx, [&]() {
        require(ms != nullptr, MAPP_DMD_E_ARG, "null output");
        for (int k = 0; k < 5; ++k) ms[k] = ctx->last_ms[k];
        return 0;
    });
}

int mapp_dmd_enable_timing(mapp_dmd_t* ctx, int on) {
    return guarded(ctx, [&]() {
        ctx->timing = on != 0;
        return 0;
    });
}

}  // extern "C"

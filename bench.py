"""Benchmark of the DMD force field on one MI355X.

A step is one energy+gradient evaluation, force_calc(1, en_st), over the
whole system with inputs resident in HBM: binary FCC Ni-Cu, 64^3 cells
(1,048,576 sites), Johnson-style EAM on 10k-point tables, 3-point
Gauss-Hermite rule, FP64 quadrature.  value = sites / step time.

Each path first runs the launch-schedule autotuner on its own system (or
takes --schedule) and reports the schedule it kept.  Also reported on the
same system: the mixed path (5-point rule, FP32
quadrature sums; errors of energy, forces, alpha-forces, E, dE and dc/dt
against the FP64 5-point path), the learned surrogate (trained here on 4M
quadrature samples, BF16 MFMA inference, quadrature fallback wherever the
networks miss the per-pair tolerances of a 5e-6 force bound; with the curve
bound -> served share, errors, site-evals/s), a 50-step L-BFGS run
with neighbour rebuilds (consistent energy mode), a batch of 96 uneven cells
(surrogate force in one launch per kernel, autotuned, 20 BDF steps), MEAM,
plain EAM, Lennard-Jones, Finnis-Sinclair and FSM force_calc on the same
lattice size, the roofline of the
dominant kernel (pass 1, timed with HIP events on the library's stream) and
the FP64 CPU oracle timed on a bounded sample (32^3 cells per thread, one
thread and then one box per thread of the process's CPU share, at most 16,
the share of one GPU on these boxes) and once on the full system, one thread.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--no-surrogate] [--no-lbfgs] [--no-batch] [--no-meam] [--fp64-only]
                    [--schedule parts,wpb,waves,wpb2,waves2]
    python bench.py --copy-calibration   (plain copy kernel for HBM counter scaling)
    python bench.py --meam-only          (MEAM force_calc alone, for profiling)
    python bench.py --md-only            (plain EAM, LJ, FS, FSM force_calc alone)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pass1_flops_per_site(P, nt, q):
    """Algorithmic FP64 operations of pass 1 per site (the reference's force_calc
    quadrature): per in-range pair nt^2 density quadratures of (20 + 19 q) and
    nt(nt+1)/2 pair quadratures of (20 + 12 q) operations."""
    return P * (nt * nt * (20 + 19 * q) + nt * (nt + 1) // 2 * (20 + 12 * q))


def site_flops(P, nt, q):
    """whole force_calc per site: pass 1 + back-pass (30 per pair and type pair) + embedding (40 per type)"""
    return pass1_flops_per_site(P, nt, q) + P * 30 * nt * nt + 40 * nt


def bytes_per_site(P_list, nt):
    """Algorithmic HBM bytes per site-evaluation: x, c, f, E/dE writes, dE read, two list passes."""
    return 8 * (3 + nt) + 8 * nt + 8 * (3 + nt) + 16 * nt + 8 * nt + 2 * 4 * P_list


def time_force(ff, steps, warmup):
    ff.enable_timing(True)
    for _ in range(warmup):
        ff.force_calc(1)
    ms_kernels = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(steps):
        en = ff.force_calc(1)
        ms_kernels += ff.last_timing()
    t1 = time.perf_counter()
    return dict(ms_per_step=(t1 - t0) * 1e3 / steps, kernel_ms=(ms_kernels / steps).tolist(), energy=float(en[0]))


def run_path(lat, tables, no_i, precision, steps, warmup, schedule=None, mobility=False, keep=False):
    """force_calc timing after the launch-schedule autotuner (or with a given schedule);
    with mobility, also one c_d_calc and one calc_g (the +dc/dt evaluation); with keep,
    E, dE and c_d of the system for error reports"""
    from mapp_amd.dmd import setup
    ff = setup(lat, tables, no_i=no_i, precision=precision)
    if schedule is None:
        ff.autotune(reps=2)
    else:
        ff.set_schedule(schedule)
    r = time_force(ff, steps, warmup)
    r["schedule"] = ff.schedule()
    if mobility:
        ff.create_2nd_neigh_lst()
        t0 = time.perf_counter()
        ff.c_d_calc()  # first call: pair values and the site table
        r["c_d_first_ms"] = (time.perf_counter() - t0) * 1e3
        reps = max(3, steps // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            ff.c_d_calc()
        r["c_d_ms"] = (time.perf_counter() - t0) * 1e3 / reps
        a = np.zeros_like(ff.get("c"))
        ff.set_vec("a", a)
        t0 = time.perf_counter()
        for _ in range(reps):
            ff.calc_g(0, 1e-3)
        r["calc_g_ms"] = (time.perf_counter() - t0) * 1e3 / reps
    r["sizes"] = ff.sizes()
    r["f"] = ff.get("f")
    if keep:
        ff.force_calc(1)
        r["E"], r["dE"] = ff.get("E"), ff.get("dE")
        ff.create_2nd_neigh_lst()
        ff.c_d_calc()
        r["c_d"] = ff.get("c_d")
    if mobility:
        # Atoms::update_0 on device-resident positions: one site moved past the
        # skin (ghosts, bins, half list, reverse list rebuilt), then small moves
        # only (ghost positions forwarded)
        x = ff.get("x")
        x[7, 0] += 0.8
        ff.set_vec("x", x)
        t0 = time.perf_counter()
        rebuilt = ff.set_x(None)
        ff.sync()
        r["list_rebuild_ms"] = (time.perf_counter() - t0) * 1e3
        assert rebuilt
        x = ff.get("x")  # the rebuild may have wrapped the moved site into the box
        x[:, :3] += 1e-3
        ff.set_vec("x", x)
        t0 = time.perf_counter()
        rebuilt = ff.set_x(None)
        ff.sync()
        r["x_update_ms"] = (time.perf_counter() - t0) * 1e3
        assert not rebuilt
    ff.close()
    return r


def run_surrogate(lat, tables, steps, warmup, n_train, f_ref, e_ref, bound=5e-6,
                  curve=(5e-6, 1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1e30)):
    """train on n_train quadrature samples (4 type pairs) with the fallback mask of a
    relative force bound, time force_calc at that operating point, then the curve
    bound -> (share of the quadrature-branch evaluations the networks serve, force and
    energy error against the FP64 5-point path, site-evals/s); bound -1 is the
    all-fallback path (the FP32 quadrature inside the surrogate kernel)"""
    from mapp_amd.dmd import setup
    ff = setup(lat, tables, no_i=5, precision="mixed")
    ff.autotune(reps=2)  # the mixed pass stands in where no cell is served
    t0 = time.perf_counter()
    rep = ff.surrogate_train(n_per_tp=n_train // 4, epochs=4, batch_per_tp=1024, lr=2e-3, seed=3, tol=bound)
    train_s = time.perf_counter() - t0
    ff.set_precision("surrogate")

    def point(b, k):
        info = ff.surrogate_set_bound(b)
        r = time_force(ff, k, 1)
        f = ff.get("f")
        return dict(bound=b, served_cells=info["served"], fallbacks=ff.surrogate_fallbacks(),
                    ms_per_step=r["ms_per_step"], pass1_ms=r["kernel_ms"][0],
                    site_evals_per_s=lat.natms / (r["ms_per_step"] / 1e3),
                    force_rel_err_vs_fp64_q5=float(np.max(np.abs(f - f_ref)) / np.max(np.abs(f_ref))),
                    energy_rel_err_vs_fp64_q5=abs(r["energy"] - e_ref) / abs(e_ref)), info

    base, _ = point(-1.0, max(3, steps // 4))
    total = base["fallbacks"]
    rows = []
    for b in curve:
        row, info = point(b, max(3, steps // 4))
        row["served_evals"] = 1.0 - row["fallbacks"] / total
        rows.append(row)
    ff.surrogate_set_bound(bound)
    r = time_force(ff, steps, warmup)
    f = ff.get("f")
    info = ff.surrogate_set_bound(bound)
    out = dict(ms_per_step=r["ms_per_step"], kernel_ms=r["kernel_ms"], bound=bound, train_seconds=train_s,
               train_report=rep, fallbacks=ff.surrogate_fallbacks(), all_fallback=base, curve=rows,
               tolerances=info["eps"].tolist(), scales=info["scales"].tolist(),
               force_rel_err_vs_fp64=float(np.max(np.abs(f - f_ref)) / np.max(np.abs(f_ref))),
               energy_rel_err_vs_fp64=abs(r["energy"] - e_ref) / abs(e_ref))
    out["served_evals"] = 1.0 - out["fallbacks"] / total
    ff.close()
    return out


def run_lbfgs(lat, tables, iters):
    """Min_LBFGS on the device from the same start, consistent energy mode, no energy tolerance"""
    from mapp_amd.dmd import setup
    ff = setup(lat, tables, no_i=3)
    ff.set_energy_mode(True)
    t0 = time.perf_counter()
    code, en, nf, ne = ff.min_lbfgs(iters - 1, 2, -1.0)
    dt = time.perf_counter() - t0
    ff.close()
    steps = len(en) - 1
    return dict(steps=steps, exit_code=code, force_calls=nf, energy_calls=ne, seconds=dt,
                ms_per_step=dt * 1e3 / max(steps, 1), energy_first=float(en[0]), energy_last=float(en[-1]))


def run_batch(tables, steps, warmup, ncell=96, bdf_steps=20, seed=5):
    """96 independent cells, n^3 FCC cells with n log-uniform in [6, 40] (864 to 256,000 sites),
    in one context: surrogate+FP32 force_calc over all cells in one launch per kernel (autotuned
    schedule), then 20 BDF steps (orders up to 2) of dc/dt on the whole batch."""
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_cells
    rng = np.random.default_rng(seed)
    ns = np.clip(np.round(np.exp(rng.uniform(np.log(6), np.log(40), ncell))).astype(int), 6, 40)
    lats = [W.fcc_alloy(int(n), seed=seed + k) for k, n in enumerate(ns)]
    natms = int(sum(l.natms for l in lats))
    t0 = time.perf_counter()
    ff = setup_cells(lats, tables, no_i=5, precision="mixed")
    build_s = time.perf_counter() - t0
    rep = ff.surrogate_train(n_per_tp=1 << 20, epochs=2, batch_per_tp=1024, lr=2e-3, seed=3, tol=5e-6)
    ff.set_precision("surrogate")
    rows = ff.autotune(reps=2)
    r = time_force(ff, steps, warmup)
    t0 = time.perf_counter()
    code, brep, ncg = ff.bdf_run(bdf_steps, 2)
    bdf_s = time.perf_counter() - t0
    ff.close()
    # the same cells with the DMD_II mobility, from an initial step at which the
    # controller holds the step size and raises the order (the DMD_I runs above stay at
    # order 1: their solver's cost test caps the step where the error is far below 1)
    ff2 = setup_cells(lats, tables, no_i=5, variant="DMD_II", precision="fp64")
    t0 = time.perf_counter()
    code2, brep2, ncg2 = ff2.bdf_run(bdf_steps, 2, initial_del_t=8.0)
    bdf2_s = time.perf_counter() - t0
    ff2.close()
    return dict(cells=ncell, sites=natms, sites_min=int(min(l.natms for l in lats)),
                sites_max=int(max(l.natms for l in lats)), build_seconds=build_s,
                force_site_evals_per_s=natms / (r["ms_per_step"] / 1e3), force_ms=r["ms_per_step"],
                kernel_ms=r["kernel_ms"][:4], schedule_candidates=int(len(rows)), surrogate_val_err=rep["val_mean_err"],
                surrogate_bound=5e-6, surrogate_masked_cells=rep["masked_frac"],
                bdf_steps=code, bdf_seconds=bdf_s, bdf_calc_g_calls=ncg,
                bdf_orders=[int(v) for v in brep[:, 0]] if code > 0 else [],
                bdf_dmd_ii=dict(steps=code2, seconds=bdf2_s, calc_g_calls=ncg2, initial_del_t=8.0,
                                orders=[int(v) for v in brep2[:, 0]] if code2 > 0 else []))


def run_meam(n, steps, warmup, cpu_n=8):
    """MEAM (ForceField_meam) force_calc on n^3 FCC cells of the two-element
    synthetic alloy (jittered, L12 cross lattice, default screening), wall
    time per call with the result on the host; the FP64 oracle on an cpu_n^3
    sample, one thread"""
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_meam
    from oracle import oracle as O
    st = W.meam_state(["Aa", "Bb"], local=W.meam_local(2))
    H, s, t = W.md_lattice(n, 3.58, "fcc", nt=2, seed=41, jitter=0.05)
    t0 = time.perf_counter()
    ff = setup_meam(H, s, t, st)
    build_s = time.perf_counter() - t0
    for _ in range(warmup):
        ff.force_calc(1)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        ff.force_calc(1)
        times.append(time.perf_counter() - t0)
    entries = int(ff.neighbor_list()[0][-1])
    ff.close()
    ms = float(np.median(times)) * 1e3
    Hc, sc, tc = W.md_lattice(cpu_n, 3.58, "fcc", nt=2, seed=41, jitter=0.05)
    o = O.Oracle(nt=2, variant="MEAM")
    o.meam_coef(st)
    o.set_system(Hc, sc.copy(), tc.reshape(-1, 1))
    o.init(0.3)
    ct = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.force_calc(1)
        ct.append(time.perf_counter() - t0)
    return dict(sites=len(t), list_entries=entries, build_seconds=build_s, ms_per_force=ms,
                site_evals_per_s=len(t) / (ms / 1e3),
                cpu_oracle=dict(sites=len(tc), site_evals_per_s=len(tc) / float(np.median(ct)), cores=1))


def run_md(n, tables, steps, warmup, cpu_n=8):
    """The MD-mode potentials on n^3 cells (1,048,576 sites at n = 64 fcc):
    plain EAM (Johnson tables, FCC a = 3.6, one site in three of type 1),
    Lennard-Jones (two types, FCC a = 3.6), Finnis-Sinclair and its
    squared-cutoff form (two types, BCC a = 2.87, 2 n^3 sites); force_calc wall
    time per call, and the FP64 oracle on a cpu_n^3 sample, one thread"""
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup_eam, setup_fs, setup_lj
    from oracle import oracle as O

    def eam(nn):
        lat = W.fcc_alloy(nn, seed=21)
        t = (np.arange(lat.natms) % 3 == 0).astype(np.float64)
        return lat.H, np.ascontiguousarray(lat.x[:, :3]), t

    cases = {
        "eam": (lambda nn: eam(nn), lambda H, s, t: setup_eam(H, s, t, tables),
                lambda: _orc_eam(O, tables)),
        "lj": (lambda nn: W.md_lattice(nn, 3.6, "fcc", seed=3), lambda H, s, t: setup_lj(H, s, t, W.LJ_COEFS, nt=2),
               lambda: _orc_lj(O, W)),
        "fs": (lambda nn: W.md_lattice(nn, 2.87, "bcc", seed=5), lambda H, s, t: setup_fs(H, s, t, W.fs_params()),
               lambda: _orc_fs(O, W, "FS")),
        "fsm": (lambda nn: W.md_lattice(nn, 2.87, "bcc", seed=5),
                lambda H, s, t: setup_fs(H, s, t, W.fs_params(), variant="FSM"), lambda: _orc_fs(O, W, "FSM")),
    }
    out = {}
    for name, (lat, make, orc) in cases.items():
        H, s, t = lat(n)
        ff = make(H, s, t)
        for _ in range(warmup):
            ff.force_calc(1)
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            ff.force_calc(1)
            times.append(time.perf_counter() - t0)
        ff.close()
        ms = float(np.median(times)) * 1e3
        Hc, sc, tc = lat(cpu_n)
        o = orc()
        o.set_system(Hc, sc.copy(), np.asarray(tc, dtype=np.float64).reshape(-1, 1))
        o.init(0.5)
        ct = []
        for _ in range(3):
            t0 = time.perf_counter()
            o.force_calc(1)
            ct.append(time.perf_counter() - t0)
        out[name] = dict(sites=len(t), ms_per_force=ms, site_evals_per_s=len(t) / (ms / 1e3),
                         cpu_oracle=dict(sites=len(tc), site_evals_per_s=len(tc) / float(np.median(ct)), cores=1))
    return out


def _orc_eam(O, tables):
    o = O.Oracle(nt=2, variant="EAM")
    o.set_fs(tables)
    o.set_params(masses=tables.masses)
    return o


def _orc_lj(O, W):
    o = O.Oracle(nt=2, variant="LJ")
    for c in W.LJ_COEFS:
        o.lj_coef(*c)
    return o


def _orc_fs(O, W, variant):
    o = O.Oracle(nt=2, variant=variant)
    o.fs_coef(W.fs_params())
    return o


def cpu_threads():
    """host threads for the multi-rank CPU baseline: the process's CPU share
    (OMP_NUM_THREADS where set), at most 16 -- the host share that comes with
    one GPU on the MI355X boxes this runs on"""
    k = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        k = min(k, int(env))
    return max(1, min(k, 16))


def cpu_full_size(lat, tables):
    """the FP64 oracle's force_calc(1) on the full benchmark system, one thread, one call
    (list build and caches set up before the timer)"""
    from oracle import oracle as O
    orc = O.make(lat, tables, no_i=3, variant="DMD_I")
    t0 = time.perf_counter()
    orc.force_calc(1)
    dt = time.perf_counter() - t0
    del orc
    return dict(value=lat.natms / dt, cores=1, seconds=dt,
                sample=f"one force_calc(1) on the full {lat.natms}-site system, 1 thread")


def cpu_baseline(tables, n=32, reps=3, threads=None):
    """The FP64 oracle's force_calc on the host.  One thread on one n^3 box
    (the reference's one rank), then K threads each with its own n^3 box run
    together, which models K ranks of the reference with no communication
    (each box periodic, so its ghosts cost what a rank's halo would);
    value = K * sites / wall time of the K concurrent calls, median of reps.
    ctypes releases the GIL for the oracle's calls, so the threads run in
    parallel."""
    import threading
    from mapp_amd import workload as W
    from oracle import oracle as O
    lat = W.fcc_alloy(n, seed=7)
    orc = O.make(lat, tables, no_i=3, variant="DMD_I")
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        orc.force_calc(1)
        times.append(time.perf_counter() - t0)
    del orc
    one = lat.natms / float(np.median(times))
    k = threads or cpu_threads()
    bar = threading.Barrier(k + 1)
    errs = []

    def rank(seed):
        try:
            o = O.make(W.fcc_alloy(n, seed=seed), tables, no_i=3, variant="DMD_I")
        except Exception as e:  # keep the barrier count intact
            errs.append(e)
            o = None
        for _ in range(reps):
            bar.wait()
            if o is not None:
                o.force_calc(1)
            bar.wait()

    ts = [threading.Thread(target=rank, args=(7 + r,)) for r in range(k)]
    for t in ts:
        t.start()
    walls = []
    for _ in range(reps):
        bar.wait()
        t0 = time.perf_counter()
        bar.wait()
        walls.append(time.perf_counter() - t0)
    for t in ts:
        t.join()
    if errs:
        raise errs[0]
    return dict(value=k * lat.natms / float(np.median(walls)), unit="site-evals/s", cores=k, kind="port",
                sample=f"FP64 oracle force_calc(1) on {k} threads at once, each on its own {n}^3 FCC box "
                       f"({lat.natms} sites), 3-point rule, median of {reps} concurrent calls",
                single_thread=dict(value=one, cores=1,
                                   sample=f"the same call on one {n}^3 box, 1 thread, median of {reps}"))


def cpu_baseline_reference(tables, n=32, reps=3, procs=None):
    """The reference's own ForceField_DMD_I::force_calc, built from its sources
    (oracle/Makefile.ref) and driven by oracle/ref_harness: one process on one
    n^3 box, then K processes at once, each on its own copy (K ranks with no
    communication, as for the port); every process builds its system first and
    the timed calls start together.  value = K * sites / the median over reps of
    the slowest process's call."""
    import tempfile
    from mapp_amd import workload as W
    from oracle import ref_run as RR
    lat = W.fcc_alloy(n, seed=7)
    with tempfile.TemporaryDirectory() as tmp:
        one = RR.time_force(tmp, lat, tables, 1, reps)[0]
    k = procs or cpu_threads()
    with tempfile.TemporaryDirectory() as tmp:
        tk = RR.time_force(tmp, lat, tables, k, reps)
    wall = float(np.median(tk.max(axis=0)))
    return dict(value=k * lat.natms / wall, unit="site-evals/s", cores=k, kind="reference",
                sample=f"the reference's ForceField_DMD_I::force_calc(1) (built from its sources, -O2) in {k} "
                       f"processes at once, each on its own {n}^3 FCC box ({lat.natms} sites), 3-point rule, "
                       f"median of {reps} calls of the slowest process",
                single_thread=dict(value=lat.natms / float(np.median(one)), cores=1,
                                   sample=f"the same call in one process on one {n}^3 box, median of {reps}"))


def run_small(tables, n=10, reps=5):
    """The 4,000-site system (10^3 FCC cells), one energy+force+dE/dalpha+dc/dt evaluation:
    force_calc(1) and c_d_calc on the GPU and, unless no_cpu, on one host thread with the
    FP64 oracle; medians of reps calls."""
    from mapp_amd import workload as W
    from mapp_amd.dmd import setup
    lat = W.fcc_alloy(n, seed=1)
    ff = setup(lat, tables, no_i=3, precision="fp64")
    ff.force_calc(1)
    ff.create_2nd_neigh_lst()
    ff.c_d_calc()
    tf, tc = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        ff.force_calc(1)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        ff.c_d_calc()
        tc.append(time.perf_counter() - t0)
    ff.close()
    return dict(sites=lat.natms, force_ms=float(np.median(tf)) * 1e3, c_d_ms=float(np.median(tc)) * 1e3,
                site_evals_per_s=lat.natms / float(np.median(tf)), lat=lat)


def cpu_small(lat, tables, reps=5):
    """the same evaluation with the FP64 oracle on one host thread"""
    from oracle import oracle as O
    orc = O.make(lat, tables, no_i=3, variant="DMD_I")
    orc.force_calc(1)
    orc.create_2nd_neigh_lst()
    tf, tc = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        orc.force_calc(1)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        orc.c_d_calc()
        tc.append(time.perf_counter() - t0)
    return dict(cores=1, force_ms=float(np.median(tf)) * 1e3, c_d_ms=float(np.median(tc)) * 1e3,
                site_evals_per_s=lat.natms / float(np.median(tf)))


def copy_calibration(mib=1024, reps=5):
    """a plain device copy of known size, profiled with the same HBM counters as the kernels"""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4, "reps": reps}))


def traffic_for(workload):
    """HBM bytes per pass-1 launch from the committed counter summary, if it covers this workload"""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    return t.get(workload)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-lbfgs", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-meam", action="store_true")
    ap.add_argument("--meam-only", action="store_true", help="time MEAM force_calc alone (profiling)")
    ap.add_argument("--md-only", action="store_true", help="time the EAM, LJ, FS and FSM force_calc alone")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="train the surrogate and time force_calc with every cell served (profiling)")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--mixed-only", action="store_true", help="time the mixed (FP32 quadrature, 5-point) force_calc alone")
    ap.add_argument("--mobility", action="store_true", help="time c_d_calc and calc_g also with --fp64-only")
    ap.add_argument("--types", type=int, default=0,
                    help="time FP64 force_calc alone on an alloy of this many types (1-4: Ni, Cu, Pd, Ag)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--schedule", type=str, default=None,
                    help="fixed launch schedule 'parts,wpb,waves,wpb2,waves2' instead of the autotuner (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one, and this path keeps one rank's whole "
              f"domain on one GPU (requested --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.meam_only:
        print(json.dumps({"meam": run_meam(args.cells, args.steps, args.warmup)}), flush=True)
        return
    if args.md_only:
        from mapp_amd import workload as W
        print(json.dumps({"md": run_md(args.cells, W.johnson_tables(), args.steps, args.warmup)}), flush=True)
        return
    if args.mixed_only:
        from mapp_amd import workload as W
        lat = W.fcc_alloy(args.cells, seed=1)
        sched = None if args.schedule is None else [int(v) for v in args.schedule.split(",")]
        r = run_path(lat, W.johnson_tables(), 5, "mixed", args.steps, args.warmup, sched)
        print(json.dumps({"mixed_fp32_q5": {k: v for k, v in r.items() if k != "f"}}), flush=True)
        return
    if args.types:
        from mapp_amd import workload as W
        nt = args.types
        if not 1 <= nt <= 4:
            sys.exit("--types takes 1 to 4")
        lat = W.fcc_alloy(args.cells, nt=nt, seed=1)
        tables = W.johnson_tables(elements=("Ni", "Cu", "Pd", "Ag")[:nt])
        sched = None if args.schedule is None else [int(v) for v in args.schedule.split(",")]
        r = run_path(lat, tables, 3, "fp64", args.steps, args.warmup, sched, mobility=True)
        r["site_evals_per_s"] = lat.natms / (r["ms_per_step"] / 1e3)
        print(json.dumps({f"fp64_q3_{nt}_types": {k: v for k, v in r.items() if k != "f"}}), flush=True)
        return
    if args.surrogate_only:
        from mapp_amd import workload as W
        from mapp_amd.dmd import setup
        lat = W.fcc_alloy(args.cells, seed=1)
        ff = setup(lat, W.johnson_tables(), no_i=5, precision="mixed")
        ff.surrogate_train(n_per_tp=1 << 20, epochs=4, batch_per_tp=1024, lr=2e-3, seed=3, tol=1e30)
        ff.set_precision("surrogate")
        r = time_force(ff, args.steps, args.warmup)
        ff.close()
        print(json.dumps({"surrogate_all_served": {k: v for k, v in r.items()}}), flush=True)
        return

    from mapp_amd import workload as W
    tables = W.johnson_tables()
    lat = W.fcc_alloy(args.cells, seed=1)
    nt = lat.c.shape[1]
    natms = lat.natms
    log(f"system: {natms} sites")

    sched = None if args.schedule is None else [int(v) for v in args.schedule.split(",")]
    r64 = run_path(lat, tables, 3, "fp64", args.steps, args.warmup, sched, mobility=args.mobility or not args.fp64_only)
    log("fp64", {k: v for k, v in r64.items() if k != "f"})
    value = natms / (r64["ms_per_step"] / 1e3)
    P = r64["sizes"]["n_inrange"] / natms
    P_list = r64["sizes"]["npairs"] / natms
    names = ["pass1", "embed", "pass2", "gather_reduce"]
    kms = dict(zip(names, r64["kernel_ms"][:4]))
    pass1_ms = kms["pass1"]
    f1 = pass1_flops_per_site(P, nt, 3)
    achieved = f1 * natms / (pass1_ms / 1e3) / 1e12
    workload = f"DMD-I force_calc, FCC Ni-Cu {args.cells}^3 cells ({natms} sites), 3-point Gauss-Hermite, FP64"
    traffic = traffic_for(f"k_pass1 fp64 q3 {args.cells}^3 parts={r64['schedule'][0]}")
    out = {
        "metric": "DMD energy+grad site-evals/s at 1M sites: surrogate/FP32 vs FP64 quadrature",
        "value": value,
        "unit": "site-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "sites": natms, "pairs_half_list": r64["sizes"]["npairs"],
                   "pairs_in_quadrature_cut": r64["sizes"]["n_inrange"],
                   "schedule": r64["schedule"]},
        "roofline": {"bound": "valu", "pipe": "fp64 vector ALU (78.6 TFLOP/s); no MFMA in this kernel",
                     "limiter": "L2 latency of the table-row gathers (TA busy, pending L1-miss stalls)",
                     "kernel": "k_pass1", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None if traffic is None else traffic["bytes_per_launch"],
                     "flops_per_site": f1, "kernel_ms": pass1_ms,
                     "hbm_gbs": None if traffic is None else traffic["bytes_per_launch"] / (pass1_ms / 1e3) / 1e9},
        "kernel_ms": kms,
        "c_d_calc_ms": r64.get("c_d_ms"),
        "calc_g_ms": r64.get("calc_g_ms"),
        "c_d_calc_first_ms": r64.get("c_d_first_ms"),
        "list_rebuild_ms": r64.get("list_rebuild_ms"),
        "x_update_ms": r64.get("x_update_ms"),
        "site_flops": site_flops(P, nt, 3),
        "site_bytes": bytes_per_site(P_list, nt),
    }
    step_traffic = traffic_for(f"force_calc fp64 q3 {args.cells}^3 step")
    if step_traffic is not None:
        alg = bytes_per_site(P_list, nt) * natms
        out["roofline"]["traffic_step"] = step_traffic["bytes_per_launch"]
        out["roofline"]["traffic_step_kernels"] = step_traffic.get("kernels")
        out["roofline"]["algorithmic_step_bytes"] = alg
        out["roofline"]["traffic_step_over_algorithmic"] = step_traffic["bytes_per_launch"] / alg
    if not args.fp64_only:
        rmix = run_path(lat, tables, 5, "mixed", args.steps, args.warmup, sched, keep=True)
        r64q5 = run_path(lat, tables, 5, "fp64", max(2, args.steps // 4), 1, sched, keep=True)
        xd = 3 + nt

        def rel(a, b):
            return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

        f5 = r64q5["f"].reshape(natms, xd)
        fm = rmix["f"].reshape(natms, xd)
        paths = {
            "fp64_q3": {"site_evals_per_s": value, "ms_per_step": r64["ms_per_step"]},
            "fp64_q5": {"site_evals_per_s": natms / (r64q5["ms_per_step"] / 1e3), "ms_per_step": r64q5["ms_per_step"],
                        "kernel_ms": r64q5["kernel_ms"][:4], "schedule": r64q5["schedule"]},
            "mixed_fp32_q5": {"site_evals_per_s": natms / (rmix["ms_per_step"] / 1e3),
                              "ms_per_step": rmix["ms_per_step"], "kernel_ms": rmix["kernel_ms"][:4],
                              "schedule": rmix["schedule"],
                              "force_rel_err_vs_fp64_q5": rel(rmix["f"], r64q5["f"]),
                              "errors_vs_fp64_q5": {
                                  "energy": abs(rmix["energy"] - r64q5["energy"]) / abs(r64q5["energy"]),
                                  "force_xyz": rel(fm[:, :3], f5[:, :3]),
                                  "force_alpha": rel(fm[:, 3:], f5[:, 3:]),
                                  "E": rel(rmix["E"], r64q5["E"]), "dE": rel(rmix["dE"], r64q5["dE"]),
                                  "dc_dt": rel(rmix["c_d"], r64q5["c_d"])}},
        }
        if not args.no_surrogate:
            rs = run_surrogate(lat, tables, args.steps, args.warmup, 4 << 20, r64q5["f"], r64q5["energy"])
            log("surrogate", rs)
            paths["surrogate_q5_fallback"] = {
                "site_evals_per_s": natms / (rs["ms_per_step"] / 1e3), "ms_per_step": rs["ms_per_step"],
                "kernel_ms": rs["kernel_ms"][:4], "bound": rs["bound"],
                "force_rel_err_vs_fp64_q5": rs["force_rel_err_vs_fp64"],
                "energy_rel_err_vs_fp64_q5": rs["energy_rel_err_vs_fp64"],
                "served_evals": rs["served_evals"], "fallback_pair_evals": rs["fallbacks"],
                "tolerances": rs["tolerances"], "bound_scales": rs["scales"],
                "train_seconds": rs["train_seconds"], "train_report": rs["train_report"],
                "speedup_vs_fp64_q5": r64q5["ms_per_step"] / rs["ms_per_step"],
                "speedup_vs_mixed_q5": rmix["ms_per_step"] / rs["ms_per_step"],
                "all_fallback": rs["all_fallback"], "curve": rs["curve"]}
        # the same FP64 step on a three-type alloy (Ni-Cu-Pd, concentrations
        # drawn over the simplex): nine type pairs per list pair
        lat3 = W.fcc_alloy(args.cells, nt=3, seed=1)
        r3 = run_path(lat3, W.johnson_tables(elements=("Ni", "Cu", "Pd")), 3, "fp64", max(3, args.steps // 2), 1)
        paths["fp64_q3_three_types"] = {"site_evals_per_s": lat3.natms / (r3["ms_per_step"] / 1e3),
                                        "ms_per_step": r3["ms_per_step"], "kernel_ms": r3["kernel_ms"][:4],
                                        "schedule": r3["schedule"], "elements": ["Ni", "Cu", "Pd"]}
        del lat3, r3
        log("fp64 three types", paths["fp64_q3_three_types"])
        out["paths"] = paths
        if not args.no_lbfgs:
            out["lbfgs"] = run_lbfgs(lat, tables, 50)
            log("lbfgs", out["lbfgs"])
        if not args.no_batch:
            out["batch"] = run_batch(tables, max(3, args.steps // 4), 1)
            log("batch", out["batch"])
        if not args.no_meam:
            out["meam"] = run_meam(args.cells, max(3, args.steps // 4), 1)
            log("meam", out["meam"])
            out["md"] = run_md(args.cells, tables, max(3, args.steps // 4), 1)
            log("md", out["md"])
    small = run_small(tables)
    small_lat = small.pop("lat")
    out["small_4000"] = small
    if not args.no_cpu:
        out["small_4000"]["cpu_oracle"] = cpu_small(small_lat, tables)
        port = cpu_baseline(tables)
        port["single_thread_full_size"] = cpu_full_size(lat, tables)
        from oracle import ref_run as RR
        ref = None
        if RR.available():
            try:
                ref = cpu_baseline_reference(tables)
            except Exception as e:  # the build did not run here: the port stands in
                log("reference baseline unavailable:", e)
        if ref is not None:
            ref["port"] = port
            out["cpu_baseline"] = ref
        else:
            out["cpu_baseline"] = port
        log("cpu", out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

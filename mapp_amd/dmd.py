"""Python view of the C ABI (include/mapp_dmd.h) through ctypes.

ForceFieldDMD mirrors the reference's ForceField interface (src/ff.h:12-29):
coef / init / fin / force_calc / energy_calc / create_2nd_neigh_lst /
c_d_calc / calc_g, with the same argument meaning.  Errors raise DMDError
carrying the library's status code and message (the reference aborts).
The library must be built in-tree (mapp_amd/lib/libmapp_dmd.so); there is no
fallback path.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmapp_dmd.so")

VARIANT = {"DMD": 0, "DMD_I": 1, "DMD_II": 2, "EAM": 3, "LJ": 4, "FS": 5, "FSM": 6, "MEAM": 7}
MD_VARIANTS = ("EAM", "LJ", "FS", "FSM", "MEAM")  # x rows [3], c = type
PREC = {"fp64": 0, "mixed": 1, "surrogate": 2}
VEC = {"x": 0, "c": 1, "f": 2, "E": 3, "dE": 4, "n": 5, "c_d": 6, "ddE": 7, "s": 8, "t": 9, "v": 10, "a": 11, "g": 12}

_lib = None

D = C.c_double
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
LP = C.POINTER(C.c_long)
FP = C.POINTER(C.c_float)
VPP = C.POINTER(C.c_void_p)
CTX = C.c_void_p

_SIGS = {
    "mapp_dmd_create": [C.POINTER(CTX), C.c_int, C.c_int, C.c_int],
    "mapp_dmd_destroy": [CTX],
    "mapp_dmd_set_fs": [CTX, C.c_int, D, C.c_int, D, DP, DP, DP],
    "mapp_dmd_read_fs": [CTX, C.c_char_p, C.POINTER(C.c_char_p)],
    "mapp_dmd_coef": [CTX, D, C.c_int, D, D, D, DP],
    "mapp_dmd_set_precision": [CTX, C.c_int],
    "mapp_dmd_set_system": [CTX, DP, C.c_int, DP, DP],
    "mapp_dmd_init": [CTX, D],
    "mapp_dmd_fin": [CTX],
    "mapp_dmd_set_x": [CTX, DP, IP],
    "mapp_dmd_set_c": [CTX, DP],
    "mapp_dmd_force_calc": [CTX, C.c_int, DP],
    "mapp_dmd_energy_calc": [CTX, DP],
    "mapp_dmd_create_2nd_neigh_lst": [CTX],
    "mapp_dmd_c_d_calc": [CTX],
    "mapp_dmd_calc_g": [CTX, C.c_int, D, C.c_void_p, C.c_void_p, DP],
    "mapp_dmd_get_vec": [CTX, C.c_int, DP, C.c_long],
    "mapp_dmd_device_ptr": [CTX, C.c_int, VPP],
    "mapp_dmd_set_vec": [CTX, C.c_int, DP, C.c_long],
    "mapp_dmd_sync": [CTX],
    "mapp_dmd_sizes": [CTX, LP],
    "mapp_dmd_get_cutoffs": [CTX, DP],
    "mapp_dmd_get_list": [CTX, IP, IP],
    "mapp_dmd_get_list2": [CTX, IP, IP],
    "mapp_dmd_get_ghosts": [CTX, DP, IP],
    "mapp_dmd_get_gh": [CTX, DP, DP, DP, DP],
    "mapp_dmd_last_timing": [CTX, DP],
    "mapp_dmd_enable_timing": [CTX, C.c_int],
    "mapp_dmd_set_energy_mode": [CTX, C.c_int],
    "mapp_dmd_set_cells": [CTX, C.c_int, DP, IP, DP, DP],
    "mapp_dmd_cells": [CTX, IP, IP, IP, C.c_int],
    "mapp_dmd_autotune": [CTX, C.c_int, DP, C.c_int, IP],
    "mapp_dmd_get_schedule": [CTX, IP],
    "mapp_dmd_set_schedule": [CTX, IP],
    "mapp_dmd_bdf_run": [CTX, C.c_int, C.c_int, D, C.c_int, C.c_int, DP, LP],
    "mapp_dmd_min_lbfgs": [CTX, C.c_int, C.c_int, D, DP, LP],
    "mapp_dmd_surrogate_train": [CTX, C.c_long, C.c_int, C.c_int, D, C.c_ulonglong, D, DP, DP],
    "mapp_dmd_surrogate_gradient": [CTX, FP, FP, C.c_long, FP, FP, FP, FP],
    "mapp_dmd_surrogate_params": [CTX, FP, FP],
    "mapp_dmd_surrogate_fallbacks": [CTX, LP],
    "mapp_dmd_surrogate_set_bound": [CTX, D, DP],
    "mapp_dmd_surrogate_cell_bounds": [CTX, DP, C.c_long],
    "mapp_dmd_calc_g_noalloc": [CTX, C.c_int, D, C.c_void_p, C.c_void_p, DP],
    "mapp_dmd_lj_coef": [CTX, C.c_int, C.c_int, D, D, D],
    "mapp_dmd_fs_coef": [CTX, DP, DP, DP, DP, DP, DP, DP, DP],
    "mapp_dmd_meam_coef": [CTX, C.c_void_p],
    "mapp_dmd_read_meam_params": [CTX, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p)],
    "mapp_dmd_meam_set_fix": [CTX, C.c_int],
    "mapp_dmd_read_fs_params": [CTX, C.c_char_p, C.POINTER(C.c_char_p)],
}


_NOCTX_SIGS = {
    "mapp_dmd_cfg_read": [C.c_char_p, C.c_int, IP, IP, DP, DP, DP, DP, C.c_char_p, C.c_int, C.c_char_p, C.c_int],
    "mapp_dmd_cfg_write": [C.c_char_p, DP, C.c_int, C.c_int, DP, DP, C.c_int, DP, C.POINTER(C.c_char_p),
                           C.c_char_p, C.c_int],
    "mapp_dmd_meam_params_read": [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.c_void_p,
                                  C.c_char_p, C.c_int],
    "mapp_dmd_fs_params_read": [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), DP, DP, DP, DP, DP, DP, DP, DP,
                                C.c_char_p, C.c_int],
}


class DMDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mapp_dmd error {code}: {msg}")
        self.code = code


def load_library(path=LIB_PATH):
    """Load libmapp_dmd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m mapp_amd.build`")
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    for name, args in _NOCTX_SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mapp_dmd_last_error.argtypes = [CTX]
    lib.mapp_dmd_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(DP)


def _ip(a):
    return a.ctypes.data_as(IP)


def _fp(a):
    return a.ctypes.data_as(FP)


SUR_NPARAM = 8966
SUR_NNORM = 56
SUR_REPORT = ("loss", "val_mean_err", "val_max_err", "masked_frac", "steps", "seconds", "n_train", "n_val")


class MeamParams(C.Structure):
    """mapp_meam_params (include/mapp_dmd.h): ForceField_meam's members after its readers"""
    _fields_ = ([("nt", C.c_int)] +
                [(k, D * 4) for k in ("Z", "A", "rho0", "beta0", "beta1", "beta2", "beta3", "t0", "t1", "t2",
                                      "t3")] +
                [(k, C.c_int * 4) for k in ("ielt", "ibar")] +
                [(k, (C.c_int * 4) * 4) for k in ("lattice", "nn2", "zbl")] +
                [(k, (D * 4) * 4) for k in ("Ec", "alpha", "re", "delta", "attrac", "repuls", "ebound")] +
                [(k, ((D * 4) * 4) * 4) for k in ("cmin", "cmax")] +
                [(k, D) for k in ("rc", "delr", "delr_inv", "gsmooth")] +
                [(k, C.c_int) for k in ("augt1", "ialloy", "mix_ref_t", "erose_form", "emb_lin_neg",
                                        "bkgd_dyn")])

    @classmethod
    def from_state(cls, st):
        """from a dict of arrays with the struct's field names"""
        p = cls()
        for name, _ in cls._fields_:
            v = st[name]
            if isinstance(v, np.ndarray):
                np.ctypeslib.as_array(getattr(p, name)).reshape(-1)[:] = np.asarray(v).reshape(-1)
            else:
                setattr(p, name, v)
        return p

    def to_state(self):
        st = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            st[name] = np.ctypeslib.as_array(v).copy() if hasattr(v, "_length_") else v
        return st


class ForceFieldDMD:
    """ForceField_DMD / _I / _II on one GPU through the C ABI."""

    def __init__(self, no_types=2, variant="DMD_I", device=0):
        self.lib = load_library()
        self.nt = no_types
        self.variant = variant
        self.ctx = CTX()
        self._call("mapp_dmd_create", C.byref(self.ctx), no_types, VARIANT[variant], device, noctx=True)

    def _call(self, name, *args, noctx=False):
        fn = getattr(self.lib, name)
        rc = fn(*args) if noctx else fn(self.ctx, *args)
        if rc != 0:
            msg = self.lib.mapp_dmd_last_error(self.ctx).decode() if not noctx and self.ctx else "create failed"
            raise DMDError(rc, msg)

    def sync(self):
        """wait for the context's stream"""
        self._call("mapp_dmd_sync")

    def close(self):
        if self.ctx:
            self.lib.mapp_dmd_destroy(self.ctx)
            self.ctx = CTX()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up (ff_coef, read, init) --
    def set_fs(self, t):
        F = np.ascontiguousarray(t.F, dtype=np.float64)
        rho = np.ascontiguousarray(t.rho, dtype=np.float64)
        rphi = np.ascontiguousarray(t.rphi, dtype=np.float64)
        self._call("mapp_dmd_set_fs", t.nrho, t.drho, t.nr, t.dr, _dp(F), _dp(rho), _dp(rphi))

    def read_fs(self, path, elements):
        arr = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
        self._call("mapp_dmd_read_fs", path.encode(), arr)

    def coef(self, r_crd, no_i, alpha_min, alpha_max, T, masses):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        self._call("mapp_dmd_coef", r_crd, no_i, alpha_min, alpha_max, T, _dp(m))

    def lj_coef(self, itype, jtype, epsilon, sigma, cut):
        """ForceField_lj::coef: itype jtype epsilon sigma cutoff"""
        self._call("mapp_dmd_lj_coef", int(itype), int(jtype), float(epsilon), float(sigma), float(cut))

    def fs_coef(self, p):
        """ForceField_fs::coef (ForceField_fsm::coef for FSM) with the parameter file's values (FSParams, type order)"""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                for a in (p.A, p.t1, p.t2, p.cut_rho, p.cut_phi, p.k1, p.k2, p.k3)]
        self._call("mapp_dmd_fs_coef", *[_dp(a) for a in arrs])

    def meam_coef(self, state):
        """ForceField_meam::coef after its readers: state as mapp_amd.workload.meam_state"""
        self._meam = MeamParams.from_state(state)
        self._call("mapp_dmd_meam_coef", C.cast(C.pointer(self._meam), C.c_void_p))

    def read_meam_params(self, library, local, elements):
        """ff_coef MEAM <library> <local>"""
        arr = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
        self._call("mapp_dmd_read_meam_params", library.encode(), None if local is None else local.encode(), arr)

    def meam_set_fix(self, fix):
        self._call("mapp_dmd_meam_set_fix", 1 if fix else 0)

    def read_fs_params(self, path, elements):
        """ff_coef FS <file>: the parameter file of ForceField_fs::coef"""
        arr = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
        self._call("mapp_dmd_read_fs_params", path.encode(), arr)

    def set_precision(self, mode):
        self._call("mapp_dmd_set_precision", PREC[mode])

    def set_system(self, H, x_reduced, c):
        H = np.ascontiguousarray(H, dtype=np.float64)
        x = np.ascontiguousarray(x_reduced, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        self._call("mapp_dmd_set_system", _dp(H), x.shape[0], _dp(x), _dp(c))

    def init(self, skin=0.5):
        self._call("mapp_dmd_init", skin)

    def fin(self):
        self._call("mapp_dmd_fin")

    def set_x(self, x):
        """Atoms::update_0 after a position change; x = None takes the positions
        already on the device (set_vec("x", ...) or a device-resident driver)"""
        flag = C.c_int(0)
        if x is None:
            self._call("mapp_dmd_set_x", None, C.byref(flag))
        else:
            x = np.ascontiguousarray(x, dtype=np.float64)
            self._call("mapp_dmd_set_x", _dp(x), C.byref(flag))
        return bool(flag.value)

    def set_c(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64)
        self._call("mapp_dmd_set_c", _dp(c))

    # -- the ForceField interface --
    def force_calc(self, st_clc=1):
        en = np.zeros(7)
        self._call("mapp_dmd_force_calc", int(st_clc), _dp(en))
        return en

    def energy_calc(self):
        e = C.c_double(0.0)
        self._call("mapp_dmd_energy_calc", C.byref(e))
        return e.value

    def create_2nd_neigh_lst(self):
        self._call("mapp_dmd_create_2nd_neigh_lst")

    def c_d_calc(self):
        self._call("mapp_dmd_c_d_calc")

    def calc_g(self, chk, alpha, a_dev_ptr=None, g_dev_ptr=None):
        """calc_g on device vectors; by default the context's own A and G buffers"""
        a_dev_ptr = self.device_ptr("a") if a_dev_ptr is None else a_dev_ptr
        g_dev_ptr = self.device_ptr("g") if g_dev_ptr is None else g_dev_ptr
        cost = C.c_double(0.0)
        self._call("mapp_dmd_calc_g", int(chk), float(alpha), C.c_void_p(a_dev_ptr), C.c_void_p(g_dev_ptr),
                   C.byref(cost))
        return cost.value

    def calc_g_noalloc(self, chk, alpha, a_dev_ptr=None, g_dev_ptr=None):
        """calc_g with the library's static scratch (no device allocation); the pair values must be
        current (a c_d_calc or calc_g since the last position change)"""
        a_dev_ptr = self.device_ptr("a") if a_dev_ptr is None else a_dev_ptr
        g_dev_ptr = self.device_ptr("g") if g_dev_ptr is None else g_dev_ptr
        cost = C.c_double(0.0)
        self._call("mapp_dmd_calc_g_noalloc", int(chk), float(alpha), C.c_void_p(a_dev_ptr), C.c_void_p(g_dev_ptr),
                   C.byref(cost))
        return cost.value

    def set_vec(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._call("mapp_dmd_set_vec", VEC[name], _dp(a), a.size)

    # -- data access --
    def sizes(self):
        out = np.zeros(6, dtype=np.int64)
        self._call("mapp_dmd_sizes", out.ctypes.data_as(LP))
        return dict(natms=int(out[0]), nghost=int(out[1]), npairs=int(out[2]), npairs2=int(out[3]),
                    nswap=int(out[4]), n_inrange=int(out[5]))

    def get(self, name):
        s = self.sizes()
        md = self.variant in MD_VARIANTS
        width = (3 if md else 3 + self.nt) if name in ("x", "f") else (1 if md else self.nt)
        out = np.zeros((s["natms"], width))
        self._call("mapp_dmd_get_vec", VEC[name], _dp(out), out.size)
        return out

    def device_ptr(self, name):
        p = C.c_void_p()
        self._call("mapp_dmd_device_ptr", VEC[name], C.byref(p))
        return p.value

    def cutoffs(self):
        """(rc, mod_rc, r_crd) after coef"""
        out = np.zeros(3)
        self._call("mapp_dmd_get_cutoffs", _dp(out))
        return tuple(float(v) for v in out)

    def neighbor_list(self):
        s = self.sizes()
        off = np.zeros(s["natms"] + 1, dtype=np.int32)
        nl = np.zeros(max(s["npairs"], 1), dtype=np.int32)
        self._call("mapp_dmd_get_list", _ip(off), _ip(nl))
        return off, nl[: s["npairs"]]

    def neighbor_list2(self):
        s = self.sizes()
        off = np.zeros(s["natms"] + 1, dtype=np.int32)
        nl = np.zeros(max(s["npairs2"], 1), dtype=np.int32)
        self._call("mapp_dmd_get_list2", _ip(off), _ip(nl))
        return off, nl[: s["npairs2"]]

    def ghosts(self):
        s = self.sizes()
        x = np.zeros((max(s["nghost"], 1), 3 if self.variant in MD_VARIANTS else 3 + self.nt))
        src = np.zeros(max(s["nghost"], 1), dtype=np.int32)
        self._call("mapp_dmd_get_ghosts", _dp(x), _ip(src))
        return x[: s["nghost"]], src[: s["nghost"]]

    def gh(self, no_i):
        arr = [np.zeros(no_i) for _ in range(4)]
        self._call("mapp_dmd_get_gh", *[_dp(a) for a in arr])
        return arr

    def enable_timing(self, on=True):
        self._call("mapp_dmd_enable_timing", 1 if on else 0)

    def last_timing(self):
        ms = np.zeros(5)
        self._call("mapp_dmd_last_timing", _dp(ms))
        return ms

    def set_energy_mode(self, full):
        """False: the reference's energy_calc type loop; True: every ordered type pair"""
        self._call("mapp_dmd_set_energy_mode", 1 if full else 0)

    def min_lbfgs(self, max_iter, m_it=2, e_tol=1e-8):
        """Min_LBFGS::run on the device: (exit code, energies per step, force and energy call counts)"""
        en = np.zeros(max_iter + 2)
        cnt = np.zeros(4, dtype=np.int64)
        self._call("mapp_dmd_min_lbfgs", int(max_iter), int(m_it), float(e_tol), _dp(en),
                   cnt.ctypes.data_as(LP))
        return int(cnt[3]), en[: cnt[0] + 2], int(cnt[1]), int(cnt[2])

    def set_cells(self, Hs, xs, cs):
        """a batch of independent cells: lists of boxes, reduced positions + alpha, concentrations"""
        H = np.ascontiguousarray(np.stack([np.asarray(h, dtype=np.float64) for h in Hs]))
        n = np.ascontiguousarray([len(x) for x in xs], dtype=np.int32)
        x = np.ascontiguousarray(np.concatenate(xs), dtype=np.float64)
        c = np.ascontiguousarray(np.concatenate(cs), dtype=np.float64)
        self._call("mapp_dmd_set_cells", len(xs), _dp(H), _ip(n), _dp(x), _dp(c))

    def cells(self):
        nc = C.c_int(0)
        self._call("mapp_dmd_cells", C.byref(nc), None, None, 0)
        na = np.zeros(nc.value, dtype=np.int32)
        ng = np.zeros(nc.value, dtype=np.int32)
        self._call("mapp_dmd_cells", C.byref(nc), _ip(na), _ip(ng), nc.value)
        return na, ng

    def autotune(self, reps=2):
        """time the candidate launch schedules of the force passes; keeps the fastest.
        Returns rows {lanes per pair, waves per block, total waves (0: one per site), pass-1 ms, pass-2 ms}."""
        rep = np.zeros((64, 5))
        n = C.c_int(0)
        self._call("mapp_dmd_autotune", int(reps), _dp(rep), 64, C.byref(n))
        return rep[: n.value]

    def schedule(self):
        """[pass-1 lanes per pair, pass-1 waves per block, pass-1 waves, pass-2 waves per block, pass-2 waves]"""
        out = np.zeros(5, dtype=np.int32)
        self._call("mapp_dmd_get_schedule", _ip(out))
        return out.tolist()

    def set_schedule(self, sched):
        arr = np.ascontiguousarray(sched, dtype=np.int32)
        self._call("mapp_dmd_set_schedule", _ip(arr))

    def bdf_run(self, nsteps, max_order=2, initial_del_t=-1.0, max_retry=100, fix=False):
        """Clock_BDF::init + run on the device: (return code, report [steps][order, del_t, cost, err, iters,
        solves], calc_g calls); the code is the number of steps or -1 / -2 as the C ABI documents"""
        rep = np.zeros((max(nsteps, 1), 6))
        cnt = np.zeros(2, dtype=np.int64)
        self._call("mapp_dmd_bdf_run", int(nsteps), int(max_order), float(initial_del_t), int(max_retry),
                   1 if fix else 0, _dp(rep), cnt.ctypes.data_as(LP))
        code = int(cnt[0])
        return code, rep[:code] if code >= 0 else rep, int(cnt[1])

    # learned surrogate of the pair quadrature (two types)
    def surrogate_train(self, n_per_tp=1 << 20, epochs=4, batch_per_tp=256, lr=2e-3, seed=1, tol=0.05, box=None):
        rep = np.zeros(8)
        b = None if box is None else _dp(np.ascontiguousarray(box, dtype=np.float64))
        self._call("mapp_dmd_surrogate_train", int(n_per_tp), int(epochs), int(batch_per_tp), float(lr), int(seed),
                   float(tol), b, _dp(rep))
        return dict(zip(SUR_REPORT, rep.tolist()))

    def surrogate_gradient(self, params, norm, X, T):
        """gradient of each network's mean squared standardised error by the training kernel"""
        params = np.ascontiguousarray(params, dtype=np.float32).reshape(4, SUR_NPARAM)
        norm = np.ascontiguousarray(norm, dtype=np.float32).reshape(SUR_NNORM)
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, 6)
        n = X.shape[0] // 4
        grad = np.zeros((4, SUR_NPARAM), dtype=np.float32)
        loss = np.zeros(4, dtype=np.float32)
        self._call("mapp_dmd_surrogate_gradient", _fp(params), _fp(norm), n, _fp(X), _fp(T), _fp(grad), _fp(loss))
        return grad, loss

    def surrogate_params(self):
        params = np.zeros((4, SUR_NPARAM), dtype=np.float32)
        norm = np.zeros(SUR_NNORM, dtype=np.float32)
        self._call("mapp_dmd_surrogate_params", _fp(params), _fp(norm))
        return params, norm

    def surrogate_set_bound(self, bound):
        """re-derive the fallback mask for a relative force bound: dict served (cell fraction), eps (six
        per-pair tolerances), bound, scales (F, F', F'', N, max|drho/dr|, max|drho/dalpha|, |U|, natms)"""
        rep = np.zeros(16)
        self._call("mapp_dmd_surrogate_set_bound", float(bound), _dp(rep))
        return dict(served=rep[0], eps=rep[1:7].copy(), bound=rep[7], scales=rep[8:16].copy())

    def surrogate_cell_bounds(self):
        """smallest relative force bound at which each cell is served: [type pair][r cell][alpha_i cell][alpha_j cell]"""
        out = np.zeros(4 * 64 * 4 * 4)
        self._call("mapp_dmd_surrogate_cell_bounds", _dp(out), out.size)
        return out.reshape(4, 64, 4, 4)

    def surrogate_fallbacks(self):
        n = C.c_long(0)
        self._call("mapp_dmd_surrogate_fallbacks", C.byref(n))
        return n.value


def read_cfg(path):
    """extended CFG of a DMD run -> (H lower triangular, x [natms][3+nt] reduced + alpha, c, masses, names)"""
    lib = load_library()
    natms, nt = C.c_int(0), C.c_int(0)
    err = C.create_string_buffer(512)
    rc = lib.mapp_dmd_cfg_read(path.encode(), 0, C.byref(natms), C.byref(nt), None, None, None, None, None, 0, err, 512)
    if rc != 0:
        raise DMDError(rc, err.value.decode())
    n, t = natms.value, nt.value
    H = np.zeros((3, 3))
    x = np.zeros((n, 3 + t))
    c = np.zeros((n, t))
    mass = np.zeros(t)
    names = C.create_string_buffer(4096)
    rc = lib.mapp_dmd_cfg_read(path.encode(), n, C.byref(natms), C.byref(nt), _dp(H), _dp(x), _dp(c), _dp(mass),
                               names, 4096, err, 512)
    if rc != 0:
        raise DMDError(rc, err.value.decode())
    nm = names.value.decode().split("\n") if names.value else []
    return H, x, c, mass[: len(nm)], nm


def write_cfg(path, H, x, c, masses, names):
    """write an extended CFG (x: reduced positions + alpha) with %.17g"""
    lib = load_library()
    H = np.ascontiguousarray(H, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    m = np.ascontiguousarray(masses, dtype=np.float64)
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    err = C.create_string_buffer(512)
    rc = lib.mapp_dmd_cfg_write(path.encode(), _dp(H), x.shape[0], c.shape[1], _dp(x), _dp(c), len(names), _dp(m),
                                arr, err, 512)
    if rc != 0:
        raise DMDError(rc, err.value.decode())


def setup_cells(lats, tables, no_i=3, variant="DMD_I", precision="fp64", skin=0.5, r_crd=3.0,
                alpha_min=25.0, alpha_max=1000.0, T=600.0, device=0):
    """a batch of independent cells (workload.Lattice each) in one context"""
    nt = lats[0].c.shape[1]
    ff = ForceFieldDMD(no_types=nt, variant=variant, device=device)
    ff.set_fs(tables)
    ff.coef(r_crd, no_i, alpha_min, alpha_max, T, tables.masses[:nt])
    ff.set_precision(precision)
    ff.set_cells([l.H for l in lats], [l.x for l in lats], [l.c for l in lats])
    ff.init(skin)
    return ff


def setup_eam(H, s3, types, tables, skin=0.5, device=0):
    """plain EAM (ForceField_eam): reduced positions [natms][3] and integer types"""
    ff = ForceFieldDMD(no_types=len(tables.elements), variant="EAM", device=device)
    ff.set_fs(tables)
    ff.coef(0.0, 1, 1.0, 1.0, 1.0, np.ones(len(tables.elements)))
    ff.set_system(H, s3, np.asarray(types, dtype=np.float64).reshape(-1, 1))
    ff.init(skin)
    return ff


def setup_lj(H, s3, types, coefs, nt, skin=0.5, device=0):
    """Lennard-Jones (ForceField_lj): coefs = [(itype, jtype, epsilon, sigma, cutoff), ...]"""
    ff = ForceFieldDMD(no_types=nt, variant="LJ", device=device)
    for c in coefs:
        ff.lj_coef(*c)
    ff.set_system(H, s3, np.asarray(types, dtype=np.float64).reshape(-1, 1))
    ff.init(skin)
    return ff


def setup_fs(H, s3, types, params, skin=0.5, device=0, variant="FS"):
    """Finnis-Sinclair with FSParams: ForceField_fs, or ForceField_fsm with variant FSM"""
    assert variant in ("FS", "FSM")
    ff = ForceFieldDMD(no_types=len(params.A), variant=variant, device=device)
    ff.fs_coef(params)
    ff.set_system(H, s3, np.asarray(types, dtype=np.float64).reshape(-1, 1))
    ff.init(skin)
    return ff


def setup_meam(H, s3, types, state, skin=0.3, fix=False, device=0):
    """MEAM (ForceField_meam) with the parameters of mapp_amd.workload.meam_state"""
    ff = ForceFieldDMD(no_types=int(state["nt"]), variant="MEAM", device=device)
    ff.meam_coef(state)
    ff.meam_set_fix(fix)
    ff.set_system(H, s3, np.asarray(types, dtype=np.float64).reshape(-1, 1))
    ff.init(skin)
    return ff


def meam_params_read(library, local, elements):
    """the library and local files of ForceField_meam::coef for the types
    `elements`, without a GPU: a dict of arrays (mapp_meam_params fields)"""
    lib = load_library()
    p = MeamParams()
    err = C.create_string_buffer(256)
    arr = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
    rc = lib.mapp_dmd_meam_params_read(library.encode(), None if local is None else local.encode(), len(elements),
                                       arr, C.cast(C.pointer(p), C.c_void_p), err, 256)
    if rc != 0:
        raise DMDError(rc, err.value.decode())
    return p.to_state()


def read_fs_params(path, elements):
    """the FS parameter file (ForceField_fs::coef layout) in elements' type order, without a GPU"""
    from .workload import FSParams
    lib = load_library()
    nt = len(elements)
    nc = nt * (nt + 1) // 2
    A, t1, t2 = np.zeros(nt), np.zeros(nt * nt), np.zeros(nt * nt)
    per = [np.zeros(nc) for _ in range(5)]
    err = C.create_string_buffer(256)
    arr = (C.c_char_p * nt)(*[e.encode() for e in elements])
    rc = lib.mapp_dmd_fs_params_read(path.encode(), nt, arr, _dp(A), _dp(t1), _dp(t2), *[_dp(a) for a in per],
                                     err, 256)
    if rc != 0:
        raise DMDError(rc, err.value.decode())
    return FSParams(elements=list(elements), A=A, t1=t1.reshape(nt, nt), t2=t2.reshape(nt, nt), cut_rho=per[0],
                    cut_phi=per[1], k1=per[2], k2=per[3], k3=per[4])


def setup(lat, tables, no_i=3, variant="DMD_I", precision="fp64", skin=0.5, r_crd=3.0,
          alpha_min=25.0, alpha_max=1000.0, T=600.0, device=0):
    """ff + ff_coef + read + init in the order of the reference's scripts."""
    ff = ForceFieldDMD(no_types=lat.c.shape[1], variant=variant, device=device)
    ff.set_fs(tables)
    ff.coef(r_crd, no_i, alpha_min, alpha_max, T, tables.masses[: lat.c.shape[1]])
    ff.set_precision(precision)
    ff.set_system(lat.H, lat.x, lat.c)
    ff.init(skin)
    return ff

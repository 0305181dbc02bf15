"""Python host layer over the MI355X stream-collide C ABI
(include/lbm_mi355x.h, lib/liblbm_mi355x.so).

It mirrors the host side of the reference's time loop (src/lbgpu_prime.cu:
construct_lattice, the odd/even half-steps with bcs_1/bcs_2, the frame
transfers and the compute.c reductions) for callers and tests written in
Python.  Every compute call goes to the HIP library; there is no CPU path:
if the library is not built this module raises on import.

The directory name is not a Python identifier; load it with
``load_package()`` from the repository root or importlib.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liblbm_mi355x.so")

STORE_F64, STORE_F32, STORE_F16 = 0, 1, 2
OBJECTIVES = {"flux": 0, "permeability": 1}
# lbm_zface_flow / lbm_zface_conc
ZFACE_DEFAULT, ZFACE_NONE, ZFACE_PRESSURE, ZFACE_VELOCITY = 0, 1, 2, 3
ZCONC_DEFAULT, ZCONC_NONE, ZCONC_DIRICHLET, ZCONC_ZEROGRAD = 0, 1, 2, 3
COMPUTE_F64, COMPUTE_F32 = 0, 1
REDUCE_REF, REDUCE_SUMS = 0, 1
MACRO_NONE, MACRO_FRAME, MACRO_PEST = 0, 1, 2
PROFILE_CLASSES = ("odd", "even", "bc", "halo", "flux", "adjoint")


class LbmError(RuntimeError):
    pass


class Health(ctypes.Structure):
    """lbm_health: what lbm_check measured"""
    _fields_ = [
        ("mass", ctypes.c_double * 2), ("min_rho", ctypes.c_double), ("min_pop", ctypes.c_double * 2),
        ("nonfinite", ctypes.c_long * 2), ("negative", ctypes.c_long * 2),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int), ("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int),
        ("nsubs", ctypes.c_int), ("halo_mode", ctypes.c_int),
        ("tau", ctypes.c_double * 2), ("rho_A", ctypes.c_double * 2),
        ("gaccel", ctypes.c_double * 6), ("sink", ctypes.c_double),
        ("sigma_t_on", ctypes.c_int), ("sigma_t_off", ctypes.c_int),
        ("pressure_n_in", ctypes.c_int * 2), ("pressure_n_out", ctypes.c_int * 2),
        ("pressure_s_in", ctypes.c_int * 2), ("pressure_s_out", ctypes.c_int * 2),
        ("zeroconcgrad_n", ctypes.c_int), ("zeroconcgrad_s", ctypes.c_int),
        ("rho_in", ctypes.c_double), ("rho_out", ctypes.c_double),
        ("pfadd", ctypes.c_double), ("pfmul", ctypes.c_double),
        ("bc_ramp_start", ctypes.c_int), ("bc_ramp_stop", ctypes.c_int),
        ("pressure_n_series", ctypes.POINTER(ctypes.c_double) * 2),
        ("pressure_n_series_len", ctypes.c_int * 2),
        ("pressure_s_series", ctypes.POINTER(ctypes.c_double) * 2),
        ("pressure_s_series_len", ctypes.c_int * 2),
        ("zface_pressure", ctypes.c_int), ("rho_bot", ctypes.c_double), ("rho_top", ctypes.c_double),
        ("zface_solute", ctypes.c_int), ("c_bot", ctypes.c_double),
        ("c_bot_map", ctypes.POINTER(ctypes.c_double)),
        ("flux_plane", ctypes.c_int),
        ("storage", ctypes.c_int), ("compute", ctypes.c_int),
        ("store_scale", ctypes.c_double * 2), ("store_shift_c", ctypes.c_double),
        ("velocity_n_in", ctypes.c_int), ("velocity_n_out", ctypes.c_int),
        ("velocity_s_in", ctypes.c_int), ("velocity_s_out", ctypes.c_int),
        ("uy_in", ctypes.c_double), ("uy_out", ctypes.c_double),
        ("zface_flow", ctypes.c_int * 2), ("zface_uz", ctypes.c_double * 2),
        ("zface_conc", ctypes.c_int * 2), ("c_top", ctypes.c_double),
        ("halo_literal", ctypes.c_int),
    ]


_LIB = None


def lib():
    """The loaded C ABI library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LbmError(f"HIP library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        lp = ctypes.POINTER(ctypes.c_long)
        sig = {
            "lbm_create": [ctypes.POINTER(Params), ctypes.POINTER(vp)],
            "lbm_destroy": [vp],
            "lbm_set_solids": [vp, u8p],
            "lbm_upload_f": [vp, ip, dp],
            "lbm_download_f": [vp, ip, dp],
            "lbm_download_f_noalloc": [vp, ip, dp],
            "lbm_init_rest": [vp, dp],
            "lbm_set_time": [vp, ip],
            "lbm_get_time": [vp, ctypes.POINTER(ctypes.c_int)],
            "lbm_step": [vp, ip],
            "lbm_run": [vp, ip, ip],
            "lbm_step_odd": [vp, ip, ip],
            "lbm_step_even": [vp, ip, ip],
            "lbm_apply_bcs": [vp, ip, ip],
            "lbm_halo_bytes": [vp, lp],
            "lbm_halo_pack": [vp, ip, vp, vp],
            "lbm_halo_unpack": [vp, ip, vp, vp],
            "lbm_halo_periodic": [vp, ip],
            "lbm_download_macro": [vp, ip, dp, dp, dp, dp],
            "lbm_upload_macro": [vp, ip, dp, dp, dp, dp],
            "lbm_set_ns": [vp, dp],
            "lbm_objective_select": [vp, ip],
            "lbm_objective_keep_states": [vp, ctypes.c_size_t],
            "lbm_objective_info": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
            "lbm_sample_macro": [vp, ip, ip, ctypes.POINTER(ctypes.c_long), dp],
            "lbm_reduce": [vp, ip, ip, dp],
            "lbm_last_flux": [vp, dp, dp],
            "lbm_check": [vp, dp, ctypes.c_double, ctypes.POINTER(Health)],
            "lbm_info": [vp, lp, lp, lp],
            "lbm_profile_enable": [vp, ip],
            "lbm_profile_read": [vp, dp, ctypes.POINTER(ctypes.c_int)],
            "lbm_copy_bandwidth": [ctypes.c_long, ip, dp],
            "lbm_objective_run": [vp, ip, ip, ip, dp],
            "lbm_objective_observations": [vp, ip, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int),
                                           dp, dp],
            "lbm_adjoint_run": [vp, dp, dp],
            "lbm_geom_discs2d": [ip, ip, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ip, u8p, dp],
            "lbm_geom_spheres3d": [ip, ip, ip, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ip, u8p, dp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.lbm_last_error.restype = ctypes.c_char_p
        L.lbm_last_error.argtypes = []
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise LbmError(lib().lbm_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(**kw):
    p = Params()
    p.dims = kw.get("dims", 3)
    p.ni, p.nj = kw["ni"], kw["nj"]
    p.nk = kw.get("nk", 1)
    p.nsubs = kw.get("nsubs", 1)
    p.halo_mode = kw.get("halo_mode", 0)
    tau, rho_A = kw.get("tau", (1.0, 1.0)), kw.get("rho_A", (1.0, 1.0))
    g = tuple(kw.get("gaccel", (0.0,) * 6)) + (0.0,) * 6
    for q in range(2):
        p.tau[q], p.rho_A[q] = tau[q], rho_A[q]
    for q in range(6):
        p.gaccel[q] = g[q]
    p.sink = kw.get("sink", 0.0)
    p.sigma_t_on = kw.get("sigma_t_on", 0)
    p.sigma_t_off = kw.get("sigma_t_off", 2**31 - 1)
    for name in ("pressure_n_in", "pressure_n_out", "pressure_s_in", "pressure_s_out"):
        v = kw.get(name, (0, 0))
        arr = getattr(p, name)
        arr[0], arr[1] = v[0], v[1]
    p.zeroconcgrad_n = kw.get("zeroconcgrad_n", 0)
    p.zeroconcgrad_s = kw.get("zeroconcgrad_s", 0)
    p.rho_in, p.rho_out = kw.get("rho_in", 1.0), kw.get("rho_out", 1.0)
    p.pfadd, p.pfmul = kw.get("pfadd", 0.0), kw.get("pfmul", 1.0)
    p.bc_ramp_start, p.bc_ramp_stop = kw.get("bc_ramp_start", -1), kw.get("bc_ramp_stop", -1)
    p.zface_pressure = kw.get("zface_pressure", 0)
    p.rho_bot, p.rho_top = kw.get("rho_bot", 1.0), kw.get("rho_top", 1.0)
    p.zface_solute = kw.get("zface_solute", 0)
    p.c_bot = kw.get("c_bot", 1.0)
    p.flux_plane = kw.get("flux_plane", -1)
    p.storage = kw.get("storage", STORE_F64)
    p.compute = kw.get("compute", COMPUTE_F64 if p.storage == STORE_F64 else COMPUTE_F32)
    sc = kw.get("store_scale", (0.0, 0.0))
    p.store_scale[0], p.store_scale[1] = sc[0], sc[1]
    p.store_shift_c = kw.get("store_shift_c", 0.0)
    for name in ("velocity_n_in", "velocity_n_out", "velocity_s_in", "velocity_s_out"):
        setattr(p, name, kw.get(name, 0))
    p.uy_in, p.uy_out = kw.get("uy_in", 0.0), kw.get("uy_out", 0.0)
    for name in ("zface_flow", "zface_uz", "zface_conc"):
        v = kw.get(name, (0, 0))
        arr = getattr(p, name)
        arr[0], arr[1] = v[0], v[1]
    p.c_top = kw.get("c_top", 0.0)
    p.halo_literal = kw.get("halo_literal", 0)
    return p


class Lattice:
    """One lattice on the GPU (lbm_create + lbm_set_solids)."""

    def __init__(self, mask=None, **kw):
        self.kw = dict(kw)
        self.p = make_params(**kw)
        keep = []
        for name in ("pressure_n_series", "pressure_s_series"):
            ser = kw.get(name)
            if ser is not None:
                for q, s in enumerate(ser):
                    if s is None:
                        continue
                    arr = np.ascontiguousarray(s, dtype=np.float64)
                    keep.append(arr)
                    getattr(self.p, name)[q] = _dp(arr)
                    getattr(self.p, name + "_len")[q] = arr.size
        cmap = kw.get("c_bot_map")
        if cmap is not None:
            arr = np.ascontiguousarray(cmap, dtype=np.float64).ravel()
            keep.append(arr)
            self.p.c_bot_map = _dp(arr)
        self.dims = self.p.dims
        self.Q = 9 if self.dims == 2 else 19
        self.N = self.p.ni * self.p.nj * (self.p.nk if self.dims == 3 else 1)
        self.nsubs = self.p.nsubs
        h = ctypes.c_void_p()
        _check(lib().lbm_create(ctypes.byref(self.p), ctypes.byref(h)))
        self.h = h
        if mask is not None:
            self.set_solids(mask)

    def close(self):
        if getattr(self, "h", None):
            lib().lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.p.nj, self.p.ni) if self.dims == 2 else (self.p.nk, self.p.nj, self.p.ni)

    def set_solids(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        if m.size != self.N:
            raise LbmError("mask size does not match the lattice")
        _check(lib().lbm_set_solids(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def info(self):
        n, nf, s = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        _check(lib().lbm_info(self.h, ctypes.byref(n), ctypes.byref(nf), ctypes.byref(s)))
        return {"nodes": n.value, "fluid": nf.value, "row_stride": s.value}

    def upload_f(self, subs, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.Q, self.N)
        _check(lib().lbm_upload_f(self.h, subs, _dp(a)))

    def download_f(self, subs):
        a = np.empty((self.Q, self.N))
        _check(lib().lbm_download_f(self.h, subs, _dp(a)))
        return a

    def download_f_noalloc(self, subs):
        """lbm_download_f_noalloc: as download_f, staged in the library's
        static device array (no allocation)"""
        a = np.empty((self.Q, self.N))
        _check(lib().lbm_download_f_noalloc(self.h, subs, _dp(a)))
        return a

    def init_rest(self, c0=None):
        if c0 is None:
            _check(lib().lbm_init_rest(self.h, None))
        else:
            c = np.ascontiguousarray(c0, dtype=np.float64).ravel()
            _check(lib().lbm_init_rest(self.h, _dp(c)))

    @property
    def time(self):
        t = ctypes.c_int()
        _check(lib().lbm_get_time(self.h, ctypes.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        _check(lib().lbm_set_time(self.h, int(t)))

    def step(self, write_mv=False):
        _check(lib().lbm_step(self.h, int(write_mv)))

    def run(self, n, write_last=True):
        _check(lib().lbm_run(self.h, int(n), int(write_last)))

    # the reference's pieces of one step pair
    def stream_collide_stream(self, time, write_mv=False):
        _check(lib().lbm_step_odd(self.h, int(time), int(write_mv)))

    def collide(self, time, write_mv=False):
        _check(lib().lbm_step_even(self.h, int(time), int(write_mv)))

    def bcs_1(self, time):
        _check(lib().lbm_apply_bcs(self.h, 1, int(time)))

    def bcs_2(self, time):
        _check(lib().lbm_apply_bcs(self.h, 2, int(time)))

    def halo_bytes(self):
        b = ctypes.c_long()
        _check(lib().lbm_halo_bytes(self.h, ctypes.byref(b)))
        return b.value

    def halo_pack(self, phase, pos_ptr, neg_ptr):
        _check(lib().lbm_halo_pack(self.h, phase, ctypes.c_void_p(pos_ptr), ctypes.c_void_p(neg_ptr)))

    def halo_unpack(self, phase, pos_ptr, neg_ptr):
        _check(lib().lbm_halo_unpack(self.h, phase, ctypes.c_void_p(pos_ptr), ctypes.c_void_p(neg_ptr)))

    def halo_periodic(self, phase):
        _check(lib().lbm_halo_periodic(self.h, phase))

    def mv(self, subs=0):
        rho, ux, uy, uz = (np.empty(self.N) for _ in range(4))
        _check(lib().lbm_download_macro(self.h, subs, _dp(rho), _dp(ux), _dp(uy), _dp(uz)))
        return rho, ux, uy, uz

    def set_ns(self, ns):
        """lbm_set_ns: Walsh partial bounce-back fraction per node (None: off)"""
        if ns is None:
            _check(lib().lbm_set_ns(self.h, None))
            return
        a = np.ascontiguousarray(ns, dtype=np.float64).reshape(-1)
        if a.size != self.N:
            raise ValueError("ns holds one value per lattice node")
        _check(lib().lbm_set_ns(self.h, _dp(a)))

    def sample_mv(self, subs, nodes):
        """lbm_sample_macro: [len(nodes), 1 + dims] values of mv at dense nodes"""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64).reshape(-1)
        out = np.empty((nodes.size, 1 + self.dims))
        _check(lib().lbm_sample_macro(self.h, subs, nodes.size,
                                      nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _dp(out)))
        return out

    def set_mv(self, subs, rho=None, ux=None, uy=None, uz=None):
        """lbm_upload_macro; None leaves a component as it is"""
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, uz)]
        for a in arr:
            if a is not None and a.size != self.N:
                raise ValueError("macroscopic arrays hold one value per lattice node")
        ptr = [None if a is None else _dp(a) for a in arr]
        _check(lib().lbm_upload_macro(self.h, subs, *ptr))

    def reduce_ref(self, subs=0):
        out = np.empty(12)
        _check(lib().lbm_reduce(self.h, REDUCE_REF, subs, _dp(out)))
        return out

    def reduce_sums(self):
        out = np.empty(6)
        _check(lib().lbm_reduce(self.h, REDUCE_SUMS, 0, _dp(out)))
        return out

    def check(self, mass_ref=None, mass_tol=0.0):
        """lbm_check: the health of the stored state; raises LbmError with
        the library's reason when it fails, else returns the lbm_health
        fields as a dict"""
        H = Health()
        ref = None
        if mass_ref is not None:
            ref = np.ascontiguousarray(mass_ref, dtype=np.float64).reshape(2)
        _check(lib().lbm_check(self.h, _dp(ref) if ref is not None else None, float(mass_tol), ctypes.byref(H)))
        return dict(mass=tuple(H.mass), min_rho=H.min_rho, min_pop=tuple(H.min_pop),
                    nonfinite=tuple(H.nonfinite), negative=tuple(H.negative))

    def last_flux(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().lbm_last_flux(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def observations(self, node, step, target, weight=None):
        """lbm_objective_observations: misfit points of the next objective runs"""
        node = np.ascontiguousarray(node, dtype=np.int64).reshape(-1)
        step = np.ascontiguousarray(step, dtype=np.int32).reshape(-1)
        target = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64).reshape(-1)
        if not (node.size == step.size == target.size and (w is None or w.size == node.size)):
            raise ValueError("observation arrays differ in length")
        _check(lib().lbm_objective_observations(
            self.h, node.size, node.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
            step.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(target), _dp(w) if w is not None else None))

    def objective(self, nsteps, window, every=100, kind="flux"):
        """J over the last `window` of `nsteps` steps (kind "flux": the solute
        flux through flux_plane; "permeability": rho_A nu <u_z,phys> / G),
        plus the observation misfit when observations are set (see
        lbm_objective_run)"""
        _check(lib().lbm_objective_select(self.h, OBJECTIVES[kind]))
        J = ctypes.c_double()
        _check(lib().lbm_objective_run(self.h, int(nsteps), int(window), int(every), ctypes.byref(J)))
        return J.value

    def keep_states(self, max_bytes):
        """device memory the next objective runs may spend on states kept
        for the adjoint (lbm_objective_keep_states)"""
        _check(lib().lbm_objective_keep_states(self.h, int(max_bytes)))

    def objective_info(self):
        """{"kept_steps", "checkpoints"} of the last objective run"""
        kept, ck = ctypes.c_int(), ctypes.c_int()
        _check(lib().lbm_objective_info(self.h, ctypes.byref(kept), ctypes.byref(ck)))
        return {"kept_steps": kept.value, "checkpoints": ck.value}

    def adjoint(self, with_map=True):
        """(dJ/d(gx, gy, gz, tau, tau_s), dJ/dc_bot_map or None)"""
        g = np.empty(5)
        m = np.empty((self.p.nj, self.p.ni)) if with_map else None
        _check(lib().lbm_adjoint_run(self.h, _dp(g), _dp(m) if m is not None else None))
        return g, m

    def profile(self, on=True):
        _check(lib().lbm_profile_enable(self.h, int(bool(on))))

    def profile_read(self):
        ms = (ctypes.c_double * len(PROFILE_CLASSES))()
        cnt = (ctypes.c_int * len(PROFILE_CLASSES))()
        _check(lib().lbm_profile_read(self.h, ms, cnt))
        return {k: (ms[q], cnt[q]) for q, k in enumerate(PROFILE_CLASSES)}


def copy_bandwidth(nbytes=4 << 30, reps=20):
    """measured device-to-device copy bandwidth in GB/s"""
    g = ctypes.c_double()
    _check(lib().lbm_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g)))
    return g.value


def discs2d(ni, nj, seed, rmin, rmax, phi, walls=True):
    m = np.empty(ni * nj, np.uint8)
    ph = ctypes.c_double()
    _check(lib().lbm_geom_discs2d(ni, nj, seed, rmin, rmax, phi, int(walls),
                                  m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(ph)))
    return m.reshape(nj, ni), ph.value


def spheres3d(ni, nj, nk, seed, rmin, rmax, phi, zbuf=4):
    m = np.empty(ni * nj * nk, np.uint8)
    ph = ctypes.c_double()
    _check(lib().lbm_geom_spheres3d(ni, nj, nk, seed, rmin, rmax, phi, zbuf,
                                    m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(ph)))
    return m.reshape(nk, nj, ni), ph.value

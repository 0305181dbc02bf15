"""ctypes binding of the CPU oracle (oracle/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the reference the GPU result is compared with, and by bench.py's CPU baseline
timing.  The product library never
imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

VX = np.array([0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0])
VY = np.array([0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1])
VZ = np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1])


class OrcParams(ctypes.Structure):
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
        ("velocity_n_in", ctypes.c_int), ("velocity_n_out", ctypes.c_int),
        ("velocity_s_in", ctypes.c_int), ("velocity_s_out", ctypes.c_int),
        ("uy_in", ctypes.c_double), ("uy_out", ctypes.c_double),
        ("zface_flow", ctypes.c_int * 2), ("zface_uz", ctypes.c_double * 2),
        ("zface_conc", ctypes.c_int * 2), ("c_top", ctypes.c_double),
        ("halo_literal", ctypes.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblbm_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_solids.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8)]
        L.orc_set_f.argtypes = [vp, ctypes.c_int, dp]
        L.orc_get_f.argtypes = [vp, ctypes.c_int, dp]
        L.orc_raw_size.restype = ctypes.c_long
        L.orc_raw_size.argtypes = [vp]
        L.orc_get_raw.argtypes = [vp, ctypes.c_int, dp]
        L.orc_init_rest.argtypes = [vp, dp]
        L.orc_set_time.argtypes = [vp, ctypes.c_int]
        L.orc_get_time.restype = ctypes.c_int
        L.orc_get_time.argtypes = [vp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_run.argtypes = [vp, ctypes.c_int]
        L.orc_bcs.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_halo_exchange.argtypes = [vp, ctypes.c_int]
        L.orc_halo_pack.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_halo_unpack.argtypes = [vp, ctypes.c_int, dp, dp]
        L.orc_get_mv.argtypes = [vp, ctypes.c_int, dp, dp, dp, dp]
        L.orc_reduce_ref.argtypes = [vp, ctypes.c_int, dp]
        L.orc_reduce_sums.argtypes = [vp, dp]
        L.orc_last_flux.argtypes = [vp, dp, dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_tlm_objective.restype = ctypes.c_int
        L.orc_tlm_objective.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint8),
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        dp, dp, dp, dp, dp, dp]
        L.orc_tlm_objective_obs.restype = ctypes.c_int
        L.orc_tlm_objective_obs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_uint8), ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int), dp, dp,
                                            dp, dp, dp, dp]
        L.orc_tlm_objective_kind.restype = ctypes.c_int
        L.orc_tlm_objective_kind.argtypes = [ctypes.c_int] + L.orc_tlm_objective_obs.argtypes
        L.orc_set_ns.restype = ctypes.c_int
        L.orc_set_ns.argtypes = [vp, dp]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(**kw):
    """OrcParams from keyword arguments (same names and meaning as the
    product's lbm_params)."""
    p = OrcParams()
    p.dims = kw.get("dims", 3)
    p.ni, p.nj = kw["ni"], kw["nj"]
    p.nk = kw.get("nk", 1)
    p.nsubs = kw.get("nsubs", 1)
    p.halo_mode = kw.get("halo_mode", 0)
    tau = kw.get("tau", (1.0, 1.0))
    rho_A = kw.get("rho_A", (1.0, 1.0))
    g = kw.get("gaccel", (0.0,) * 6)
    for q in range(2):
        p.tau[q] = tau[q]
        p.rho_A[q] = rho_A[q]
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
    p.rho_in = kw.get("rho_in", 1.0)
    p.rho_out = kw.get("rho_out", 1.0)
    p.pfadd = kw.get("pfadd", 0.0)
    p.pfmul = kw.get("pfmul", 1.0)
    p.bc_ramp_start = kw.get("bc_ramp_start", -1)
    p.bc_ramp_stop = kw.get("bc_ramp_stop", -1)
    p.zface_pressure = kw.get("zface_pressure", 0)
    p.rho_bot = kw.get("rho_bot", 1.0)
    p.rho_top = kw.get("rho_top", 1.0)
    p.zface_solute = kw.get("zface_solute", 0)
    p.c_bot = kw.get("c_bot", 1.0)
    p.flux_plane = kw.get("flux_plane", -1)
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


class Oracle:
    """Owns one oracle lattice.  Keep references to series/maps alive."""

    def __init__(self, **kw):
        self.kw = dict(kw)
        self.p = make_params(**kw)
        self._keep = []
        for name in ("pressure_n_series", "pressure_s_series"):
            ser = kw.get(name)
            if ser is not None:
                for q, s in enumerate(ser):
                    if s is None:
                        continue
                    arr = np.ascontiguousarray(s, dtype=np.float64)
                    self._keep.append(arr)
                    getattr(self.p, name)[q] = _dp(arr)
                    getattr(self.p, name + "_len")[q] = arr.size
        cmap = kw.get("c_bot_map")
        if cmap is not None:
            arr = np.ascontiguousarray(cmap, dtype=np.float64).ravel()
            self._keep.append(arr)
            self.p.c_bot_map = _dp(arr)
        self.dims = self.p.dims
        self.Q = 9 if self.dims == 2 else 19
        self.N = self.p.ni * self.p.nj * (self.p.nk if self.dims == 3 else 1)
        self.nsubs = self.p.nsubs
        self.h = lib().orc_create(ctypes.byref(self.p))
        if not self.h:
            raise MemoryError("orc_create failed")

    def __del__(self):
        try:
            if self.h:
                lib().orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def shape(self):
        if self.dims == 2:
            return (self.p.nj, self.p.ni)
        return (self.p.nk, self.p.nj, self.p.ni)

    def set_solids(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        assert m.size == self.N
        self._mask = m
        lib().orc_set_solids(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))

    def set_f(self, subs, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.Q, self.N)
        lib().orc_set_f(self.h, subs, _dp(a))

    def get_f(self, subs):
        a = np.empty((self.Q, self.N))
        lib().orc_get_f(self.h, subs, _dp(a))
        return a

    def get_raw(self, subs):
        a = np.empty(lib().orc_raw_size(self.h))
        lib().orc_get_raw(self.h, subs, _dp(a))
        return a

    def init_rest(self, c0=None):
        if c0 is None:
            lib().orc_init_rest(self.h, None)
        else:
            c = np.ascontiguousarray(c0, dtype=np.float64).ravel()
            lib().orc_init_rest(self.h, _dp(c))

    @property
    def time(self):
        return lib().orc_get_time(self.h)

    @time.setter
    def time(self, t):
        lib().orc_set_time(self.h, t)

    def set_ns(self, ns):
        """Walsh partial bounce-back fraction per node (None: off)"""
        if ns is None:
            lib().orc_set_ns(self.h, None)
            return
        self._ns = np.ascontiguousarray(ns, dtype=np.float64).reshape(-1)
        assert self._ns.size == self.N
        if lib().orc_set_ns(self.h, _dp(self._ns)):
            raise MemoryError("orc_set_ns")

    def step(self, write_mv=False):
        lib().orc_step(self.h, int(write_mv))

    def run(self, n, write_last=True):
        for q in range(n):
            self.step(write_last if q == n - 1 else 0)

    def mv(self, subs=0):
        rho, ux, uy, uz = (np.empty(self.N) for _ in range(4))
        lib().orc_get_mv(self.h, subs, _dp(rho), _dp(ux), _dp(uy), _dp(uz))
        return rho, ux, uy, uz

    def reduce_ref(self, subs=0):
        out = np.empty(12)
        lib().orc_reduce_ref(self.h, subs, _dp(out))
        return out

    def reduce_sums(self):
        out = np.empty(6)
        lib().orc_reduce_sums(self.h, _dp(out))
        return out

    def last_flux(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().orc_last_flux(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def bcs(self, phase, t):
        lib().orc_bcs(self.h, phase, t)

    def halo_pack(self, phase):
        E = self.p.ni if self.dims == 2 else self.p.ni * self.p.nj
        nb = 3 if self.dims == 2 else 5
        pos = np.zeros(self.nsubs * nb * E)
        neg = np.zeros(self.nsubs * nb * E)
        lib().orc_halo_pack(self.h, phase, _dp(pos), _dp(neg))
        return pos, neg

    def halo_unpack(self, phase, pos, neg):
        lib().orc_halo_unpack(self.h, phase, _dp(np.ascontiguousarray(pos)), _dp(np.ascontiguousarray(neg)))


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def tlm_objective(mask, rho_bot, rho_top, zf, nsteps, window, params, cmap, dparams=None, dcmap=None,
                  observations=None, objective="flux"):
    """(J, dJ) of the window objective (window > 0: the solute flux through
    plane zf, or with objective="permeability" the permeability
    rho_A nu <u_z,phys> / G) plus the misfit of observations = (dense node,
    step, target, weight or None) from the FP64 tangent-linear model
    (oracle/lbm_tlm.cpp); params = (gx, gy, gz, tau, tau_s)."""
    kind = {"flux": 0, "permeability": 1}[objective]
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    nk, nj, ni = m.shape
    p = np.ascontiguousarray(list(params) + [0.0], dtype=np.float64)
    c = np.ascontiguousarray(cmap, dtype=np.float64).ravel()
    J, dJ = ctypes.c_double(), ctypes.c_double()
    dpp = np.ascontiguousarray(dparams, dtype=np.float64) if dparams is not None else None
    dcc = np.ascontiguousarray(dcmap, dtype=np.float64).ravel() if dcmap is not None else None
    if observations is None:
        on, os_, ot, ow = (np.zeros(0, np.int64), np.zeros(0, np.int32), np.zeros(0), None)
    else:
        on = np.ascontiguousarray(observations[0], dtype=np.int64)
        os_ = np.ascontiguousarray(observations[1], dtype=np.int32)
        ot = np.ascontiguousarray(observations[2], dtype=np.float64)
        ow = None if observations[3] is None else np.ascontiguousarray(observations[3], dtype=np.float64)
    rc = lib().orc_tlm_objective_kind(
        kind, ni, nj, nk, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), rho_bot, rho_top, zf, nsteps, window,
        _dp(p), _dp(c), on.size, on.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
        os_.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(ot), _dp(ow) if ow is not None else None,
        _dp(dpp) if dpp is not None else None, _dp(dcc) if dcc is not None else None, ctypes.byref(J),
        ctypes.byref(dJ))
    if rc:
        raise ValueError("orc_tlm_objective_kind: bad arguments")
    return J.value, dJ.value

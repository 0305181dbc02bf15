"""Reference input decks and output files from Python (include/lbm_io.h).

``read_deck`` is read_params (src/params.h:40-661), ``read_solids_bmp`` /
``read_solids_raw`` are spy_bmp and read_solids (src/lbio.c:3073-3363,
507-565), ``read_series`` reads the pressure series (src/latman.c:536-631),
and the writers produce frames.dat, rho/u .txt, .dat and 8-bit .raw files
(src/lbio.c:37-505, 934-1060, 2507-2561, 2864-2973) under the names of
gen_filename (src/lattice.c:371-452).  ``deck_main`` runs a whole deck as the
reference's main() does.  All work is done by the C library.
"""
import ctypes

import numpy as np

from . import LbmError, Params, lib as _lib

_i, _d = ctypes.c_int, ctypes.c_double
_I2, _D2 = ctypes.c_int * 2, ctypes.c_double * 2

_FACES = ["pressure_t_in", "pressure_b_in", "pressure_t_out", "pressure_b_out",
          "velocity_t_in", "velocity_b_in", "velocity_t_out", "velocity_b_out",
          "pressure_n_in", "pressure_s_in", "pressure_n_out", "pressure_s_out",
          "velocity_n_in", "velocity_s_in", "velocity_n_out", "velocity_s_out",
          "pressure_e_in", "pressure_w_in", "pressure_e_out", "pressure_w_out",
          "velocity_e_in", "velocity_w_in", "velocity_e_out", "velocity_w_out"]
_SOLUTE = [f"{k}_{f}_{io}" for f2 in ("tb", "ns", "ew") for k in ("constcon", "constflx")
           for f, io in ((f2[0], "in"), (f2[1], "in"), (f2[0], "out"), (f2[1], "out"))]


class Deck(ctypes.Structure):
    """lbm_deck: the fields of the reference's param struct that
    params.in sets, named as there"""
    _fields_ = (
        [(n, _i) for n in ("LX", "LY", "LZ", "BX", "BY", "BZ", "length_scale", "NumFrames", "FrameRate")]
        + [("tau", _D2), ("gforce", (ctypes.c_double * 3) * 2), ("end_grav", _I2)]
        + [(n, _i) for n in ("buoyancy", "incompressible", "simple_diffusion")]
        + [("rho_A", _D2), ("rho_B", _D2)]
        + [(n, _d) for n in ("rho_sigma", "rho_sigma_in", "rho_sigma_out", "u_sigma", "u_sigma_in", "u_sigma_out")]
        + [(n, _i) for n in ("sigma_t_on", "sigma_t_off", "sigma_start", "sigma_stop", "sigma_btc_rate",
                             "sigma_btc_spot", "GZL", "PressureBC", "AllBoundaryPeriodic")]
        + [(n, _d) for n in ("sink", "pfmul", "pfadd", "rho_in", "rho_out", "ux_in", "ux_out", "uy_in", "uy_out",
                             "uz_in", "uz_out", "big_V0")]
        + [("big_V0_solid", _D2), ("ns_flag", _i), ("ns", _d)]
        + [(n, _i) for n in ("ic_poisseuille", "bc_poisseuille", "bc_slip_north", "bc_sigma_slip",
                             "bc_ramp_start", "bc_ramp_stop")]
        + [(n, _I2) for n in _FACES]
        + [(n, _i) for n in _SOLUTE]
        + [(f"zeroconcgrad_{s}", _i) for s in ("t", "b", "n", "s", "e", "w", "full")]
        + [(n, _i) for n in ("plot_scale_dynamic", "use_colormap", "initial_condition", "x0", "y0", "z0", "r0")]
        + [("cut", _d)]
        + [(n, _i) for n in ("x1", "x2", "y1", "y2", "z1", "z2")]
        + [(n, _d) for n in ("rel_x1", "rel_x2", "rel_y1", "rel_y2", "rel_z1", "rel_z2")]
        + [(n, _i) for n in ("dump_rho", "dump_u", "dump_force", "dump_vor", "NumTimeSteps")]
        + [("NumNodes", ctypes.c_long)]
    )


_BOUND = False


def lib():
    global _BOUND
    L = _lib()
    if not _BOUND:
        cp, ip, dp = ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        sig = {
            "lbm_io_read_deck": [cp, ip, ip, ctypes.POINTER(Deck)],
            "lbm_io_deck_params": [ctypes.POINTER(Deck), ip, ip, ctypes.POINTER(Params)],
            "lbm_io_read_series": [cp, dp, ip, ctypes.POINTER(ctypes.c_int)],
            "lbm_io_read_solids_raw": [cp, ip, ip, ip, u8p],
            "lbm_io_read_solids_bmp": [cp, ip, ip, u8p],
            "lbm_io_filename": [ctypes.c_char_p, ctypes.c_size_t, cp, ip, ip, ip, ip, ip, ip, cp, ip, ip, cp],
            "lbm_io_frame_stats": [dp, dp, dp, dp, u8p, ctypes.c_long, dp],
            "lbm_io_frames_dat": [cp, ip, dp, dp],
            "lbm_io_write_txt": [cp, dp, ip, ip, ip],
            "lbm_io_write_dat": [cp, dp, u8p, ctypes.c_long],
            "lbm_io_write_raw8": [cp, dp, dp, dp, u8p, ctypes.c_long, dp, dp],
            "lbm_deck_main": [ip, ctypes.POINTER(ctypes.c_char_p)],
            "lbm_io_read_observations": [cp, cp, cp, ip, ip, ip, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
            "lbm_io_write_observations": [cp, ip, dp, dp, ctypes.POINTER(ctypes.c_int)],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.lbm_io_deck_bytes.restype = ctypes.c_size_t
        L.lbm_io_deck_bytes.argtypes = []
        if L.lbm_io_deck_bytes() != ctypes.sizeof(Deck):
            raise LbmError("Deck does not mirror lbm_deck")
        _BOUND = True
    return L


def _check(rc):
    if rc != 0:
        raise LbmError(_lib().lbm_last_error().decode())


def _b(s):
    return str(s).encode()


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def read_deck(path, dims=2, nsubs=1):
    d = Deck()
    _check(lib().lbm_io_read_deck(_b(path), dims, nsubs, ctypes.byref(d)))
    return d


def deck_params(deck, dims=2, nsubs=1):
    p = Params()
    _check(lib().lbm_io_deck_params(ctypes.byref(deck), dims, nsubs, ctypes.byref(p)))
    return p


def read_series(path):
    n = ctypes.c_int()
    _check(lib().lbm_io_read_series(_b(path), None, 0, ctypes.byref(n)))
    out = np.empty(max(n.value, 1))
    _check(lib().lbm_io_read_series(_b(path), _dp(out), n.value, ctypes.byref(n)))
    return out[:n.value]


def read_solids_raw(path, LX, LY, LZ):
    m = np.empty(LX * LY * LZ, np.uint8)
    _check(lib().lbm_io_read_solids_raw(_b(path), LX, LY, LZ, _u8(m)))
    return m.reshape(LZ, LY, LX)


def read_solids_bmp(path, LX, LY):
    m = np.empty(LX * LY, np.uint8)
    _check(lib().lbm_io_read_solids_bmp(_b(path), LX, LY, _u8(m)))
    return m.reshape(LY, LX)


def filename(out_path, dims, LX, LY, nsubs, prefix, frame, subs, suffix, num_procs=1, proc_id=0):
    buf = ctypes.create_string_buffer(1024)
    _check(lib().lbm_io_filename(buf, 1024, _b(out_path), dims, LX, LY, nsubs, num_procs, proc_id, _b(prefix),
                                 frame, subs, _b(suffix)))
    return buf.value.decode()


def frame_stats(rho, ux, uy, uz, mask):
    rho, ux, uy, uz = (_f64(a) for a in (rho, ux, uy, uz))
    m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
    out = np.empty(12)
    _check(lib().lbm_io_frame_stats(_dp(rho), _dp(ux), _dp(uy), _dp(uz), _u8(m), m.size, _dp(out)))
    return out


def frames_dat(path, time, stats, flux=(0.0, 0.0, 0.0, 0.0)):
    st, fx = _f64(stats), _f64(flux)
    _check(lib().lbm_io_frames_dat(_b(path), time, _dp(st), _dp(fx)))


def write_txt(path, a, LX, LY, LZ=1):
    a = _f64(a)
    _check(lib().lbm_io_write_txt(_b(path), _dp(a), LX, LY, LZ))


def write_dat(path, a, mask):
    a = _f64(a)
    m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
    _check(lib().lbm_io_write_dat(_b(path), _dp(a), _u8(m), m.size))


def write_raw8(path, ux, mask, a_max, a_min, uy=None, uz=None):
    ux, uy, uz = _f64(ux), _f64(uy), _f64(uz)
    m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
    hi, lo = _f64(np.resize(np.asarray(a_max, float), 3)), _f64(np.resize(np.asarray(a_min, float), 3))
    _check(lib().lbm_io_write_raw8(_b(path), _dp(ux), _dp(uy), _dp(uz), _u8(m), m.size, _dp(hi), _dp(lo)))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def read_observations(t_path, x_path, y_path, LX, LY):
    """(times, x, y, order) of the kept observation points"""
    n = ctypes.c_int()
    _check(lib().lbm_io_read_observations(_b(t_path), _b(x_path), _b(y_path), LX, LY, 0, None, None, None,
                                          None, ctypes.byref(n)))
    arrs = [np.zeros(max(n.value, 1), np.int32) for _ in range(4)]
    _check(lib().lbm_io_read_observations(_b(t_path), _b(x_path), _b(y_path), LX, LY, n.value,
                                          *(_ip(a) for a in arrs), ctypes.byref(n)))
    return tuple(a[:n.value] for a in arrs)


def write_observations(out_dir, conc, vel, order):
    conc, vel = _f64(conc), _f64(vel)
    order = np.ascontiguousarray(order, dtype=np.int32)
    _check(lib().lbm_io_write_observations(_b(out_dir), order.size, _dp(conc), _dp(vel), _ip(order)))


def deck_main(*args):
    """lbm_deck_main with the given arguments; returns its status"""
    argv = [b"lbgpu_prime"] + [_b(a) for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return lib().lbm_deck_main(len(argv), arr)

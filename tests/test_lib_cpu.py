"""CPU-side checks of the product library: it loads without a GPU, exports
every entry point the C header declares, and its seeded geometry generator
is reproducible and meets its stated properties."""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

import _pkg

ROOT = _pkg.ROOT


def header_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for name in sorted(os.listdir(inc)):
        if name.endswith(".h"):
            src = open(os.path.join(inc, name)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            syms |= set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", src))
    return sorted(syms)


def test_library_exports_header_symbols():
    L = _pkg.load()
    lib = L.lib()
    syms = header_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(lib, s), s
    assert "lbm_io_read_deck" in syms and "lbm_deck_main" in syms


def test_headers_compile_as_c(tmp_path):
    """the C ABI headers are plain C (the reference's host code is C): a C99
    translation unit that includes both and links against nothing compiles
    with warnings as errors"""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "use_headers.c"
    src.write_text('#include "lbm_mi355x.h"\n#include "lbm_io.h"\n'
                   "int use(lbm_handle *h) { lbm_params p = {0}; (void)p; return lbm_destroy(h); }\n")
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                        os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "use_headers.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_discs2d_properties_and_hash():
    L = _pkg.load()
    m, phi = L.discs2d(128, 128, 101, 4.0, 10.0, 0.70, walls=True)
    assert phi <= 0.70 and phi > 0.66
    assert m[0].all() and m[-1].all()
    assert abs(1 - m.mean() - phi) < 1e-12
    # discs never reach the wall rows: the rows next to the walls are not all solid
    assert not m[1].all() and not m[-2].all()
    h = hashlib.sha256(m.tobytes()).hexdigest()
    assert h == "7d02560889c5969c741a1501ce1b9ef35b44f6d163b68275fa249b467d4c87a6"
    m2, _ = L.discs2d(128, 128, 101, 4.0, 10.0, 0.70, walls=True)
    assert np.array_equal(m, m2)


def test_spheres3d_properties():
    L = _pkg.load()
    m, phi = L.spheres3d(48, 40, 64, 202, 6.0, 14.0, 0.40, zbuf=4)
    assert phi <= 0.40 and phi > 0.35
    assert m[:4].sum() == 0 and m[-4:].sum() == 0
    sample = m[4:-4]
    assert abs((1 - sample.mean()) - phi) < 1e-12
    m2, _ = L.spheres3d(48, 40, 64, 202, 6.0, 14.0, 0.40, zbuf=4)
    assert np.array_equal(m, m2)
    m3, _ = L.spheres3d(48, 40, 64, 203, 6.0, 14.0, 0.40, zbuf=4)
    assert not np.array_equal(m, m3)


@pytest.mark.parametrize("name,args,sha", [
    ("256^3 flow, porosity 0.40", (256, 256, 256, 202, 6.0, 14.0, 0.40),
     "58d3c9c5de2e132090d9702f02a17a6e2bba6167678a723a13e2f3f19a259d0e"),
    ("192^3 flow + solute, porosity 0.45", (192, 192, 192, 303, 6.0, 14.0, 0.45),
     "a3f40ce3a7fcd4651967e742b88ae9fccfdd7d4d57f027e4f2731882d10a4e84"),
    ("128^3 adjoint, porosity 0.40", (128, 128, 128, 404, 6.0, 14.0, 0.40),
     "3c1fc171307dd4657989c624b4d869f5dac9a6e1640a4a983cee306d428ce890"),
    ("448^3 slab, porosity 0.35", (448, 448, 448, 505, 8.0, 20.0, 0.35),
     "a1c6f0410c42e6daad980d701bbb5adf040cbcc44965e8422acd9049b093b166"),
])
def test_benchmark_geometries_hash(name, args, sha):
    """the seeded sphere packs of the benchmark configurations (4-plane fluid
    buffers at each z face) are reproducible byte for byte"""
    L = _pkg.load()
    m, phi = L.spheres3d(*args, zbuf=4)
    assert phi <= args[-1]
    assert hashlib.sha256(m.tobytes()).hexdigest() == sha, name


def test_geometry_rejects_bad_arguments():
    L = _pkg.load()
    with pytest.raises(L.LbmError):
        L.discs2d(16, 16, 1, 9.0, 10.0, 0.7, walls=True)


def test_profile_class_count_matches_header():
    import re as _re
    L = _pkg.load()
    src = open(os.path.join(ROOT, "include", "lbm_mi355x.h")).read()
    n = int(_re.search(r"#define LBM_PROFILE_CLASSES (\d+)", src).group(1))
    assert n == len(L.PROFILE_CLASSES)


@pytest.mark.parametrize("name,args,sha,phi", [
    ("256^3 flow, seed 202", (256, 256, 256, 202, 6.0, 14.0, 0.40),
     "58d3c9c5de2e132090d9702f02a17a6e2bba6167678a723a13e2f3f19a259d0e", 0.3999288620487336),
    ("192^3 flow + solute, seed 303", (192, 192, 192, 303, 6.0, 14.0, 0.45),
     "a3f40ce3a7fcd4651967e742b88ae9fccfdd7d4d57f027e4f2731882d10a4e84", 0.44973209399532005),
    ("128^3 adjoint, seed 404", (128, 128, 128, 404, 6.0, 14.0, 0.40),
     "3c1fc171307dd4657989c624b4d869f5dac9a6e1640a4a983cee306d428ce890", 0.3986236572265625),
    ("448^3 slab, seed 505", (448, 448, 448, 505, 8.0, 20.0, 0.35),
     "a1c6f0410c42e6daad980d701bbb5adf040cbcc44965e8422acd9049b093b166", 0.34993869307310993),
])
def test_benchmark_geometries_recorded(name, args, sha, phi):
    """the seeded sphere packs of the measured workloads, as .raw bytes
    (i fastest), are the recorded ones"""
    L = _pkg.load()
    m, p = L.spheres3d(*args, zbuf=4)
    assert hashlib.sha256(m.tobytes()).hexdigest() == sha, name
    assert p == phi


@pytest.mark.parametrize("bad,match", [
    (dict(dims=4), "dims must be 2 or 3"),
    (dict(ni=0), "bad lattice size"),
    (dict(dims=3, nk=0), "bad lattice size"),
    (dict(nsubs=3), "nsubs must be 1 or 2"),
    (dict(tau=(0.5, 1.0)), "must exceed 1/2"),
    (dict(nsubs=2, tau=(0.8, 0.4)), "must exceed 1/2"),
    (dict(storage=3), "bad storage/compute"),
    (dict(storage=2, compute=0), "FP16 storage runs the FP32 collision"),
    (dict(dims=3, ni=2048, nj=1024, nk=1024), "too large for 32-bit node indices"),
    (dict(ni=40000, nj=2), "LX = 40000 above 32767 is not supported"),
    (dict(dims=3, nk=5, flux_plane=5), "flux_plane 5 lies outside"),
    (dict(flux_plane=6), "flux_plane 6 lies outside"),
    (dict(halo_literal=2), "halo_literal must be 0 or 1"),
])
def test_create_rejects_bad_parameters(bad, match):
    """lbm_create checks its parameters before it touches the GPU, and
    reports the reason (the reference's process_exit(1) on bad input)"""
    L = _pkg.load()
    kw = dict(dims=2, ni=8, nj=6, nk=1, tau=(0.8, 0.8))
    kw.update(bad)
    with pytest.raises(L.LbmError, match=match):
        L.Lattice(**kw)


def test_handle_calls_reject_null_arguments():
    """calls on a handle check their pointers before anything else (no GPU
    needed): a null handle or output is an error with its reason"""
    L = _pkg.load()
    lib = L.lib()
    kept, ck = ctypes.c_int(-1), ctypes.c_int(-1)
    assert lib.lbm_objective_info(None, ctypes.byref(kept), ctypes.byref(ck)) != 0
    assert b"null argument" in lib.lbm_last_error()
    assert lib.lbm_objective_keep_states(None, 1) != 0
    assert b"null handle" in lib.lbm_last_error()

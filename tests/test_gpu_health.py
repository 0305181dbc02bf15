"""lbm_check: the library's failure detection (the reference stops the run
on any CUDA error, checkCUDAError -> process_exit(1), src/lattice.c:3123-3142,
and can stop on negative populations, PUKE_NEGATIVE_DENSITIES,
src/flags.h:264-270).  A healthy state passes and reports the same masses as
the FP64 reductions; NaN, infinite and negative flow populations and a mass
drift beyond the tolerance make it return nonzero with the reason."""
import numpy as np
import pytest

import _pkg
from test_gpu_parity import random_state, say

pytestmark = pytest.mark.gpu


def lattice(L, storage=0, compute=0):
    ni, nj, nk = 12, 10, 14
    mask, _ = L.spheres3d(ni, nj, nk, 77, 1.5, 3.0, 0.7, zbuf=2)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), gaccel=(1e-5, 0, 2e-5, 0, 0, 0),
              storage=storage, compute=compute)
    if storage == 2:
        kw.update(store_shift_c=0.0)
    return L.Lattice(mask=mask, **kw), mask


@pytest.mark.parametrize("storage,compute", [(0, 0), (1, 1), (2, 1)])
def test_healthy_state_passes_and_reports_masses(storage, compute):
    say(f"case health healthy storage={storage}")
    L = _pkg.load()
    g, mask = lattice(L, storage, compute)
    rng = np.random.default_rng(5)
    N = mask.size
    f, c = random_state(rng, 19, N, 2)
    c = np.abs(c) + 0.01 * np.array([1 / 3] + [1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] * 2 + [1 / 36] * 8)[:, None]
    g.upload_f(0, f)
    g.upload_f(1, c)
    h0 = g.check()
    assert h0["nonfinite"] == (0, 0) and h0["negative"][0] == 0 and h0["min_rho"] > 0.9
    # the masses are the stored populations' sums; the macroscopic write of
    # the next step (before its collision) sums the same populations
    g.run(40, write_last=True)
    # relative drift bounds: FP64 round-off, FP32 / FP16 storage rounding
    h = g.check(mass_ref=h0["mass"], mass_tol={0: 1e-12, 1: 1e-6, 2: 1e-3}[storage])
    fl = mask.reshape(-1) != 1
    rho = g.mv(0)[0][fl]
    say(f"storage={storage}: mass {h['mass']}, from mv {np.sum(rho - 1.0):.6e}, min rho {h['min_rho']:.6f}")
    # periodic, bounce-back: both masses stay (round-off of the storage)
    assert abs(h["mass"][0] - h0["mass"][0]) <= (1e-10 if storage == 0 else 1e-2) * fl.sum()
    assert abs(h["mass"][1] - h0["mass"][1]) <= (1e-12 if storage == 0 else 1e-3) * abs(h0["mass"][1])


def test_nan_and_infinite_populations_fail():
    say("case health nan")
    L = _pkg.load()
    g, mask = lattice(L)
    N = mask.size
    fl = np.flatnonzero(mask.reshape(-1) != 1)
    g.init_rest()
    g.check()
    f = np.zeros((19, N))
    f[7, fl[5]] = np.nan
    g.upload_f(0, f)
    with pytest.raises(L.LbmError, match="1 flow populations are NaN or infinite"):
        g.check()
    g.init_rest()
    c = np.zeros((19, N))
    c[3, fl[9]] = np.inf
    c[4, fl[11]] = -np.inf
    g.upload_f(1, c)
    with pytest.raises(L.LbmError, match="2 solute populations are NaN or infinite"):
        g.check()
    # a NaN spreads by stepping; the check still names it
    g.init_rest()
    f[:] = 0.0
    f[0, fl[0]] = np.nan
    g.upload_f(0, f)
    g.run(5)
    with pytest.raises(L.LbmError, match="NaN or infinite"):
        g.check()


def test_negative_flow_populations_and_density_fail():
    say("case health negative")
    L = _pkg.load()
    g, mask = lattice(L)
    N = mask.size
    fl = np.flatnonzero(mask.reshape(-1) != 1)
    g.init_rest()
    f = np.zeros((19, N))
    f[2, fl[3]] = -0.06  # f~ = -0.06 against w rho_A = 1/18: f = -0.0044
    g.upload_f(0, f)
    with pytest.raises(L.LbmError, match="1 negative flow populations"):
        g.check()
    f[2, fl[3]] = 0.0
    f[0, fl[4]] = -1.2  # the density rho_A + delta-rho = -0.2
    g.upload_f(0, f)
    with pytest.raises(L.LbmError, match="negative flow populations .* least density -0.2"):
        g.check()
    # negative solute populations are reported, not an error
    g.init_rest()
    c = np.zeros((19, N))
    c[5, fl[6]] = -1e-4
    g.upload_f(1, c)
    h = g.check()
    assert h["negative"] == (0, 1) and h["min_pop"][1] == -1e-4


def test_mass_drift_fails_beyond_tolerance():
    say("case health mass drift")
    L = _pkg.load()
    g, mask = lattice(L)
    g.init_rest(c0=np.full(mask.size, 0.5))
    h = g.check()
    nf = int(np.sum(mask != 1))
    assert h["mass"][0] == 0.0 and abs(h["mass"][1] - 0.5 * nf) <= 1e-12 * nf
    g.check(mass_ref=h["mass"], mass_tol=0.0)
    ref = (h["mass"][0] + 1e-6 * nf, np.nan)  # flow: 1e-6 of the total mass rho_A N_f
    g.check(mass_ref=ref, mass_tol=1.01e-6)
    with pytest.raises(L.LbmError, match="flow mass .* drifted"):
        g.check(mass_ref=ref, mass_tol=0.99e-6)
    ref = (np.nan, h["mass"][1] * (1 + 1e-3))
    with pytest.raises(L.LbmError, match="solute mass .* drifted"):
        g.check(mass_ref=ref, mass_tol=1e-4)
    g.check(mass_ref=ref, mass_tol=2e-3)

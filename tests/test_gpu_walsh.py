"""Walsh partial bounce-back on the GPU against the FP64 CPU oracle.

The oracle's Walsh update is pinned on the CPU (ns = 0 is the plain update,
ns = 1 reverses every population, mass conservation and the closed-form
momentum recursion of a uniform ns).  Here the HIP half-step kernels with a
per-node ns field must match it to 1e-12 (FP64 storage and collision), on
random states, with the reference's boundary conditions, and after the field
is switched off again."""
import numpy as np
import pytest

import _pkg
from test_gpu_parity import build_pair, compare, load_both, random_state, relnorm, say

pytestmark = pytest.mark.gpu


def _kw(dims, shape, nsubs, **extra):
    return dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=nsubs,
                tau=(0.8, 0.7), gaccel=(1e-4, -2e-4, 3e-4, 0, 0, 0), **extra)


@pytest.mark.parametrize("dims,shape,nsubs", [(2, (12, 17), 1), (2, (9, 10), 2), (3, (7, 6, 9), 1),
                                               (3, (8, 5, 6), 2)])
def test_random_ns_field(dims, shape, nsubs):
    say(f"case Walsh dims={dims} shape={shape} nsubs={nsubs}")
    rng = np.random.default_rng(400 + dims * 10 + nsubs)
    mask = (rng.random(shape) < 0.2).astype(np.uint8)
    o, g = build_pair(mask, **_kw(dims, shape, nsubs))
    ns = rng.uniform(0.0, 1.0, shape)
    ns[rng.random(shape) < 0.2] = 0.0
    o.set_ns(ns)
    g.set_ns(ns)
    Q = 9 if dims == 2 else 19
    f, gg = random_state(rng, Q, mask.size, nsubs)
    load_both(o, g, f, gg)
    for t in range(7):
        o.step(True)
        g.step(True)
    compare(o, g, mask, nsubs, f"Walsh {dims}D nsubs={nsubs}")


def test_ns_one_is_period_two_on_gpu():
    """ns = 1 everywhere: each step reverses every population, so two steps
    (one odd and one even half-step) return the start state"""
    shape = (10, 13)
    mask = np.zeros(shape, np.uint8)
    o, g = build_pair(mask, **_kw(2, shape, 1))
    g.set_ns(np.ones(shape))
    rng = np.random.default_rng(9)
    f, _ = random_state(rng, 9, mask.size, 1)
    g.upload_f(0, f)
    g.run(2)
    assert np.array_equal(g.download_f(0), f)


def test_ns_with_reference_bcs_and_switch_off():
    """Zou-He inflow/outflow and zero-gradient solute faces with a partial
    bounce-back field, then lbm_set_ns(NULL) back to the plain update"""
    rng = np.random.default_rng(31)
    ni, nj = 22, 18
    mask = (rng.random((nj, ni)) < 0.1).astype(np.uint8)
    mask[:2] = 0
    mask[-2:] = 0
    kw = dict(dims=2, ni=ni, nj=nj, nsubs=2, tau=(0.9, 0.75), zeroconcgrad_n=1, zeroconcgrad_s=1,
              pressure_s_in=(1, 0), pressure_n_out=(1, 0), rho_in=1.002, rho_out=0.999)
    o, g = build_pair(mask, **kw)
    ns = np.where(rng.random((nj, ni)) < 0.5, rng.uniform(0.0, 0.6, (nj, ni)), 0.0)
    o.set_ns(ns)
    g.set_ns(ns)
    f, gg = random_state(rng, 9, mask.size, 2)
    load_both(o, g, f, gg)
    for t in range(20):
        o.step(True)
        g.step(True)
    compare(o, g, mask, 2, "Walsh + Zou-He")
    o.set_ns(None)
    g.set_ns(None)
    for t in range(6):
        o.step(True)
        g.step(True)
    compare(o, g, mask, 2, "Walsh switched off")


@pytest.mark.parametrize("halo", [0, 1])
def test_ns_with_3d_zface_conditions(halo):
    """the z-face pressure and solute conditions formed inside the half-step
    kernel, with partial bounce-back at the face nodes too"""
    say(f"case Walsh 3D z faces halo={halo}")
    L = _pkg.load()
    ni, nj, nk = 14, 12, 20
    mask, _ = L.spheres3d(ni, nj, nk, 77, 2.5, 4.0, 0.6, zbuf=3)
    rng = np.random.default_rng(8)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.85, 0.7), zface_pressure=1, rho_bot=1.003,
              rho_top=0.998, zface_solute=1, c_bot=1.0, halo_mode=halo)
    o, g = build_pair(mask, **kw)
    ns = rng.uniform(0.0, 0.5, mask.shape)
    o.set_ns(ns)
    g.set_ns(ns)
    f, gg = random_state(rng, 19, mask.size, 2)
    load_both(o, g, f, gg)
    for t in range(10):
        o.step(True)
        g.step(True)
    compare(o, g, mask, 2, f"Walsh 3D z faces halo={halo}")


def test_ns_rejected_values_and_adjoint():
    L = _pkg.load()
    shape = (6, 7)
    g = L.Lattice(mask=np.zeros(shape, np.uint8), **_kw(2, shape, 1))
    bad = np.zeros(shape)
    bad[2, 3] = 1.5
    with pytest.raises(L.LbmError):
        g.set_ns(bad)
    bad[2, 3] = np.nan
    with pytest.raises(L.LbmError):
        g.set_ns(bad)
    with pytest.raises(ValueError):
        g.set_ns(np.zeros(5))
    # the discrete adjoint does not cover the Walsh update
    ni, nj, nk = 8, 6, 12
    mask, _ = L.spheres3d(ni, nj, nk, 5, 1.5, 2.5, 0.7, zbuf=3)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, zface_solute=1, c_bot=1.0, flux_plane=nk - 2)
    a = L.Lattice(mask=mask, **kw)
    a.init_rest()
    a.set_ns(np.full(mask.shape, 0.1))
    with pytest.raises(L.LbmError):
        a.objective(4, 2)


def test_reduced_precision_with_ns():
    """FP32-shifted storage with FP64 collision under a partial bounce-back
    field stays within the storage format's rounding of the FP64 result"""
    L = _pkg.load()
    rng = np.random.default_rng(12)
    shape = (16, 20)
    mask = (rng.random(shape) < 0.15).astype(np.uint8)
    kw = _kw(2, shape, 1)
    a = L.Lattice(mask=mask, **kw)
    b = L.Lattice(mask=mask, storage=1, compute=0, **kw)
    ns = rng.uniform(0.0, 0.4, shape)
    a.set_ns(ns)
    b.set_ns(ns)
    a.init_rest()
    b.init_rest()
    a.run(50)
    b.run(50)
    fl = mask.reshape(-1) != 1
    r = relnorm(b.mv(0)[1][fl], a.mv(0)[1][fl])
    say(f"fp32s vs fp64 with ns: u_x {r:.3e}")
    assert r < 1e-4

"""The kernels' 64-bit offset variants (used above 4 GiB of populations)
on small lattices, selected by LBM_OFFSET64=1: the FP64 step, the z-face
conditions formed in the half-step, Walsh partial bounce-back and the
discrete adjoint must give what the oracle (or the 32-bit variants, for the
adjoint) gives, bit for bit."""
import numpy as np
import pytest

import _pkg
from test_gpu_adjoint import PARAMS, gpu_objective, setup
from test_gpu_parity import build_pair, compare, load_both, random_state, say

pytestmark = pytest.mark.gpu


@pytest.fixture
def off64(monkeypatch):
    monkeypatch.setenv("LBM_OFFSET64", "1")


@pytest.mark.parametrize("dims,shape,nsubs", [(2, (12, 17), 1), (2, (9, 10), 2), (3, (7, 6, 9), 1),
                                               (3, (8, 5, 6), 2)])
@pytest.mark.parametrize("walsh", [False, True])
def test_random_state_offset64(off64, dims, shape, nsubs, walsh):
    say(f"case offset64 dims={dims} shape={shape} nsubs={nsubs} walsh={walsh}")
    rng = np.random.default_rng(900 + dims * 10 + nsubs)
    mask = (rng.random(shape) < 0.25).astype(np.uint8)
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=nsubs,
              tau=(0.8, 0.7), gaccel=(1e-4, -2e-4, 3e-4, 0, 0, 0))
    o, g = build_pair(mask, **kw)
    if walsh:
        ns = rng.uniform(0.0, 1.0, shape)
        o.set_ns(ns)
        g.set_ns(ns)
    Q = 9 if dims == 2 else 19
    f, gg = random_state(rng, Q, mask.size, nsubs)
    load_both(o, g, f, gg)
    for t in range(5):
        o.step(True)
        g.step(True)
    _, bitwise = compare(o, g, mask, nsubs, f"offset64 {dims}D nsubs={nsubs} walsh={walsh}")
    assert bitwise


def test_zface_conditions_offset64(off64):
    L = _pkg.load()
    ni, nj, nk = 14, 12, 24
    mask, _ = L.spheres3d(ni, nj, nk, 404, 2.5, 4.5, 0.55, zbuf=3)
    cmap = np.random.default_rng(5).uniform(0.8, 1.2, (nj, ni))
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
    o, g = build_pair(mask, **kw)
    o.init_rest()
    g.init_rest()
    for t in range(30):
        o.step(True)
        g.step(True)
        fo, fg = o.last_flux(), g.last_flux()
        assert abs(fg[1] - fo[1]) <= 1e-12 * abs(fo[1]) + 1e-300
    _, bitwise = compare(o, g, mask, 2, "offset64 3D z faces")
    assert bitwise


def test_adjoint_offset64_equals_offset32(monkeypatch):
    L, mask, cmap = setup()
    lat, J = gpu_objective(L, mask, cmap, PARAMS, 40, 16, 15)
    g32, m32 = lat.adjoint()
    lat.close()
    monkeypatch.setenv("LBM_OFFSET64", "1")
    lat, J64 = gpu_objective(L, mask, cmap, PARAMS, 40, 16, 15)
    g64, m64 = lat.adjoint()
    lat.close()
    say(f"J {J:.15e} / {J64:.15e}; grad {g32} / {g64}")
    assert J64 == J
    assert np.array_equal(g64, g32) and np.array_equal(m64, m32)

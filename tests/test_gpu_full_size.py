"""Size-independent properties at the full sizes of the benchmark
configurations, where the FP64 oracle would take minutes:

- 256^3 flow (porosity 0.40): mass conserved to round-off in a periodic,
  body-force-driven run;
- 192^3 flow + solute (porosity 0.45): solute mass conserved to round-off;
- 448^3 flow + solute slab in FP32-shifted storage (porosity 0.35; its
  populations take 4.8 GB, so the 64-bit offset kernels run): the slab in
  halo mode with the in-device periodic exchange equals the periodic run
  bit for bit;
- 128^3 flow + solute (porosity 0.40), 2000 steps with checkpoints every
  100: the discrete-adjoint gradient equals central differences of the
  forward run to 1e-5;
and, at reduced step counts, parity with the FP64 oracle on the 256^3 flow
and 192^3 flow + solute configurations' own seeded geometry and boundary
conditions."""
import sys

import numpy as np
import pytest

import _pkg
from test_gpu_parity import build_pair, compare

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_c2_mass_conserved_to_roundoff():
    say("case 256^3 flow mass")
    L = _pkg.load()
    n = 256
    mask, _ = L.spheres3d(n, n, n, 202, 6.0, 14.0, 0.40, zbuf=4)
    lat = L.Lattice(mask=mask, dims=3, ni=n, nj=n, nk=n, tau=(0.8, 0.75), gaccel=(2e-6, -1e-6, 1e-5, 0, 0, 0))
    lat.init_rest()
    lat.run(400, write_last=True)
    s = lat.reduce_sums()
    rho = lat.mv(0)[0]
    fl = mask.reshape(-1) != 1
    scale = np.abs(rho[fl] - 1.0).sum()
    say(f"sum of drho {s[0]:.3e}, sum of |drho| {scale:.3e}, fluid nodes {s[4]:.0f}")
    assert s[4] == fl.sum()
    assert scale > 1e-6  # the force has moved the density
    assert abs(s[0]) <= 1e-11 * scale
    lat.close()


def test_c3_solute_mass_conserved_to_roundoff():
    say("case 192^3 flow + solute mass")
    L = _pkg.load()
    n = 192
    mask, _ = L.spheres3d(n, n, n, 303, 3.0, 6.0, 0.45, zbuf=4)
    lat = L.Lattice(mask=mask, dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75),
                    gaccel=(0, 0, 1e-5, 0, 0, 0))
    k = np.arange(n)[:, None, None]
    c0 = np.broadcast_to(np.where((k > 40) & (k < 80), 1.0, 0.0), (n, n, n)).reshape(-1)
    lat.init_rest(c0)
    lat.run(2, write_last=True)
    m0 = lat.reduce_sums()
    lat.run(400, write_last=True)
    m1 = lat.reduce_sums()
    say(f"solute mass {m0[5]:.15e} -> {m1[5]:.15e}; flow {m0[0]:.3e} -> {m1[0]:.3e}")
    assert m0[5] > 1e5
    assert abs(m1[5] - m0[5]) <= 1e-13 * m0[5]
    lat.close()


def test_c5_slab_halo_exchange_equals_periodic():
    say("case 448^3 slab halo vs periodic, FP32-shifted")
    L = _pkg.load()
    n = 448
    mask, _ = L.spheres3d(n, n, n, 505, 8.0, 20.0, 0.35, zbuf=4)
    base = dict(dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
                rho_top=1.0, zface_solute=1, c_bot=1.0, flux_plane=n - 2, storage=1, compute=0)
    fl = np.flatnonzero(mask.reshape(-1) != 1)
    nodes = np.random.default_rng(7).choice(fl, 200000, replace=False)
    # interior and the planes next to the slab faces, where the exchange acts
    k = fl // (n * n)
    nodes = np.concatenate([nodes, fl[k <= 1][:20000], fl[k >= n - 2][:20000]])
    out = []
    for halo in (0, 1):
        lat = L.Lattice(mask=mask, halo_mode=halo, **base)
        lat.init_rest()
        lat.run(21, write_last=True)
        out.append((lat.reduce_sums(), lat.last_flux(), lat.sample_mv(0, nodes), lat.sample_mv(1, nodes)))
        lat.close()
    (s0, f0, a0, b0), (s1, f1, a1, b1) = out
    say(f"sums {s0} / {s1}; outlet flux {f0} / {f1}")
    assert np.array_equal(s0, s1)
    assert f0 == f1
    assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
    assert np.abs(a0[:, 3]).max() > 0  # the flow has started


def test_c4_adjoint_matches_central_differences_full_size():
    say("case 128^3 adjoint vs central differences, 2000 steps")
    L = _pkg.load()
    n, steps, window, every = 128, 2000, 200, 100
    mask, _ = L.spheres3d(n, n, n, 404, 6.0, 14.0, 0.40, zbuf=4)
    cmap = np.random.default_rng(4040).uniform(0.8, 1.2, (n, n))
    # a flux plane the solute reaches within the run
    plane = 8

    def objective(params, cm):
        lat = L.Lattice(mask=mask, dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(params[3], params[4]),
                        gaccel=(params[0], params[1], params[2], 0, 0, 0), zface_pressure=1, rho_bot=1.001,
                        rho_top=1.0, zface_solute=1, c_bot_map=cm, flux_plane=plane)
        lat.init_rest()
        return lat, lat.objective(steps, window, every)

    P = np.array([0.0, 0.0, 1e-6, 0.8, 0.75])
    lat, J = objective(P, cmap)
    grad, gmap = lat.adjoint()
    lat.close()
    say(f"J {J:.12e}")
    assert J > 1e-9
    for q, h in ((2, 1e-8), (3, 1e-6), (4, 1e-6)):
        dp = np.zeros(5)
        dp[q] = h
        a, Jp = objective(P + dp, cmap)
        a.close()
        b, Jm = objective(P - dp, cmap)
        b.close()
        fd = (Jp - Jm) / (2 * h)
        say(f"param {q}: adjoint {grad[q]:.10e} fd {fd:.10e}")
        assert abs(grad[q] - fd) <= 1e-5 * abs(fd)
    d = np.random.default_rng(9).uniform(-1, 1, cmap.shape)
    h = 1e-6
    a, Jp = objective(P, cmap + h * d)
    a.close()
    b, Jm = objective(P, cmap - h * d)
    b.close()
    fd = (Jp - Jm) / (2 * h)
    adj = float(np.sum(gmap * d))
    say(f"inlet map direction: adjoint {adj:.10e} fd {fd:.10e}")
    assert abs(adj - fd) <= 1e-5 * abs(fd)


@pytest.mark.parametrize("case", ["256^3 flow", "192^3 flow + solute"])
def test_parity_on_benchmark_geometry(case):
    """the GPU path against the FP64 oracle on the benchmark configurations'
    seeded sphere packs, z-face conditions and sizes, 20 steps (a few
    seconds of oracle time); bit for bit"""
    say(f"case {case} geometry, 20 steps against the oracle")
    L = _pkg.load()
    if case.startswith("256"):
        n, nsubs = 256, 1
        mask, _ = L.spheres3d(n, n, n, 202, 6.0, 14.0, 0.40, zbuf=4)
        kw = dict(dims=3, ni=n, nj=n, nk=n, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001, rho_top=1.0)
    else:
        n, nsubs = 192, 2
        mask, _ = L.spheres3d(n, n, n, 303, 6.0, 14.0, 0.45, zbuf=4)
        kw = dict(dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
                  rho_top=1.0, zface_solute=1, c_bot=1.0, flux_plane=n - 2)
    o, g = build_pair(mask, **kw)
    o.init_rest()
    g.init_rest()
    o.run(20)
    g.run(20)
    _, bitwise = compare(o, g, mask, nsubs, f"{case}, 20 steps")
    assert bitwise
    g.close()

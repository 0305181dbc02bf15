"""Pins of the oracle's Zou-He velocity conditions and per-face z conditions
(the reference has them only on its CPU path, src/bcs.c:766-862, in the
compressible form; the GPU path it drives has no velocity condition).

In the incompressible shifted model the velocity is j / rho_A, so a velocity
face prescribes the node's momentum rho_A u; the unknowns then follow from
the same Zou-He lines as the pressure face with that momentum.  The checks:
  - identities: after the condition the face nodes collide with exactly the
    prescribed velocity and no tangential velocity (3-D faces, 2-D rows,
    both phases, with the ramp of the 2-D rows);
  - plug flow: an empty column with a velocity inlet and a pressure outlet
    reaches the uniform velocity everywhere;
  - a square duct fed by a uniform velocity inlet carries the prescribed
    flux through every plane and, away from the inlet, the same profile as
    the pressure-driven duct with that flux;
  - 2-D: a velocity row drives plug flow in an empty channel, and in a
    walled channel the profile of the pressure-driven one at the same flux;
  - per-face solute kinds: a Dirichlet top face holds c_top, a zero-gradient
    bottom face has no net first moment."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle

VEL, PRES, NONE = 3, 2, 1
DIRI, ZGRAD = 2, 3


@pytest.mark.parametrize("phase_even", [0, 1])
def test_zface_velocity_identities(phase_even):
    """both z faces velocity: u_z = zface_uz exactly on each face, u_x = u_y
    = 0; a sign or factor slip (u_z = -u, u_z = u / rho) fails"""
    rng = np.random.default_rng(23)
    ni, nj, nk = 5, 4, 8
    uz = (2.5e-3, -1.5e-3)
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=1, tau=(0.8, 1.0), rho_A=(1.3, 1.0),
               zface_flow=(VEL, VEL), zface_uz=uz)
    o.set_solids(np.zeros((nk, nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (19, ni * nj * nk)))
    if phase_even:
        o.step(False)
    o.step(True)
    rho, ux, uy, u = (x.reshape(nk, nj, ni) for x in o.mv(0))
    np.testing.assert_allclose(u[0], uz[0], rtol=0, atol=1e-17)
    np.testing.assert_allclose(u[-1], uz[1], rtol=0, atol=1e-17)
    for v in (ux, uy):
        np.testing.assert_allclose(v[0], 0, atol=1e-17)
        np.testing.assert_allclose(v[-1], 0, atol=1e-17)
    # the density is free on a velocity face: it is not pinned to rho_A
    assert np.ptp(rho[0]) > 1e-5


def test_zface_mixed_kinds_identities():
    """velocity at z = 0 with pressure at z = nk-1: each face obeys its own
    condition"""
    rng = np.random.default_rng(29)
    ni, nj, nk = 4, 5, 7
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, tau=(0.9, 1.0), zface_flow=(VEL, PRES), zface_uz=(1e-3, 0.0),
               rho_top=0.999)
    o.set_solids(np.zeros((nk, nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (19, ni * nj * nk)))
    o.step(True)
    rho, ux, uy, u = (x.reshape(nk, nj, ni) for x in o.mv(0))
    np.testing.assert_allclose(u[0], 1e-3, rtol=0, atol=1e-17)
    np.testing.assert_allclose(rho[-1], 0.999, rtol=0, atol=2e-16)
    np.testing.assert_allclose(ux[[0, -1]], 0, atol=1e-17)


def test_velocity_inlet_plug_flow():
    """empty column, periodic in x and y: a velocity inlet at z = 0 and a
    pressure outlet at z = nk-1 drive the uniform velocity u0 through the
    whole column at steady state, with the outlet density throughout.  Both
    faces reflect sound, so the start-up wave rings down slowly (about a
    factor 1.6 per 1000 steps here: 3.5e-2 after 6000 steps, 5e-5 after
    20000); 40000 steps leave it far below the bound."""
    nk, u0 = 24, 1e-3
    o = Oracle(dims=3, ni=2, nj=2, nk=nk, tau=(0.8, 1.0), zface_flow=(VEL, PRES), zface_uz=(u0, 0.0),
               rho_top=1.0)
    o.set_solids(np.zeros((nk, 2, 2), np.uint8))
    o.init_rest()
    o.run(40000)
    rho, ux, uy, u = (x.reshape(nk, 2, 2) for x in o.mv(0))
    assert np.max(np.abs(u - u0)) / u0 < 1e-6
    assert np.max(np.abs(rho - 1.0)) < 1e-8
    assert np.abs(ux).max() < 1e-15 and np.abs(uy).max() < 1e-15


def _duct(n, nk):
    mask = np.zeros((nk, n, n), np.uint8)
    mask[:, 0, :] = 1
    mask[:, -1, :] = 1
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    return mask


def test_velocity_inlet_duct_matches_pressure_drive():
    """a square duct fed through a uniform velocity inlet carries the
    prescribed flux rho_A u0 A through every plane at steady state, and
    several widths downstream its profile is the developed one of the
    pressure-driven duct at the same flux (the two-plane means cancel the
    period-2 mode of the pressure face)"""
    n, nk, tau = 10, 48, 0.8
    mask = _duct(n, nk)
    fl = mask[0] == 0
    p = Oracle(dims=3, ni=n, nj=n, nk=nk, tau=(tau, 1), zface_pressure=1, rho_bot=1.0002, rho_top=1.0)
    p.set_solids(mask)
    p.init_rest()
    p.run(5000)
    up = p.mv(0)[3].reshape(nk, n, n)
    m = nk // 2
    ref = (up[m][fl] + up[m + 1][fl]) / 2
    u0 = ref.mean()
    v = Oracle(dims=3, ni=n, nj=n, nk=nk, tau=(tau, 1), zface_flow=(VEL, PRES), zface_uz=(u0, 0.0), rho_top=1.0)
    v.set_solids(mask)
    v.init_rest()
    v.run(5000)
    uv = v.mv(0)[3].reshape(nk, n, n)
    prof = np.array([uv[k][fl].sum() for k in range(nk)])
    pair = (prof[:-1] + prof[1:]) / 2
    # the inlet plane carries exactly the prescribed flux; the others too at steady state
    assert abs(prof[0] / (u0 * fl.sum()) - 1) < 1e-12
    assert np.max(np.abs(pair / (u0 * fl.sum()) - 1)) < 1e-4
    k = 3 * nk // 4
    dev = (uv[k][fl] + uv[k + 1][fl]) / 2
    assert np.max(np.abs(dev - ref)) / ref.max() < 3e-3
    # near the inlet the uniform profile has not developed yet
    assert np.max(np.abs(uv[1][fl] - ref)) / ref.max() > 0.1


@pytest.mark.parametrize("phase_even", [0, 1])
def test_2d_velocity_rows_identities(phase_even):
    """2-D rows: u_y = uy_in on the south row (velocity_s_in) and uy_out on
    the north row (velocity_n_out), u_x = 0; ramped from 0"""
    rng = np.random.default_rng(31)
    ni, nj = 7, 6
    kw = dict(dims=2, ni=ni, nj=nj, tau=(0.9, 1.0), rho_A=(1.2, 1.0), velocity_s_in=1, velocity_n_out=1,
              uy_in=2e-3, uy_out=1e-3)
    o = Oracle(**kw)
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (9, ni * nj)))
    if phase_even:
        o.step(False)
    o.step(True)
    rho, ux, uy, _ = (x.reshape(nj, ni) if x is not None else None for x in o.mv(0))
    np.testing.assert_allclose(uy[0], 2e-3, rtol=0, atol=1e-17)
    np.testing.assert_allclose(uy[-1], 1e-3, rtol=0, atol=1e-17)
    np.testing.assert_allclose(ux[[0, -1]], 0, atol=1e-17)
    # the ramp: 0 before bc_ramp_start, then the linear fraction
    o = Oracle(bc_ramp_start=4, bc_ramp_stop=12, **kw)
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (9, ni * nj)))
    for t in range(1, 9):
        o.step(True)
        uy = o.mv(0)[2].reshape(nj, ni)
        frac = 0.0 if t < 4 else (t - 4) / 8
        np.testing.assert_allclose(uy[0], 2e-3 * frac, rtol=0, atol=1e-17)


def test_2d_velocity_row_plug_flow():
    """2-D, periodic in x, no walls: a velocity row at j = 0 and a pressure
    row at j = nj-1 give the uniform velocity everywhere at steady state"""
    ni, nj, u0 = 3, 24, 1e-3
    o = Oracle(dims=2, ni=ni, nj=nj, tau=(0.8, 1.0), velocity_s_in=1, uy_in=u0, pressure_n_out=(1, 0), rho_out=1.0)
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.init_rest()
    o.run(40000)
    rho, ux, uy, _ = o.mv(0)
    assert np.max(np.abs(uy - u0)) / u0 < 1e-6
    assert np.max(np.abs(rho - 1.0)) < 1e-8 and np.abs(ux).max() < 1e-15


def test_2d_velocity_row_channel_matches_pressure_drive():
    """a 2-D channel with side walls fed by a velocity row: the inlet row
    carries exactly the prescribed flux, and downstream the two-row mean of
    the profile is that of the pressure-driven channel with the same mean
    flux.  (The reference's pressure rows beside bounce-back walls leave a
    period-2 alternation of the row sums, +-9 % here; in the velocity-fed
    channel it starts at the pressure outlet row and dies out towards the
    velocity row, so the two-row means agree to a few 1e-3.)"""
    ni, nj = 12, 40
    mask = np.zeros((nj, ni), np.uint8)
    mask[:, 0] = 1
    mask[:, -1] = 1
    p = Oracle(dims=2, ni=ni, nj=nj, tau=(0.8, 1.0), pressure_s_in=(1, 0), rho_in=1.0005, pressure_n_out=(1, 0),
               rho_out=1.0)
    p.set_solids(mask)
    p.init_rest()
    p.run(20000)
    up = p.mv(0)[2].reshape(nj, ni)
    u0 = up.sum(axis=1).mean() / (ni - 2)
    v = Oracle(dims=2, ni=ni, nj=nj, tau=(0.8, 1.0), velocity_s_in=1, uy_in=u0, pressure_n_out=(1, 0), rho_out=1.0)
    v.set_solids(mask)
    v.init_rest()
    v.run(20000)
    uv = v.mv(0)[2].reshape(nj, ni)
    assert abs(uv[0].sum() / (u0 * (ni - 2)) - 1) < 1e-12
    k = 3 * nj // 4
    dev = np.abs((up[k] + up[k + 1]) / 2 - (uv[k] + uv[k + 1]) / 2).max() / up[k].max()
    assert dev < 5e-3
    # the developed profile is not the uniform inflow
    assert np.abs(uv[k] - u0)[1:-1].max() / u0 > 0.2


def test_zface_solute_kinds_per_face():
    """Dirichlet at z = nk-1 holds c_top; zero gradient at z = 0 leaves no net
    first moment sum_a c_a,z g_a on that face"""
    rng = np.random.default_rng(37)
    ni, nj, nk = 4, 3, 6
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.7), zface_conc=(ZGRAD, DIRI), c_top=0.7,
               flux_plane=0)
    o.set_solids(np.zeros((nk, nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (19, ni * nj * nk)))
    o.set_f(1, rng.uniform(0, 0.05, (19, ni * nj * nk)))
    o.step(True)
    c = o.mv(1)[0].reshape(nk, nj, ni)
    np.testing.assert_allclose(c[-1], 0.7, rtol=0, atol=1e-15)
    flux, _ = o.last_flux()
    assert abs(flux) < 1e-16
    # the flow has no condition: the faces are periodic for it
    o2 = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.7))
    o2.set_solids(np.zeros((nk, nj, ni), np.uint8))
    rng = np.random.default_rng(37)
    o2.set_f(0, rng.uniform(-1e-3, 1e-3, (19, ni * nj * nk)))
    o2.set_f(1, rng.uniform(0, 0.05, (19, ni * nj * nk)))
    o2.step(True)
    assert np.array_equal(o.mv(0)[0], o2.mv(0)[0])

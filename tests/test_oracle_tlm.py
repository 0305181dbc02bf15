"""The FP64 tangent-linear model (oracle/lbm_tlm.cpp) that checks the GPU
adjoint: its objective equals the AA oracle's on the same lattice, and its
directional derivatives equal central finite differences."""
import numpy as np
import pytest

import _pkg
from oracle.pyoracle import Oracle, tlm_objective


def case():
    L = _pkg.load()
    ni, nj, nk = 10, 8, 16
    mask, _ = L.spheres3d(ni, nj, nk, 11, 2.0, 3.5, 0.6, zbuf=3)
    rng = np.random.default_rng(4)
    cmap = rng.uniform(0.8, 1.2, (nj, ni))
    return mask, cmap


def oracle_objective(mask, cmap, params, nsteps, window, zf):
    nk, nj, ni = mask.shape
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(params[3], params[4]),
               gaccel=tuple(params[:3]) + (0.0, 0.0, 0.0), zface_pressure=1, rho_bot=1.001, rho_top=1.0,
               zface_solute=1, c_bot_map=cmap, flux_plane=zf)
    o.set_solids(mask)
    o.init_rest()
    acc = 0.0
    for t in range(1, nsteps + 1):
        o.step(False)
        if t > nsteps - window:
            acc += o.last_flux()[0]
    return acc / (window * ni * nj)


def test_tlm_objective_equals_oracle():
    mask, cmap = case()
    params = (0.0, 0.0, 1e-5, 0.8, 0.75)
    zf = mask.shape[0] - 2
    J_tlm, _ = tlm_objective(mask, 1.001, 1.0, zf, 60, 20, params, cmap)
    J_orc = oracle_objective(mask, cmap, params, 60, 20, zf)
    assert J_orc != 0.0
    assert abs(J_tlm - J_orc) <= 1e-12 * abs(J_orc)


@pytest.mark.parametrize("which", range(6))
def test_tlm_derivative_equals_central_difference(which):
    mask, cmap = case()
    params = np.array([0.0, 0.0, 1e-5, 0.8, 0.75])
    zf = mask.shape[0] - 2
    rng = np.random.default_rng(7)
    dp = np.zeros(5)
    dc = np.zeros_like(cmap)
    if which < 5:
        dp[which] = 1.0
        h = 1e-6 if which >= 3 else 1e-6
    else:
        dc = rng.uniform(-1, 1, cmap.shape)
        h = 1e-5
    _, dJ = tlm_objective(mask, 1.001, 1.0, zf, 50, 15, params, cmap, dp, dc)
    Jp, _ = tlm_objective(mask, 1.001, 1.0, zf, 50, 15, params + h * dp, cmap + h * dc)
    Jm, _ = tlm_objective(mask, 1.001, 1.0, zf, 50, 15, params - h * dp, cmap - h * dc)
    fd = (Jp - Jm) / (2 * h)
    assert abs(dJ - fd) <= 1e-6 * abs(fd) + 1e-12, (dJ, fd)


def observation_case(mask):
    """observation points at fluid nodes of a few planes and steps"""
    rng = np.random.default_rng(21)
    fluid = np.flatnonzero(mask.reshape(-1) != 1)
    nodes = rng.choice(fluid, 7, replace=False)
    steps = np.array([5, 12, 12, 30, 31, 44, 50])
    target = rng.uniform(0.0, 0.3, 7)
    weight = rng.uniform(0.5, 2.0, 7)
    return nodes, steps, target, weight


def test_tlm_observation_misfit_equals_oracle():
    """the misfit term of the TLM equals the AA oracle's PEST-written
    concentrations at the observed steps"""
    mask, cmap = case()
    nk, nj, ni = mask.shape
    params = (0.0, 0.0, 1e-5, 0.8, 0.75)
    obs = observation_case(mask)
    J_tlm, _ = tlm_objective(mask, 1.001, 1.0, nk - 2, 50, 0, params, cmap, observations=obs)
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), gaccel=(0.0, 0.0, 1e-5, 0, 0, 0),
               zface_pressure=1, rho_bot=1.001, rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
    o.set_solids(mask)
    o.init_rest()
    J = 0.0
    for t in range(1, 51):
        here = obs[1] == t
        o.step(2 if here.any() else 0)
        if here.any():
            c = o.mv(1)[0]
            for q in np.flatnonzero(here):
                J += obs[3][q] * (c[obs[0][q]] - obs[2][q]) ** 2
    assert J > 0
    assert abs(J_tlm - J) <= 1e-12 * J


@pytest.mark.parametrize("which", [2, 3, 4, 5])
def test_tlm_observation_derivative_equals_central_difference(which):
    mask, cmap = case()
    params = np.array([0.0, 0.0, 1e-5, 0.8, 0.75])
    zf = mask.shape[0] - 2
    obs = observation_case(mask)
    rng = np.random.default_rng(9)
    dp = np.zeros(5)
    dc = np.zeros_like(cmap)
    if which < 5:
        dp[which] = 1.0
    else:
        dc = rng.uniform(-1, 1, cmap.shape)
    h = 1e-6
    _, dJ = tlm_objective(mask, 1.001, 1.0, zf, 50, 10, params, cmap, dp, dc, observations=obs)
    Jp, _ = tlm_objective(mask, 1.001, 1.0, zf, 50, 10, params + h * dp, cmap + h * dc, observations=obs)
    Jm, _ = tlm_objective(mask, 1.001, 1.0, zf, 50, 10, params - h * dp, cmap - h * dc, observations=obs)
    fd = (Jp - Jm) / (2 * h)
    assert dJ != 0.0
    assert abs(dJ - fd) <= 1e-6 * abs(dJ)


def oracle_permeability(mask, cmap, params, nsteps, window, rho_bot=1.001, rho_top=1.0):
    """rho_A nu <u_z,phys> / G over the window from the AA oracle's
    macroscopic writes: u_phys = u_eq - (tau - 1/2) g (superficial average),
    G = (rho_bot - rho_top)/(3 (LZ - 1)) + rho_A g_z, rho_A = 1"""
    nk, nj, ni = mask.shape
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(params[3], params[4]),
               gaccel=tuple(params[:3]) + (0.0, 0.0, 0.0), zface_pressure=1, rho_bot=rho_bot, rho_top=rho_top,
               zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
    o.set_solids(mask)
    o.init_rest()
    fl = mask.reshape(-1) != 1
    acc = 0.0
    for t in range(1, nsteps + 1):
        o.step(t > nsteps - window)
        if t > nsteps - window:
            acc += np.sum(o.mv(0)[3][fl] - (params[3] - 0.5) * params[2])
    nu = (params[3] - 0.5) / 3
    G = (rho_bot - rho_top) / (3 * (nk - 1)) + params[2]
    return nu * acc / (window * mask.size) / G


def test_tlm_permeability_equals_oracle():
    mask, cmap = case()
    params = (0.0, 0.0, 1e-5, 0.8, 0.75)
    J_tlm, _ = tlm_objective(mask, 1.001, 1.0, 0, 60, 20, params, cmap, objective="permeability")
    J_orc = oracle_permeability(mask, cmap, params, 60, 20)
    assert J_orc > 0.0
    assert abs(J_tlm - J_orc) <= 1e-12 * abs(J_orc)


@pytest.mark.parametrize("which", range(6))
def test_tlm_permeability_derivative_equals_central_difference(which):
    """includes the explicit dependence on tau (through nu) and on g_z
    (through u_phys and the driving gradient)"""
    mask, cmap = case()
    params = np.array([0.0, 0.0, 1e-5, 0.8, 0.75])
    rng = np.random.default_rng(8)
    dp = np.zeros(5)
    dc = np.zeros_like(cmap)
    if which < 5:
        dp[which] = 1.0
        h = 1e-8 if which < 3 else 1e-6
    else:
        dc = rng.uniform(-1, 1, cmap.shape)
        h = 1e-5
    kw = dict(objective="permeability")
    _, dJ = tlm_objective(mask, 1.001, 1.0, 0, 50, 15, params, cmap, dp, dc, **kw)
    Jp, _ = tlm_objective(mask, 1.001, 1.0, 0, 50, 15, params + h * dp, cmap + h * dc, **kw)
    Jm, _ = tlm_objective(mask, 1.001, 1.0, 0, 50, 15, params - h * dp, cmap - h * dc, **kw)
    fd = (Jp - Jm) / (2 * h)
    if which >= 4:
        # the flow does not see the solute: no dependence on tau_s or the inlet map
        assert dJ == 0.0 and fd == 0.0
    else:
        assert abs(dJ - fd) <= 1e-6 * abs(fd) + 1e-12, (dJ, fd)

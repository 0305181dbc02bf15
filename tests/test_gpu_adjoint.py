"""GPU discrete adjoint of the solute-flux objective (lbm_objective_run /
lbm_adjoint_run) against the FP64 tangent-linear model (oracle/lbm_tlm.cpp)
and against central finite differences of the GPU forward run."""
import sys

import numpy as np
import pytest

import _pkg
from oracle.pyoracle import tlm_objective

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(ni=10, nj=8, nk=16, seed=11):
    L = _pkg.load()
    mask, _ = L.spheres3d(ni, nj, nk, seed, 2.0, 3.5, 0.6, zbuf=3)
    cmap = np.random.default_rng(4).uniform(0.8, 1.2, (nj, ni))
    return L, mask, cmap


def gpu_objective(L, mask, cmap, params, nsteps, window, every):
    nk, nj, ni = mask.shape
    lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(params[3], params[4]),
                    gaccel=tuple(params[:3]) + (0.0, 0.0, 0.0), zface_pressure=1, rho_bot=1.001,
                    rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
    lat.init_rest()
    J = lat.objective(nsteps, window, every)
    return lat, J


PARAMS = np.array([2e-6, -1e-6, 1e-5, 0.8, 0.75])


def test_adjoint_matches_tangent_linear_model():
    say("case adjoint vs TLM 10x8x16, 60 steps, window 20, checkpoints every 25")
    L, mask, cmap = setup()
    zf = mask.shape[0] - 2
    lat, J = gpu_objective(L, mask, cmap, PARAMS, 60, 20, 25)
    grad, gmap = lat.adjoint()
    J_tlm, _ = tlm_objective(mask, 1.001, 1.0, zf, 60, 20, PARAMS, cmap)
    say(f"J gpu {J:.15e} tlm {J_tlm:.15e}")
    assert abs(J - J_tlm) <= 1e-12 * abs(J_tlm)
    for q in range(5):
        dp = np.zeros(5)
        dp[q] = 1.0
        _, dJ = tlm_objective(mask, 1.001, 1.0, zf, 60, 20, PARAMS, cmap, dp, np.zeros_like(cmap))
        say(f"param {q}: adjoint {grad[q]:.12e} tlm {dJ:.12e}")
        say(f"param {q}: relative difference {abs(grad[q] - dJ) / max(abs(dJ), 1e-300):.2e}")
        assert abs(grad[q] - dJ) <= 1e-12 * abs(dJ) + 1e-22
    dc = np.random.default_rng(9).uniform(-1, 1, cmap.shape)
    _, dJ = tlm_objective(mask, 1.001, 1.0, zf, 60, 20, PARAMS, cmap, np.zeros(5), dc)
    adj = float(np.sum(gmap * dc))
    say(f"inlet map direction: adjoint {adj:.12e} tlm {dJ:.12e}")
    assert abs(adj - dJ) <= 1e-12 * abs(dJ)
    # the map gradient lives on the inlet face's fluid nodes only
    assert np.all(gmap[mask[0] == 1] == 0.0)


def test_adjoint_matches_central_differences():
    say("case adjoint vs central differences 12x10x20, 80 steps")
    L, mask, cmap = setup(12, 10, 20, seed=5)
    lat, J = gpu_objective(L, mask, cmap, PARAMS, 80, 30, 40)
    grad, gmap = lat.adjoint()
    for q, h in ((2, 1e-8), (3, 1e-6), (4, 1e-6)):
        dp = np.zeros(5)
        dp[q] = h
        _, Jp = gpu_objective(L, mask, cmap, PARAMS + dp, 80, 30, 40)
        _, Jm = gpu_objective(L, mask, cmap, PARAMS - dp, 80, 30, 40)
        fd = (Jp - Jm) / (2 * h)
        say(f"param {q}: adjoint {grad[q]:.10e} fd {fd:.10e}")
        assert abs(grad[q] - fd) <= 1e-5 * abs(fd)
    i, j = 3, 4
    h = 1e-6
    cp, cm = cmap.copy(), cmap.copy()
    cp[j, i] += h
    cm[j, i] -= h
    _, Jp = gpu_objective(L, mask, cp, PARAMS, 80, 30, 40)
    _, Jm = gpu_objective(L, mask, cm, PARAMS, 80, 30, 40)
    fd = (Jp - Jm) / (2 * h)
    say(f"inlet node ({i},{j}): adjoint {gmap[j, i]:.10e} fd {fd:.10e}")
    assert abs(gmap[j, i] - fd) <= 1e-5 * abs(fd)


def observations(mask):
    """points in the interior, on both z faces, and one repeated at the same
    step (its two terms add); steps odd and even, inside and outside the
    flux window, in several checkpoint segments"""
    nk, nj, ni = mask.shape
    rng = np.random.default_rng(33)
    fl = mask.reshape(-1) != 1
    flat = np.arange(mask.size)
    k = flat // (ni * nj)
    interior = flat[fl & (k > 1) & (k < nk - 2)]
    bottom = flat[fl & (k == 0)]
    top = flat[fl & (k == nk - 1)]
    nodes = np.concatenate([rng.choice(interior, 5, replace=False), bottom[:1], top[:1], interior[:1]])
    steps = np.array([3, 14, 27, 27, 46, 9, 51, 3])
    target = rng.uniform(0.0, 0.4, nodes.size)
    weight = rng.uniform(0.5, 2.0, nodes.size)
    nodes[-1] = nodes[0]  # a second observation of the first point, same step
    return nodes, steps, target, weight


@pytest.mark.parametrize("window", [0, 20])
def test_observation_adjoint_matches_tangent_linear_model(window):
    """observation misfit (the PEST points as a calibration objective), with
    and without the flux term"""
    say(f"case observation adjoint vs TLM, window {window}")
    L, mask, cmap = setup()
    zf = mask.shape[0] - 2
    obs = observations(mask)
    nk, nj, ni = mask.shape
    lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(PARAMS[3], PARAMS[4]),
                    gaccel=tuple(PARAMS[:3]) + (0.0, 0.0, 0.0), zface_pressure=1, rho_bot=1.001,
                    rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=zf)
    lat.init_rest()
    lat.observations(*obs)
    J = lat.objective(60, window, 25)
    grad, gmap = lat.adjoint()
    J_tlm, _ = tlm_objective(mask, 1.001, 1.0, zf, 60, window, PARAMS, cmap, observations=obs)
    say(f"J gpu {J:.15e} tlm {J_tlm:.15e}")
    assert J_tlm > 0 and abs(J - J_tlm) <= 1e-12 * abs(J_tlm)
    for q in range(5):
        dp = np.zeros(5)
        dp[q] = 1.0
        _, dJ = tlm_objective(mask, 1.001, 1.0, zf, 60, window, PARAMS, cmap, dp, np.zeros_like(cmap),
                              observations=obs)
        say(f"param {q}: adjoint {grad[q]:.12e} tlm {dJ:.12e}")
        say(f"param {q}: relative difference {abs(grad[q] - dJ) / max(abs(dJ), 1e-300):.2e}")
        assert abs(grad[q] - dJ) <= 1e-12 * abs(dJ) + 1e-22
    dc = np.random.default_rng(19).uniform(-1, 1, cmap.shape)
    _, dJ = tlm_objective(mask, 1.001, 1.0, zf, 60, window, PARAMS, cmap, np.zeros(5), dc, observations=obs)
    adj = float(np.sum(gmap * dc))
    say(f"inlet map direction: adjoint {adj:.12e} tlm {dJ:.12e}")
    assert abs(adj - dJ) <= 1e-12 * abs(dJ)
    # a second sweep of the same run gives the same gradient (the seeds are
    # cleared after their step)
    grad2, gmap2 = lat.adjoint()
    assert np.array_equal(grad, grad2) and np.array_equal(gmap, gmap2)


def test_observation_arguments_checked():
    L, mask, cmap = setup()
    nk, nj, ni = mask.shape
    lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1,
                    zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
    lat.init_rest()
    solid = int(np.flatnonzero(mask.reshape(-1) == 1)[0])
    with pytest.raises(L.LbmError, match="fluid node"):
        lat.observations([solid], [3], [0.1])
    fluid = int(np.flatnonzero(mask.reshape(-1) != 1)[0])
    with pytest.raises(L.LbmError, match="steps count from 1"):
        lat.observations([fluid], [0], [0.1])
    lat.observations([fluid], [40], [0.1])
    with pytest.raises(L.LbmError, match="after the run"):
        lat.objective(30, 10, 10)
    lat.observations([], [], [])
    with pytest.raises(L.LbmError, match="bad objective"):
        lat.objective(30, 0, 10)


def test_adjoint_rejects_unsupported_configurations():
    L, mask, cmap = setup()
    nk, nj, ni = mask.shape
    lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), storage=1,
                    flux_plane=nk - 2, zface_solute=1)
    with pytest.raises(L.LbmError):
        lat.objective(10, 5)
    lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), flux_plane=nk - 2)
    with pytest.raises(L.LbmError):
        lat.adjoint()


def test_permeability_adjoint_matches_tlm_and_differences():
    """the permeability objective: the adjoint gradient against the FP64
    TLM (1e-12) and central differences of the GPU forward run (1e-5); the
    flow does not see the solute, so the tau_s and inlet-map gradients are
    exactly zero"""
    say("case permeability adjoint 10x8x16, 60 steps, window 20, checkpoints every 25")
    L, mask, cmap = setup()
    nk, nj, ni = mask.shape

    def objective(params):
        lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(params[3], params[4]),
                        gaccel=tuple(params[:3]) + (0.0, 0.0, 0.0), zface_pressure=1, rho_bot=1.001,
                        rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
        lat.init_rest()
        return lat, lat.objective(60, 20, 25, kind="permeability")

    lat, J = objective(PARAMS)
    grad, gmap = lat.adjoint()
    J_tlm, _ = tlm_objective(mask, 1.001, 1.0, 0, 60, 20, PARAMS, cmap, objective="permeability")
    say(f"J gpu {J:.15e} tlm {J_tlm:.15e}")
    assert J > 0 and abs(J - J_tlm) <= 1e-12 * abs(J_tlm)
    for q in range(5):
        dp = np.zeros(5)
        dp[q] = 1.0
        _, dJ = tlm_objective(mask, 1.001, 1.0, 0, 60, 20, PARAMS, cmap, dp, np.zeros_like(cmap),
                              objective="permeability")
        say(f"param {q}: adjoint {grad[q]:.12e} tlm {dJ:.12e}")
        # dJ/dg_z is the small difference of the explicit terms (about 2e4)
        # and the dynamics' part, so its relative round-off is ~40x larger
        assert abs(grad[q] - dJ) <= 1e-11 * abs(dJ) + 1e-22
    assert grad[4] == 0.0 and not gmap.any()
    for q, h in ((2, 1e-8), (3, 1e-6)):
        dp = np.zeros(5)
        dp[q] = h
        _, Jp = objective(PARAMS + dp)
        _, Jm = objective(PARAMS - dp)
        fd = (Jp - Jm) / (2 * h)
        say(f"param {q}: adjoint {grad[q]:.10e} fd {fd:.10e}")
        assert abs(grad[q] - fd) <= 1e-5 * abs(fd)
    # the objective kind is part of the handle's state: back to the flux
    _, Jf = gpu_objective(L, mask, cmap, PARAMS, 60, 20, 25)
    J_flux, _ = tlm_objective(mask, 1.001, 1.0, nk - 2, 60, 20, PARAMS, cmap)
    assert abs(Jf - J_flux) <= 1e-12 * abs(J_flux)


@pytest.mark.parametrize("case", ["half_kept", "all_kept", "short_last_segment_kept", "none_kept",
                                  "half_kept_permeability_observations",
                                  "two_thirds_kept_permeability_observations"])
def test_kept_states_give_the_same_gradient(case):
    """lbm_objective_keep_states: the objective run keeps the states of its
    last whole segments and the sweep uses them instead of replaying; J, the
    gradients, the inlet-map gradient and the lattice's state after the run
    equal those of the replaying sweep bit for bit"""
    say(f"case kept states {case}")
    L, mask, cmap = setup()
    # (nsteps, window, every, states the budget holds): 40 of 80 steps kept;
    # all 70 (the last of 3 segments has 20 steps); only the last segment of
    # 10 steps (60 = 2 x 25 + 10); none (9 states hold no whole segment);
    # 20 of 60; 40 of 60.  With at least every + 1 states kept and segments
    # left to replay (40 of 80, 40 of 60) the sweep replays them into the
    # first kept buffers.
    nsteps, window, every, nkeep = {"half_kept": (80, 20, 20, 40), "all_kept": (70, 15, 25, 1000),
                                    "short_last_segment_kept": (60, 20, 25, 10),
                                    "none_kept": (60, 20, 25, 9),
                                    "half_kept_permeability_observations": (60, 20, 20, 30),
                                    "two_thirds_kept_permeability_observations": (60, 20, 20, 40)}[case]
    kept_replay = case in ("half_kept", "two_thirds_kept_permeability_observations")
    # steps kept and checkpoints saved: the last whole segments the budget holds
    kept_steps = {"half_kept": 40, "all_kept": 70, "short_last_segment_kept": 10, "none_kept": 0,
                  "half_kept_permeability_observations": 20,
                  "two_thirds_kept_permeability_observations": 40}[case]
    # the permeability window writes the velocity on its steps, the observed
    # steps the concentration (in kept and in replayed segments)
    kind = "permeability" if "permeability" in case else "flux"
    out = []
    for keep in (False, True):
        nk, nj, ni = mask.shape
        lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(PARAMS[3], PARAMS[4]),
                        gaccel=tuple(PARAMS[:3]) + (0.0, 0.0, 0.0), zface_pressure=1, rho_bot=1.001,
                        rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2)
        if keep:
            # one state: the population array, 2048-node tiles x 2 components x 19 x 8 B
            nf = lat.info()["fluid"]
            state = (nf + 2047) // 2048 * 2048 * 2 * 19 * 8
            lat.keep_states(nkeep * state)
        if "observations" in case:
            lat.observations(*observations(mask))
        lat.init_rest()
        J = lat.objective(nsteps, window, every, kind=kind)
        info = lat.objective_info()
        k = kept_steps if keep else 0
        assert info == {"kept_steps": k, "checkpoints": (nsteps - k + every - 1) // every}, info
        grad, gmap = lat.adjoint()
        f_after = lat.download_f(0), lat.download_f(1)
        out.append((J, grad, gmap, f_after))
        if keep and kept_replay:
            # the kept buffers now hold replayed states: no second sweep
            with pytest.raises(L.LbmError, match="since a sweep replayed into its kept states"):
                lat.adjoint()
        else:
            grad2, gmap2 = lat.adjoint()
            assert np.array_equal(grad2, grad) and np.array_equal(gmap2, gmap)
        lat.close()
    (J0, g0, m0, f0), (J1, g1, m1, f1) = out
    say(f"J {J0:.15e} / {J1:.15e}; grad {g0} / {g1}")
    assert J0 == J1
    assert np.array_equal(g0, g1)
    assert np.array_equal(m0, m1)
    assert np.array_equal(f0[0], f1[0]) and np.array_equal(f0[1], f1[1])

"""Closed forms and invariants that pin the oracle's collision, force,
boundary conditions, halo exchange and reductions."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from npmodel import lattice


def one_node(dims, nsubs, tau, gacc, F, G=None, sink=0.0):
    """one periodic node: the odd half-step's pull and push are identities,
    so the stored slots after one step are the post-collision values in
    opposite slots"""
    o = Oracle(dims=dims, ni=1, nj=1, nk=1, nsubs=nsubs, tau=tau, gaccel=gacc, sink=sink)
    o.set_solids(np.zeros(1, np.uint8))
    o.set_f(0, F.reshape(-1, 1))
    if nsubs == 2:
        o.set_f(1, G.reshape(-1, 1))
    o.step(write_mv=True)
    _, _, opp = lattice(dims)
    post = o.get_f(0)[:, 0][opp]
    postg = o.get_f(1)[:, 0][opp] if nsubs == 2 else None
    return post, postg, o


@pytest.mark.parametrize("dims", [2, 3])
def test_collision_moments_closed_form(dims):
    """M0 = drho, M1 = rho_A u_eq, M2 = rho_A u u + drho/3 I relaxed with 1/tau"""
    rng = np.random.default_rng(11 + dims)
    c, w, _ = lattice(dims)
    tau = 0.83
    for trial in range(20):
        F = rng.uniform(-2e-2, 2e-2, len(w))
        g = np.zeros(6)
        g[:dims] = rng.uniform(-1e-3, 1e-3, dims)
        post, _, o = one_node(dims, 1, (tau, 1.0), tuple(g), F)
        drho = F.sum()
        u = c.T @ F + tau * g[:dims]
        assert abs(post.sum() - drho) < 1e-15
        M1 = c.T @ post
        np.testing.assert_allclose(M1, (1 - 1 / tau) * (c.T @ F) + u / tau, rtol=0, atol=1e-15)
        # post-collision momentum = pre-collision momentum + rho_A g
        np.testing.assert_allclose(M1, c.T @ F + g[:dims], rtol=0, atol=1e-15)
        M2 = np.einsum("a,ai,aj->ij", post, c, c)
        M2pre = np.einsum("a,ai,aj->ij", F, c, c)
        eq2 = np.outer(u, u) + drho / 3 * np.eye(dims)
        np.testing.assert_allclose(M2, (1 - 1 / tau) * M2pre + eq2 / tau, rtol=0, atol=1e-15)
        rho_mv, ux, uy, uz = o.mv(0)
        assert abs(rho_mv[0] - (1 + drho)) < 1e-15
        np.testing.assert_allclose([ux[0], uy[0], uz[0]][:dims], u, atol=1e-16, rtol=0)


@pytest.mark.parametrize("dims", [2, 3])
def test_solute_collision_moments(dims):
    """solute: c conserved, flux relaxes to c u_eq (flow u_eq incl. tau g),
    second moment to c/3 I"""
    rng = np.random.default_rng(5 + dims)
    c, w, _ = lattice(dims)
    tau, taus = 0.9, 0.75
    for trial in range(10):
        F = rng.uniform(-1e-2, 1e-2, len(w))
        G = rng.uniform(0.0, 0.1, len(w))
        g = np.zeros(6)
        g[:dims] = rng.uniform(-1e-3, 1e-3, dims)
        _, postg, o = one_node(dims, 2, (tau, taus), tuple(g), F, G)
        conc = G.sum()
        u = c.T @ F + tau * g[:dims]
        assert abs(postg.sum() - conc) < 1e-15
        np.testing.assert_allclose(c.T @ postg, (1 - 1 / taus) * (c.T @ G) + conc * u / taus,
                                   rtol=0, atol=1e-15)
        M2 = np.einsum("a,ai,aj->ij", postg, c, c)
        M2pre = np.einsum("a,ai,aj->ij", G, c, c)
        np.testing.assert_allclose(M2, (1 - 1 / taus) * M2pre + conc / 3 * np.eye(dims) / taus,
                                   rtol=0, atol=1e-15)


def test_sink_applied_once_odd_twice_even():
    """src/k_stream_collide_stream.c:254-255 (once) vs src/k_collide.c:96-97
    and :259-262 (twice): for a uniform resting solute the concentration
    follows c' = c (1 - 1/tau_s + (1 - s tau_s)^k / tau_s), k = 1 odd, 2 even"""
    taus, sink = 0.75, 0.01
    o = Oracle(dims=2, ni=3, nj=3, nsubs=2, tau=(1.0, taus), sink=sink)
    o.set_solids(np.zeros(9, np.uint8))
    o.init_rest(np.full(9, 0.6))
    cval = 0.6
    for t in range(1, 7):
        o.step(write_mv=True)
        k = 1 if t % 2 else 2
        c_mv = o.mv(1)[0]
        pre = cval
        np.testing.assert_allclose(c_mv, pre * (1 - sink * taus) ** k, rtol=1e-15)
        cval = pre * (1 - 1 / taus + (1 - sink * taus) ** k / taus)


@pytest.mark.parametrize("dims,shape", [(2, (9, 11)), (3, (6, 5, 7))])
def test_rest_state_exact_and_mass_invariant(dims, shape):
    rng = np.random.default_rng(3)
    mask = (rng.random(shape) < 0.35).astype(np.uint8)
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, tau=(0.7, 1))
    o = Oracle(**kw)
    o.set_solids(mask)
    o.init_rest()
    o.run(10)
    assert np.all(o.get_f(0) == 0.0)
    Q = 9 if dims == 2 else 19
    f = rng.uniform(-1e-3, 1e-3, (Q, mask.size))
    o = Oracle(**kw)
    o.set_solids(mask)
    o.set_f(0, f)
    o.step(True)
    m0 = o.reduce_sums()[0]
    for _ in range(9):
        o.run(7)
        assert abs(o.reduce_sums()[0] - m0) < 1e-15 * mask.size


@pytest.mark.parametrize("dims,shape", [(2, (8, 10)), (3, (6, 4, 5))])
def test_force_impulse_law(dims, shape):
    """all-fluid periodic box: sum_x u_eq(t) = sum_x u_eq(1) + (t-1) g N"""
    rng = np.random.default_rng(21)
    tau = 0.8
    g = (2e-5, -1e-5, 3e-5)[:dims]
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1,
              tau=(tau, 1), gaccel=tuple(g) + (0.0,) * (6 - dims))
    o = Oracle(**kw)
    o.set_solids(np.zeros(shape, np.uint8))
    Q = 9 if dims == 2 else 19
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (Q, int(np.prod(shape)))))
    o.step(True)
    s1 = o.reduce_sums()[1:1 + dims].copy()
    N = np.prod(shape)
    for t in range(2, 25):
        o.step(True)
        s = o.reduce_sums()[1:1 + dims]
        np.testing.assert_allclose(s, s1 + (t - 1) * np.asarray(g) * N, rtol=0, atol=1e-14)


def test_shear_wave_viscosity():
    """u_x(y) = U sin(k y) decays as exp(-nu k^2 t), nu = (tau - 1/2)/3"""
    ni, nj, tau, U = 4, 64, 0.9, 1e-4
    c, w, _ = lattice(2)
    y = np.arange(nj)
    k = 2 * np.pi / nj
    ux = U * np.sin(k * y)[:, None] * np.ones((nj, ni))
    cu = c[:, 0, None, None] * ux
    f = w[:, None, None] * (3 * cu + 4.5 * cu ** 2 - 1.5 * ux ** 2)
    o = Oracle(dims=2, ni=ni, nj=nj, tau=(tau, 1))
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.set_f(0, f.reshape(9, -1))
    T = 400
    o.run(T)
    u = o.mv(0)[1].reshape(nj, ni)[:, 0]
    amp = 2 / nj * np.sum(u * np.sin(k * y))
    nu = (tau - 0.5) / 3
    rate = -np.log(amp / U) / T
    assert abs(rate / (nu * k * k) - 1) < 0.02


def test_poiseuille_channel():
    """body-force channel between bounce-back walls (halfway), 2-D:
    u_phys(y) = g/(2 nu) (y - 1/2)(H + 1/2 - y), u_phys = u_eq - (tau - 1/2) g"""
    ni, H, tau, g = 3, 24, 0.8, 1e-6
    nj = H + 2
    mask = np.zeros((nj, ni), np.uint8)
    mask[0] = 1
    mask[-1] = 1
    o = Oracle(dims=2, ni=ni, nj=nj, tau=(tau, 1), gaccel=(g, 0, 0, 0, 0, 0))
    o.set_solids(mask)
    o.init_rest()
    o.run(30000)
    ueq = o.mv(0)[1].reshape(nj, ni)[1:-1, 0]
    uphys = ueq - (tau - 0.5) * g
    nu = (tau - 0.5) / 3
    yy = np.arange(1, H + 1)
    exact = g / (2 * nu) * (yy - 0.5) * (H + 0.5 - yy)
    assert np.max(np.abs(uphys - exact)) / exact.max() < 5e-3
    # profile is symmetric about the centre line
    np.testing.assert_allclose(uphys, uphys[::-1], rtol=1e-10)


def _zou_he_2d_case(north, phase_even):
    rng = np.random.default_rng(2 + north)
    ni, nj = 6, 8
    o = Oracle(dims=2, ni=ni, nj=nj, tau=(0.9, 1),
               pressure_n_out=(1, 0) if north else (0, 0),
               pressure_s_in=(0, 0) if north else (1, 0), rho_in=1.002, rho_out=0.999)
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (9, ni * nj)))
    if phase_even:
        o.step(False)
    o.step(True)
    rho, ux, _, _ = o.mv(0)
    j = nj - 1 if north else 0
    return rho.reshape(nj, ni)[j], ux.reshape(nj, ni)[j], (0.999 if north else 1.002)


@pytest.mark.parametrize("north", [0, 1])
@pytest.mark.parametrize("phase_even", [0, 1])
def test_zou_he_2d_identity(north, phase_even):
    """after bcs_1/bcs_2 the boundary row collides with rho = rho_b, u_x = 0"""
    rho, ux, rb = _zou_he_2d_case(north, phase_even)
    np.testing.assert_allclose(rho, rb, rtol=0, atol=2e-16)
    np.testing.assert_allclose(ux, 0.0, rtol=0, atol=1e-17)


def test_bc_ramp_and_series():
    """rho_b = pfadd + pfmul series[t mod len], ramped from rho_A
    (src/bcs.c:293-319)"""
    ni, nj = 4, 6
    series = np.array([1.001, 1.003, 0.998])
    o = Oracle(dims=2, ni=ni, nj=nj, tau=(1.0, 1), pressure_s_in=(2, 0), pfadd=0.0, pfmul=1.0,
               pressure_s_series=[series, None], bc_ramp_start=2, bc_ramp_stop=6)
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.init_rest()
    for t in range(1, 9):
        o.step(True)
        rho_b = series[t % 3]
        if t < 2:
            rho_b = 1.0
        elif t < 6:
            r = (t - 2) / 4
            rho_b = rho_b * r + (1 - r) * 1.0
        np.testing.assert_allclose(o.mv(0)[0].reshape(nj, ni)[0], rho_b, rtol=0, atol=3e-16)


@pytest.mark.parametrize("north", [0, 1])
def test_zero_gradient_2d_identity(north):
    """unknown incoming solute populations equal w_i c*, c* from the outgoing
    three (src/system_boundary_kernels.c:508-516)"""
    rng = np.random.default_rng(9)
    ni, nj = 5, 7
    o = Oracle(dims=2, ni=ni, nj=nj, nsubs=2, tau=(1.0, 0.8),
               zeroconcgrad_n=north, zeroconcgrad_s=1 - north)
    o.set_solids(np.zeros((nj, ni), np.uint8))
    o.set_f(0, np.zeros((9, ni * nj)))
    G = rng.uniform(0, 0.1, (9, ni * nj))
    o.set_f(1, G)
    o.bcs(1, 1)
    c, w, opp = lattice(2)
    raw = o.get_f(1).reshape(9, nj, ni)
    j = nj - 1 if north else 0
    unk = [4, 6, 8] if north else [3, 5, 7]
    out = [3, 5, 7] if north else [4, 6, 8]
    src_row = 0 if north else nj - 1
    for i in range(ni):
        pulled_out = [raw[a, (j - c[a, 1]) % nj, (i - c[a, 0]) % ni] for a in out]
        cs = sum(pulled_out) / sum(w[a] for a in unk)
        for a in unk:
            val = raw[a, src_row, (i - c[a, 0]) % ni]
            assert abs(val - w[a] * cs) < 1e-16


@pytest.mark.parametrize("phase_even", [0, 1])
def test_zface_bcs_identities(phase_even):
    """3-D z faces: rho = rho_b and j_x = j_y = 0 on both faces; solute c =
    c_b(x,y) on the inlet face and zero net first moment on the outlet face"""
    rng = np.random.default_rng(17)
    ni, nj, nk = 5, 4, 8
    cmap = rng.uniform(0.8, 1.2, (nj, ni))
    o = Oracle(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75),
               zface_pressure=1, rho_bot=1.001, rho_top=1.0, zface_solute=1, c_bot_map=cmap,
               flux_plane=nk - 1)
    o.set_solids(np.zeros((nk, nj, ni), np.uint8))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (19, ni * nj * nk)))
    o.set_f(1, rng.uniform(0, 0.05, (19, ni * nj * nk)))
    if phase_even:
        o.step(False)
    o.step(True)
    rho, ux, uy, _ = o.mv(0)
    rho = rho.reshape(nk, nj, ni)
    np.testing.assert_allclose(rho[0], 1.001, atol=2e-16, rtol=0)
    np.testing.assert_allclose(rho[-1], 1.0, atol=2e-16, rtol=0)
    for u in (ux, uy):
        u = u.reshape(nk, nj, ni)
        np.testing.assert_allclose(u[0], 0, atol=1e-17)
        np.testing.assert_allclose(u[-1], 0, atol=1e-17)
    conc = o.mv(1)[0].reshape(nk, nj, ni)
    np.testing.assert_allclose(conc[0], cmap, atol=1e-15, rtol=0)
    flux, _ = o.last_flux()
    assert abs(flux) < 1e-15


@pytest.mark.parametrize("dims,shape,nsubs", [(3, (6, 5, 4), 2), (2, (7, 6), 2), (3, (5, 4, 6), 1)])
def test_halo_self_exchange_equals_wrap(dims, shape, nsubs):
    """k_bound_* pack/unpack with a periodic self-exchange reproduces the
    fully wrapped run bit for bit"""
    rng = np.random.default_rng(31)
    mask = (rng.random(shape) < 0.3).astype(np.uint8)
    Q = 9 if dims == 2 else 19
    N = mask.size
    f = rng.uniform(-1e-3, 1e-3, (Q, N))
    g = rng.uniform(0, 0.1, (Q, N))
    res = []
    for halo in (0, 1):
        o = Oracle(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1,
                   nsubs=nsubs, tau=(0.8, 0.7), gaccel=(1e-4, 2e-4, -1e-4, 0, 0, 0), halo_mode=halo)
        o.set_solids(mask)
        o.set_f(0, f)
        if nsubs == 2:
            o.set_f(1, g)
        mvs = []
        for t in range(12):
            o.step(True)
            mvs.append(o.mv(0)[0])
            mvs.append(o.mv(0)[3])
        # after an even step the pushed halo contents are back in the planes
        res.append([o.get_f(s) for s in range(nsubs)] + mvs)
    for a, b in zip(*res):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dims,shape", [(3, (6, 5, 4)), (2, (7, 6))])
def test_halo_literal_unpack_rule(dims, shape):
    """halo_literal (the phase-2 unpack of k_bound_HtD_2, which stores every
    received position) changes a run only through face nodes whose
    cross-face upstream node is solid: with fluid face planes it equals the
    masked exchange bit for bit over 12 steps; with solids on the faces, one
    phase-2 unpack after the same odd step leaves states that differ, and
    only in the two face planes"""
    rng = np.random.default_rng(41)
    Q = 9 if dims == 2 else 19
    N = int(np.prod(shape))
    E = N // shape[0]  # nodes per plane across the last axis
    f = rng.uniform(-1e-3, 1e-3, (Q, N))
    g = rng.uniform(0, 0.1, (Q, N))
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=2,
              tau=(0.8, 0.7), gaccel=(1e-4, 2e-4, -1e-4, 0, 0, 0), halo_mode=1)

    def oracle(mask, lit):
        o = Oracle(halo_literal=lit, **kw)
        o.set_solids(mask)
        o.set_f(0, f)
        o.set_f(1, g)
        return o

    open_faces = (rng.random(shape) < 0.3).astype(np.uint8)
    open_faces[0] = 0
    open_faces[-1] = 0
    res = []
    for lit in (0, 1):
        o = oracle(open_faces, lit)
        o.run(12)
        res.append([o.get_f(0), o.get_f(1), o.mv(0)[0]])
    for a, b in zip(*res):
        assert np.array_equal(a, b)

    solid_faces = (rng.random(shape) < 0.35).astype(np.uint8)
    assert solid_faces[0].any() and solid_faces[-1].any()
    states = []
    for lit in (0, 1):
        o = oracle(solid_faces, lit)
        o.step()
        pos, neg = o.halo_pack(2)
        o.halo_unpack(2, pos, neg)
        states.append(np.concatenate([o.get_f(0), o.get_f(1)]))
    diff = np.any(states[0] != states[1], axis=0)
    assert diff.any()
    assert not diff[E:N - E].any()


def test_reductions_brute_force():
    rng = np.random.default_rng(8)
    shape = (6, 7, 5)
    mask = (rng.random(shape) < 0.4).astype(np.uint8)
    o = Oracle(dims=3, ni=5, nj=7, nk=6, nsubs=2, tau=(0.8, 0.8), gaccel=(1e-4, 0, 2e-4, 0, 0, 0))
    o.set_solids(mask)
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (19, mask.size)))
    o.set_f(1, rng.uniform(0, 0.1, (19, mask.size)))
    o.run(3)
    fl = mask.reshape(-1) == 0
    rho, ux, uy, uz = o.mv(0)
    out = o.reduce_ref(0)
    np.testing.assert_allclose(out[:3], [np.abs(rho[fl]).max(), np.abs(rho[fl]).min(), rho[fl].mean()],
                               rtol=1e-14)
    for c, u in enumerate((ux, uy, uz)):
        np.testing.assert_allclose([out[3 + c], out[6 + c], out[9 + c]],
                                   [np.abs(u[fl]).max(), np.abs(u[fl]).min(), u[fl].mean()],
                                   rtol=1e-12, atol=1e-20)
    s = o.reduce_sums()
    conc = o.mv(1)[0]
    np.testing.assert_allclose(s, [(rho[fl] - 1).sum(), ux[fl].sum(), uy[fl].sum(), uz[fl].sum(),
                                   fl.sum(), conc[fl].sum()], rtol=1e-12, atol=1e-18)

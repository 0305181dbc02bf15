"""The oracle's AA-pattern half-steps against the independent two-lattice
numpy model (tests/npmodel.py): streaming, bounce-back, slot swaps, moments,
body force and the solute coupling."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from npmodel import NpLBM


def random_case(dims, shape, nsubs, seed, solid_frac=0.3):
    rng = np.random.default_rng(seed)
    mask = (rng.random(shape) < solid_frac).astype(np.uint8)
    Q = 9 if dims == 2 else 19
    N = int(np.prod(shape))
    f = rng.uniform(-1e-3, 1e-3, (Q, N))
    g = rng.uniform(0.0, 0.05, (Q, N)) if nsubs == 2 else None
    return mask, f, g


def make(dims, shape, nsubs, tau, gacc, mask, f, g):
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1,
              nsubs=nsubs, tau=tau, rho_A=(1.0, 1.0), gaccel=tuple(gacc) + (0.0,) * 3)
    o = Oracle(**kw)
    o.set_solids(mask)
    o.set_f(0, f)
    if nsubs == 2:
        o.set_f(1, g)
    return o


def aa_view(o, subs, t, opp):
    """natural-order view of the oracle state after t steps"""
    raw = o.get_f(subs)
    if t % 2 == 0:
        return raw
    return raw[opp]


@pytest.mark.parametrize("dims,shape", [(2, (7, 9)), (3, (5, 6, 7))])
@pytest.mark.parametrize("nsubs", [1, 2])
@pytest.mark.parametrize("nsteps", [1, 2, 5, 12])
def test_aa_matches_two_lattice(dims, shape, nsubs, nsteps):
    tau = (0.8, 0.7)
    gacc = (1e-4, -2e-4, 3e-4)[:dims] + ((0.0,) if dims == 2 else ())
    mask, f, g = random_case(dims, shape, nsubs, seed=dims * 100 + nsubs * 10 + nsteps)
    o = make(dims, shape, nsubs, tau, gacc[:3], mask, f, g)
    m = NpLBM(dims, shape, mask, tau, g=gacc[:3], nsubs=nsubs)
    fluid = mask.reshape(-1) == 0

    P = f.reshape((m.Q,) + shape).copy()
    G = g.reshape((m.Q,) + shape).copy() if nsubs == 2 else None
    for t in range(1, nsteps + 1):
        F = m.stream(P)
        if nsubs == 2:
            Gs = m.stream(G)
        P, drho, u = m.collide_flow(F)
        if nsubs == 2:
            G, conc = m.collide_solute(Gs, u)
        o.step(write_mv=True)
        rho_o, ux_o, uy_o, uz_o = o.mv(0)
        np.testing.assert_allclose(rho_o[fluid], (drho.reshape(-1) + 1.0)[fluid], rtol=0, atol=1e-15)
        uo = [ux_o, uy_o, uz_o]
        for d in range(dims):
            np.testing.assert_allclose(uo[d][fluid], u[d].reshape(-1)[fluid], rtol=0, atol=1e-17)
        if nsubs == 2:
            c_o = o.mv(1)[0]
            np.testing.assert_allclose(c_o[fluid], conc.reshape(-1)[fluid], rtol=0, atol=1e-14)

    # stored populations: natural after even steps, streamed+swapped after odd
    if nsteps % 2 == 0:
        ref_f, ref_g = P, G
    else:
        ref_f = m.stream(P)
        ref_g = m.stream(G) if nsubs == 2 else None
    got = aa_view(o, 0, nsteps, m.opp)
    np.testing.assert_allclose(got[:, fluid], ref_f.reshape(m.Q, -1)[:, fluid], rtol=0, atol=1e-17)
    if nsubs == 2:
        got = aa_view(o, 1, nsteps, m.opp)
        np.testing.assert_allclose(got[:, fluid], ref_g.reshape(m.Q, -1)[:, fluid], rtol=0, atol=1e-15)


def test_openmp_equals_serial_bitwise():
    from oracle import pyoracle
    dims, shape = 3, (8, 6, 5)
    mask, f, g = random_case(dims, shape, 2, seed=7)
    res = []
    for nt in (1, 4):
        pyoracle.set_threads(nt)
        o = make(dims, shape, 2, (0.8, 0.75), (1e-5, 0, 2e-5), mask, f, g)
        o.run(9)
        res.append((o.get_f(0), o.get_f(1), o.mv(0)[3]))
    pyoracle.set_threads(0)
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dims,shape", [(2, (7, 9)), (3, (5, 6, 7))])
def test_pest_write_selects_components(dims, shape):
    """write_mv 2 (the PEST observation switch) stores the flow velocity and
    the solute concentration of a full frame write, keeps the flow density
    and the solute velocity from before, and steps the populations exactly
    as a step without writes"""
    mask, f, g = random_case(dims, shape, 2, seed=50 + dims)
    a = make(dims, shape, 2, (0.8, 0.7), (1e-4, 2e-4, -1e-4), mask, f, g)
    b = make(dims, shape, 2, (0.8, 0.7), (1e-4, 2e-4, -1e-4), mask, f, g)
    a.run(3)
    b.run(3)
    before = (b.mv(0), b.mv(1))
    for _ in range(2):
        a.step(1)
        b.step(2)
        fa, fb = (a.mv(0), a.mv(1)), (b.mv(0), b.mv(1))
        assert np.array_equal(fb[0][0], before[0][0])
        assert np.array_equal(fb[1][0], fa[1][0])
        for c in range(1, 1 + dims):
            assert np.array_equal(fb[0][c], fa[0][c])
            assert np.array_equal(fb[1][c], before[1][c])
    for s in range(2):
        assert np.array_equal(a.get_f(s), b.get_f(s))

"""The oracle's Walsh partial bounce-back (the reference's WALSH_NS_ON path,
src/k_stream_collide_stream.c:129-135,363-383, src/lattice.c:2936-2939),
pinned by what the model fixes: ns = 0 is the plain update, ns = 1 reverses
every population, mass is conserved, and the total momentum of a uniform-ns
periodic box follows a closed recursion."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle


def make(dims, shape, nsubs, mask=None, tau=(0.8, 0.7), g=(1e-4, -2e-4, 3e-4), seed=3):
    rng = np.random.default_rng(seed)
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=nsubs, tau=tau,
              gaccel=tuple(g) + (0.0, 0.0, 0.0))
    o = Oracle(**kw)
    o.set_solids(np.zeros(shape, np.uint8) if mask is None else mask)
    Q = 9 if dims == 2 else 19
    N = int(np.prod(shape))
    o.set_f(0, rng.uniform(-1e-3, 1e-3, (Q, N)))
    if nsubs == 2:
        o.set_f(1, rng.uniform(0.0, 0.1, (Q, N)))
    return o


@pytest.mark.parametrize("dims,shape", [(2, (9, 11)), (3, (5, 6, 7))])
def test_ns_zero_is_the_plain_update(dims, shape):
    mask = (np.random.default_rng(1).random(shape) < 0.25).astype(np.uint8)
    a = make(dims, shape, 2, mask)
    b = make(dims, shape, 2, mask)
    b.set_ns(np.zeros(shape))
    for t in range(5):
        a.step(True)
        b.step(True)
    for s in range(2):
        assert np.array_equal(a.get_f(s), b.get_f(s))
        for c in range(1 + dims):
            assert np.array_equal(a.mv(s)[c], b.mv(s)[c])


@pytest.mark.parametrize("dims,shape", [(2, (9, 11)), (3, (5, 6, 7))])
def test_ns_one_reverses_every_population(dims, shape):
    """full bounce-back in place: the state repeats every two steps exactly,
    with halfway bounce-back solids present and a body force"""
    mask = (np.random.default_rng(2).random(shape) < 0.25).astype(np.uint8)
    o = make(dims, shape, 2, mask)
    o.set_ns(np.ones(shape))
    f0 = [o.get_f(s) for s in range(2)]
    fl = mask.reshape(-1) != 1
    o.run(2)
    for s in range(2):
        assert np.array_equal(o.get_f(s)[:, fl], f0[s][:, fl])
    o.run(4)
    for s in range(2):
        assert np.array_equal(o.get_f(s)[:, fl], f0[s][:, fl])


def test_mass_conserved_with_fractional_ns():
    shape = (6, 7, 8)
    mask = (np.random.default_rng(5).random(shape) < 0.2).astype(np.uint8)
    o = make(3, shape, 2, mask)
    o.set_ns(np.random.default_rng(6).uniform(0, 1, shape))
    fl = mask.reshape(-1) != 1
    o.step(True)
    m0 = o.mv(0)[0][fl].sum(), o.mv(1)[0][fl].sum()
    for t in range(9):
        o.step(True)
    m1 = o.mv(0)[0][fl].sum(), o.mv(1)[0][fl].sum()
    assert abs(m1[0] - m0[0]) <= 1e-13 * fl.sum()
    assert abs(m1[1] - m0[1]) <= 1e-13 * abs(m0[1])


@pytest.mark.parametrize("ns", [0.1, 0.35, 0.8])
def test_uniform_ns_momentum_recursion(ns):
    """all-fluid periodic box, uniform ns: the pre-collision momentum obeys
    J' = J ((1-ns)(1-w ns) - ns) + (1-ns)^2 rho_A g N (w = 1/tau)"""
    shape = (6, 5, 7)
    tau, g = 0.8, np.array([2e-5, -1e-5, 3e-5])
    o = make(3, shape, 1, tau=(tau, 1.0), g=g)
    o.set_ns(np.full(shape, ns))
    N = int(np.prod(shape))
    w = 1.0 / tau
    r = (1 - ns) * (1 - w * ns) - ns
    J = []
    for t in range(12):
        o.step(True)
        m = o.mv(0)
        U = np.array([m[1].sum(), m[2].sum(), m[3].sum()])
        J.append((U / N - (1 - ns) * tau * g) * N / (1 - ns))  # rho_A = 1
    for t in range(1, 12):
        pred = J[t - 1] * r + (1 - ns) ** 2 * g * N
        assert np.allclose(J[t], pred, rtol=1e-10, atol=1e-16), (t, J[t], pred)

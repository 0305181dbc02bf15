"""The 2-D open boundaries of the reference (Zou-He pressure rows and the
solute zero-gradient rows, src/system_boundary_kernels.c:4-831) in an
independent two-lattice numpy model, against the oracle's in-place (AA)
restatement, to round-off.

The model applies the reference's rules as read from the kernels, in
two-lattice terms (F = populations after streaming, before collision):
- the unknowns of a boundary node come from its knowns after streaming,
  bounce-back values included (north pressure: rho u_y = sum of C, E, W and
  twice N, NE, NW, minus rho_b, plus rho_A; f_S = f_N - 2/3 rho u_y,
  f_SW = f_NE + (f_E - f_W)/2 - rho u_y/6, f_SE = f_NW + (f_W - f_E)/2 -
  rho u_y/6; zero gradient: c* = (sum of the three outgoing)/(sum of the
  three unknown weights), unknown = w c*);
- odd half-step (the _1 kernels write where the pull will read, through
  set_f1d_d, src/lattice.c:2765-2848): an unknown replaces F_a(x) when its
  wrapped source x - c_a is fluid.  When that source is solid, set_f1d_d
  falls back to the node's own slot a, which holds its outgoing population
  P_a(x): F_a(x) keeps the bounce-back value, and the Zou-He value replaces
  what the node x + c_a pulls from that slot (F_a(x + c_a)), or, when
  x + c_a is solid too, the bounce-back input F_opp(a)(x) of x itself;
- even half-step (the _2 kernels write the swapped own slots): the
  unknown always replaces F_a(x).
"""
import numpy as np
import pytest

from npmodel import NpLBM
from oracle.pyoracle import Oracle

RHO_IN, RHO_OUT = 1.002, 0.999


def _unknowns(F, j, north, rho_b, kind):
    """{a: value per column} for the boundary row j"""
    g = F[:, j, :]
    if kind == "pressure":
        if north:
            ruy = g[0] + g[1] + g[2] + 2 * (g[3] + g[5] + g[7]) - rho_b + 1.0
            return {4: g[3] - 2 / 3 * ruy, 6: g[5] + 0.5 * (g[1] - g[2]) - ruy / 6,
                    8: g[7] + 0.5 * (g[2] - g[1]) - ruy / 6}
        ruy = -(g[0] + g[1] + g[2] + 2 * (g[4] + g[6] + g[8]) - rho_b + 1.0)
        return {3: g[4] + 2 / 3 * ruy, 5: g[6] + 0.5 * (g[2] - g[1]) + ruy / 6,
                7: g[8] + 0.5 * (g[1] - g[2]) + ruy / 6}
    w = {3: 1 / 9, 4: 1 / 9, 5: 1 / 36, 6: 1 / 36, 7: 1 / 36, 8: 1 / 36}
    if north:
        cs = (g[3] + g[5] + g[7]) / (w[4] + w[6] + w[8])
        return {4: w[4] * cs, 6: w[6] * cs, 8: w[8] * cs}
    cs = (g[4] + g[6] + g[8]) / (w[3] + w[5] + w[7])
    return {3: w[3] * cs, 5: w[5] * cs, 7: w[7] * cs}


def _apply(F, m, t, rows):
    """rows: list of (j, north, rho_b, kind); t = 1-based step"""
    nj, ni = m.shape
    c, opp, solid = m.c, m.opp, m.solid
    new = [(j, north, _unknowns(F, j, north, rho_b, kind)) for j, north, rho_b, kind in rows]
    for j, north, vals in new:
        for a, v in vals.items():
            for i in range(ni):
                if solid[j, i]:
                    continue
                src_solid = solid[(j - c[a, 1]) % nj, (i - c[a, 0]) % ni]
                if t % 2 == 0 or not src_solid:
                    F[a, j, i] = v[i]
                    continue
                tj, ti = (j + c[a, 1]) % nj, (i + c[a, 0]) % ni
                if not solid[tj, ti]:
                    F[a, tj, ti] = v[i]
                else:
                    F[opp[a], j, i] = v[i]
    return F


def _run_model(mask, P0, G0, T, tau, flow_rows, solute_rows):
    nj, ni = mask.shape
    m = NpLBM(2, (nj, ni), mask, tau, nsubs=2 if G0 is not None else 1)
    P, G = P0.copy(), None if G0 is None else G0.copy()
    for t in range(1, T + 1):
        F = _apply(m.stream(P), m, t, flow_rows)
        P, drho, u = m.collide_flow(F)
        if G is not None:
            H = _apply(m.stream(G), m, t, solute_rows)
            G, conc = m.collide_solute(H, u)
    return drho, u, (conc if G0 is not None else None)


@pytest.mark.parametrize("T", [1, 2, 3, 8, 31])
def test_pressure_rows_match_two_lattice_model(T):
    """south inlet and north outlet with side walls and solids next to the
    boundary rows (the corner rules are exercised)"""
    ni, nj, tau = 11, 12, (0.8, 1.0)
    rng = np.random.default_rng(30 + T)
    mask = np.zeros((nj, ni), np.uint8)
    mask[:, 0] = 1
    mask[:, -1] = 1
    mask[1, 4] = 1
    mask[nj - 2, 7] = 1
    mask[5, 5] = 1
    # corner nodes whose x + c_a is solid as well as their wrapped source
    mask[1, 2] = 1
    mask[nj - 2, ni - 3] = 1
    P0 = rng.uniform(-1e-3, 1e-3, (9, nj, ni))
    P0[:, mask == 1] = 0
    o = Oracle(dims=2, ni=ni, nj=nj, tau=tau, pressure_s_in=(1, 0), pressure_n_out=(1, 0), rho_in=RHO_IN,
               rho_out=RHO_OUT)
    o.set_solids(mask)
    o.set_f(0, P0.reshape(9, -1))
    o.run(T)
    rows = [(0, False, RHO_IN, "pressure"), (nj - 1, True, RHO_OUT, "pressure")]
    drho, u, _ = _run_model(mask, P0, None, T, tau, rows, [])
    fl = mask.reshape(-1) == 0
    rho_o, ux_o, uy_o, _ = o.mv(0)
    assert np.max(np.abs(rho_o - 1.0 - drho.reshape(-1))[fl]) < 1e-15
    assert np.max(np.abs(ux_o - u[0].reshape(-1))[fl]) < 1e-15
    assert np.max(np.abs(uy_o - u[1].reshape(-1))[fl]) < 1e-15


@pytest.mark.parametrize("T", [1, 2, 9])
def test_pressure_and_zero_gradient_rows_with_solute(T):
    ni, nj, tau = 9, 10, (0.9, 0.75)
    rng = np.random.default_rng(70 + T)
    mask = np.zeros((nj, ni), np.uint8)
    mask[:, 0] = 1
    mask[:, -1] = 1
    mask[1, 3] = 1
    mask[nj - 2, 5] = 1
    P0 = rng.uniform(-1e-3, 1e-3, (9, nj, ni))
    G0 = rng.uniform(0.0, 0.1, (9, nj, ni))
    P0[:, mask == 1] = 0
    G0[:, mask == 1] = 0
    o = Oracle(dims=2, ni=ni, nj=nj, nsubs=2, tau=tau, pressure_s_in=(1, 0), pressure_n_out=(1, 0),
               rho_in=RHO_IN, rho_out=RHO_OUT, zeroconcgrad_n=1, zeroconcgrad_s=1)
    o.set_solids(mask)
    o.set_f(0, P0.reshape(9, -1))
    o.set_f(1, G0.reshape(9, -1))
    o.run(T)
    flow = [(0, False, RHO_IN, "pressure"), (nj - 1, True, RHO_OUT, "pressure")]
    sol = [(0, False, None, "zerograd"), (nj - 1, True, None, "zerograd")]
    drho, u, conc = _run_model(mask, P0, G0, T, tau, flow, sol)
    fl = mask.reshape(-1) == 0
    assert np.max(np.abs(o.mv(0)[0] - 1.0 - drho.reshape(-1))[fl]) < 1e-15
    assert np.max(np.abs(o.mv(0)[2] - u[1].reshape(-1))[fl]) < 1e-15
    assert np.max(np.abs(o.mv(1)[0] - conc.reshape(-1))[fl]) < 1e-14


def test_corner_rule_changes_the_interior():
    """the odd-phase fallback is what the reference does, and it matters: a
    model that writes every unknown at the boundary node (textbook Zou-He
    after streaming) differs from the oracle next to the corners"""
    ni, nj, tau = 11, 12, (0.8, 1.0)
    rng = np.random.default_rng(5)
    mask = np.zeros((nj, ni), np.uint8)
    mask[:, 0] = 1
    mask[:, -1] = 1
    P0 = rng.uniform(-1e-3, 1e-3, (9, nj, ni))
    P0[:, mask == 1] = 0
    o = Oracle(dims=2, ni=ni, nj=nj, tau=tau, pressure_s_in=(1, 0), pressure_n_out=(1, 0), rho_in=RHO_IN,
               rho_out=RHO_OUT)
    o.set_solids(mask)
    o.set_f(0, P0.reshape(9, -1))
    o.run(1)
    m = NpLBM(2, (nj, ni), mask, tau)
    F = m.stream(P0)
    for j, north, rho_b in ((0, False, RHO_IN), (nj - 1, True, RHO_OUT)):
        for a, v in _unknowns(F, j, north, rho_b, "pressure").items():
            F[a, j, 1:-1] = v[1:-1]
    _, drho, _ = m.collide_flow(F)
    d = np.abs(o.mv(0)[0].reshape(nj, ni) - 1.0 - drho)
    d[mask == 1] = 0
    assert d[1, 2] > 1e-6 and d[1, ni - 3] > 1e-6  # the nodes pulling the corner slots

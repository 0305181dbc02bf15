"""Physical checks of the 3-D z-face conditions (no reference counterpart):
a pressure-driven square duct against the same duct driven by the
equivalent body force, the steady flux along it, and pure diffusion of the
solute from the Dirichlet inlet against the erfc closed form.

Zou-He pressure faces with halfway bounce-back walls leave a period-2 mode
in the plane sums of u_z (alternating between neighbouring planes and from
step to step; about +-4 % in this duct, +-1.3 % for textbook Zou-He in a
2-D channel) that persists unchanged at steady state.  The checks therefore
use the mean of two neighbouring planes, in which the mode cancels."""
from math import erfc, sqrt

import numpy as np

from oracle.pyoracle import Oracle


def _duct(n, nk):
    mask = np.zeros((nk, n, n), np.uint8)
    mask[:, 0, :] = 1
    mask[:, -1, :] = 1
    mask[:, :, 0] = 1
    mask[:, :, -1] = 1
    return mask


def test_duct_pressure_drive_equals_body_force():
    """steady duct flow driven by rho_bot - rho_top equals the flow driven by
    g = (rho_bot - rho_top) / (3 (LZ - 1)) with periodic ends: the face
    planes are LZ - 1 apart.  An off-by-one length (LZ or LZ - 2) misses by
    about 3 %."""
    n, nk, tau, drho = 10, 32, 0.8, 2e-4
    mask = _duct(n, nk)
    fl = mask[0] == 0
    p = Oracle(dims=3, ni=n, nj=n, nk=nk, tau=(tau, 1), zface_pressure=1, rho_bot=1.0 + drho, rho_top=1.0)
    p.set_solids(mask)
    p.init_rest()
    p.run(4000)
    uz = p.mv(0)[3].reshape(nk, n, n)
    g = drho / 3 / (nk - 1)
    b = Oracle(dims=3, ni=n, nj=n, nk=nk, tau=(tau, 1), gaccel=(0, 0, g, 0, 0, 0))
    b.set_solids(mask)
    b.init_rest()
    b.run(4000)
    ub = b.mv(0)[3].reshape(nk, n, n) - (tau - 0.5) * g  # u_phys = u_eq - (tau - 1/2) g
    m = nk // 2
    pair = (uz[m][fl] + uz[m + 1][fl]) / 2
    ref = ub[m][fl]
    assert np.max(np.abs(pair - ref)) / ref.max() < 2e-3
    assert abs(pair.sum() / ref.sum() - 1) < 1e-3


def test_duct_steady_flux_uniform_along_z():
    """mass conservation at steady state: the two-plane mean of the plane
    flux is the same at every z, the face planes included"""
    n, nk, tau, drho = 10, 32, 0.8, 2e-4
    mask = _duct(n, nk)
    fl = mask[0] == 0
    p = Oracle(dims=3, ni=n, nj=n, nk=nk, tau=(tau, 1), zface_pressure=1, rho_bot=1.0 + drho, rho_top=1.0)
    p.set_solids(mask)
    p.init_rest()
    p.run(4000)
    uz = p.mv(0)[3].reshape(nk, n, n)
    prof = np.array([uz[k][fl].sum() for k in range(nk)])
    pair = (prof[:-1] + prof[1:]) / 2
    assert (pair.max() - pair.min()) / pair.mean() < 1e-4
    # the density falls linearly between the faces
    rho = p.mv(0)[0].reshape(nk, n, n)[:, n // 2, n // 2]
    slope = np.polyfit(np.arange(4, nk - 4), rho[4:nk - 4], 1)[0]
    assert abs(slope / (-drho / (nk - 1)) - 1) < 1e-3


def test_solute_inlet_diffusion_erfc():
    """flow at rest, solute entering by diffusion from the Dirichlet inlet
    (c = 1 on the z = 0 plane): c(z, t) = erfc(z / (2 sqrt(D t))) with
    D = (tau_s - 1/2)/3.  Moving the boundary half a node off z = 0 misses
    by 3.5e-2, a 5 % error in D by about 1e-2."""
    nk, tau_s, T = 160, 1.0, 400
    o = Oracle(dims=3, ni=2, nj=2, nk=nk, nsubs=2, tau=(0.8, tau_s), zface_pressure=1, rho_bot=1.0,
               rho_top=1.0, zface_solute=1, c_bot=1.0)
    o.set_solids(np.zeros((nk, 2, 2), np.uint8))
    o.init_rest()
    o.run(T)
    c = o.mv(1)[0].reshape(nk, 2, 2)
    # the column is uniform across x and y
    assert np.ptp(c, axis=(1, 2)).max() < 1e-15
    D = (tau_s - 0.5) / 3
    exact = np.array([erfc(z / (2 * sqrt(D * T))) for z in range(nk)])
    assert np.max(np.abs(c[:, 0, 0] - exact)) < 2e-3
    # the flow stays at rest
    assert np.abs(o.mv(0)[3]).max() < 1e-15

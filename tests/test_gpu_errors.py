"""Error behaviour of the C ABI past lbm_create: every failure returns a
nonzero status with the reason in lbm_last_error (the reference's
checkCUDAError -> process_exit(1)), and the handle stays usable."""
import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu


def lattice(L, **extra):
    kw = dict(dims=2, ni=9, nj=7, tau=(0.8, 0.8), nsubs=2)
    kw.update(extra)
    return L.Lattice(**kw)


def test_calls_before_and_with_bad_solids():
    L = _pkg.load()
    g = lattice(L)
    with pytest.raises(L.LbmError, match="lbm_set_solids has not been called"):
        g.run(1)
    with pytest.raises(L.LbmError, match="no fluid node"):
        g.set_solids(np.ones((7, 9), np.uint8))
    # a byte other than 1 is fluid (the reference's test is == 1)
    g.set_solids(np.full((7, 9), 255, np.uint8))
    g.init_rest()
    g.run(2)
    assert np.all(np.isfinite(g.mv(0)[0]))
    # a rejected lbm_set_solids releases the previous lattice and leaves
    # none behind: steps fail cleanly until a later call succeeds
    with pytest.raises(L.LbmError, match="no fluid node"):
        g.set_solids(np.ones((7, 9), np.uint8))
    with pytest.raises(L.LbmError, match="lbm_set_solids has not been called"):
        g.run(1)
    with pytest.raises(L.LbmError, match="lbm_set_solids has not been called"):
        g.mv(0)
    g.set_solids(np.zeros((7, 9), np.uint8))
    g.init_rest()
    g.run(2)
    assert g.reduce_sums()[4] == 63


def test_bad_subs_nodes_and_modes():
    L = _pkg.load()
    g = lattice(L, mask=np.zeros((7, 9), np.uint8))
    g.init_rest()
    with pytest.raises(L.LbmError, match="bad subs 2"):
        g.upload_f(2, np.zeros((9, 63)))
    with pytest.raises(L.LbmError, match="bad subs"):
        g.mv(3)
    with pytest.raises(L.LbmError, match="outside the lattice"):
        g.sample_mv(0, [63])
    with pytest.raises(L.LbmError, match="halo_mode = 1"):
        g.halo_periodic(1)
    with pytest.raises(L.LbmError, match="no plane flux"):
        g.last_flux()
    # the handle still works after the failures
    g.run(3)
    assert g.reduce_sums()[4] == 63


def test_objective_errors():
    L = _pkg.load()
    mask, _ = L.spheres3d(8, 6, 12, 5, 1.5, 2.5, 0.7, zbuf=3)
    g = L.Lattice(mask=mask, dims=3, ni=8, nj=6, nk=12, nsubs=2, tau=(0.8, 0.75), zface_pressure=1,
                  rho_bot=1.0, rho_top=1.0, zface_solute=1, c_bot=1.0, flux_plane=9)
    g.init_rest()
    with pytest.raises(L.LbmError, match="unknown objective"):
        L._check(L.lib().lbm_objective_select(g.h, 7))
    with pytest.raises(L.LbmError, match="driving gradient"):
        g.objective(10, 4, 5, kind="permeability")
    with pytest.raises(L.LbmError, match="has not been called"):
        g.adjoint()
    J = g.objective(10, 4, 5)
    assert np.isfinite(J)

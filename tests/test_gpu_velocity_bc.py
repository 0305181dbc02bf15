"""GPU parity of the Zou-He velocity conditions and the per-face z
conditions against the FP64 oracle (bit for bit, the 1e-12 normwise bar
asserted): 3-D faces with velocity and pressure kinds mixed, solute
Dirichlet / zero-gradient faces either way round, fused into the half-step
and as separate launches (flux plane on a face, solute starting late), in
periodic and halo mode; 2-D velocity rows with a ramp beside side walls;
and the adjoint's refusal of face kinds it has no transpose for."""
import numpy as np
import pytest

import _pkg
from oracle.pyoracle import Oracle
from test_gpu_parity import build_pair, compare, load_both, random_state, say

pytestmark = pytest.mark.gpu

VEL, PRES, NONE = 3, 2, 1
DIRI, ZGRAD = 2, 3

CASES = {
    "vel_in_pres_out": dict(zface_flow=(VEL, PRES), zface_uz=(1.5e-3, 0.0), rho_top=0.9995,
                            zface_solute=1),
    "pres_in_vel_out": dict(zface_flow=(PRES, VEL), zface_uz=(0.0, -2e-3), rho_bot=1.001,
                            zface_conc=(ZGRAD, DIRI), c_top=0.8),
    "vel_both": dict(zface_flow=(VEL, VEL), zface_uz=(1e-3, 1e-3), zface_conc=(DIRI, DIRI), c_bot=1.0,
                     c_top=0.25),
    "one_face_only": dict(zface_flow=(NONE, VEL), zface_uz=(0.0, 1e-3), zface_conc=(ZGRAD, NONE)),
    "flux_on_face": dict(zface_flow=(VEL, PRES), zface_uz=(1e-3, 0.0), zface_solute=1, flux_plane=0),
    "solute_starts_late": dict(zface_flow=(VEL, PRES), zface_uz=(1e-3, 0.0), zface_conc=(DIRI, ZGRAD),
                               c_bot=0.9, sigma_t_on=7, sigma_t_off=25),
}


@pytest.mark.parametrize("halo", [0, 1])
@pytest.mark.parametrize("case", sorted(CASES))
def test_3d_zface_kinds_match_oracle(case, halo):
    say(f"case 3D z-face kinds {case} halo={halo}")
    L = _pkg.load()
    ni, nj, nk = 14, 12, 20
    mask, _ = L.spheres3d(ni, nj, nk, 707, 2.5, 5.0, 0.6, zbuf=2)
    rng = np.random.default_rng(9)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), rho_A=(1.0, 1.0), flux_plane=nk - 2,
              c_bot_map=rng.uniform(0.8, 1.2, (nj, ni)), halo_mode=halo, gaccel=(1e-6, 0, 1e-6, 0, 0, 0))
    kw.update(CASES[case])
    if "c_bot" in kw:
        del kw["c_bot_map"]
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 19, mask.size, 2)
    load_both(o, g, f, gg)
    on, off = kw.get("sigma_t_on", 0), kw.get("sigma_t_off", 10**9)
    for t in range(30):
        o.step(True)
        g.step(True)
        if on <= t + 1 <= off:
            fo, fg = o.last_flux(), g.last_flux()
            assert abs(fg[0] - fo[0]) <= 1e-12 * abs(fo[0]) + 1e-300, (t, fo, fg)
    _, bitwise = compare(o, g, mask, 2, f"z-face kinds {case} halo={halo}")
    assert bitwise
    g.close()


@pytest.mark.parametrize("storage", [1, 2])
def test_3d_velocity_faces_reduced_storage(storage):
    """FP32-shifted and FP16-scaled storage with the FP32 collision: the
    velocity inlet holds the prescribed u_z at the face to the storage's
    rounding, and the fields stay within the storage's error of the oracle"""
    say(f"case 3D velocity faces storage={storage}")
    L = _pkg.load()
    ni, nj, nk = 12, 10, 24
    mask, _ = L.spheres3d(ni, nj, nk, 708, 2.0, 4.0, 0.65, zbuf=3)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_flow=(VEL, PRES), zface_uz=(2e-3, 0),
              rho_top=1.0, zface_solute=1, c_bot=1.0)
    o = Oracle(**kw)
    o.set_solids(mask)
    g = L.Lattice(mask=mask, storage=storage, compute=1, store_shift_c=0.0, **kw)
    o.init_rest()
    g.init_rest()
    o.run(200)
    g.run(200)
    fl = mask.reshape(nk, nj, ni) != 1
    uz_o, uz_g = o.mv(0)[3].reshape(nk, nj, ni), g.mv(0)[3].reshape(nk, nj, ni)
    face = fl[0]
    tol = 1e-6 if storage == 1 else 2e-3
    assert np.max(np.abs(uz_g[0][face] - 2e-3)) <= tol * 2e-3
    err = np.max(np.abs(uz_g[fl] - uz_o[fl])) / np.abs(uz_o[fl]).max()
    say(f"storage={storage}: u_z error {err:.2e}")
    assert err <= (1e-5 if storage == 1 else 5e-2)


@pytest.mark.parametrize("nsubs", [1, 2])
def test_2d_velocity_rows_match_oracle(nsubs):
    """2-D: velocity row at j = 0 (ramped), pressure row at j = nj-1, side
    walls beside both rows, solute zero gradient at both rows"""
    say(f"case 2D velocity rows nsubs={nsubs}")
    rng = np.random.default_rng(12)
    ni, nj = 20, 26
    mask = (rng.random((nj, ni)) < 0.12).astype(np.uint8)
    mask[:, 0] = 1
    mask[:, -1] = 1
    kw = dict(dims=2, ni=ni, nj=nj, nsubs=nsubs, tau=(0.9, 0.75), velocity_s_in=1, uy_in=1.5e-3,
              pressure_n_out=(1, 0), rho_out=0.9998, bc_ramp_start=3, bc_ramp_stop=20)
    if nsubs == 2:
        kw.update(zeroconcgrad_n=1, zeroconcgrad_s=1)
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 9, mask.size, nsubs)
    load_both(o, g, f, gg)
    o.run(40)
    g.run(40)
    _, bitwise = compare(o, g, mask, nsubs, f"2D velocity rows nsubs={nsubs}")
    assert bitwise
    # the velocity row holds the full u_y after the ramp (pre-collision)
    uy = g.mv(0)[2].reshape(nj, ni)
    row = mask[0] != 1
    np.testing.assert_allclose(uy[0][row], 1.5e-3, rtol=0, atol=1e-17)


def test_create_rejects_conflicting_conditions():
    L = _pkg.load()
    with pytest.raises(L.LbmError, match="both a pressure and a velocity condition"):
        L.Lattice(dims=2, ni=8, nj=6, tau=(0.8, 0.8), velocity_n_in=1, pressure_n_out=(1, 0))
    with pytest.raises(L.LbmError, match="velocity rows are 2-D"):
        L.Lattice(dims=3, ni=8, nj=6, nk=5, tau=(0.8, 0.8), velocity_s_out=1)
    with pytest.raises(L.LbmError, match="not an lbm_zface_flow"):
        L.Lattice(dims=3, ni=8, nj=6, nk=5, tau=(0.8, 0.8), zface_flow=(4, 0))


def test_adjoint_rejects_face_kinds_without_transpose():
    L = _pkg.load()
    mask, _ = L.spheres3d(8, 6, 12, 5, 1.5, 2.5, 0.7, zbuf=3)
    g = L.Lattice(mask=mask, dims=3, ni=8, nj=6, nk=12, nsubs=2, tau=(0.8, 0.75), zface_flow=(VEL, PRES),
                  zface_uz=(1e-3, 0), zface_solute=1, c_bot=1.0, flux_plane=9)
    g.init_rest()
    with pytest.raises(L.LbmError, match="pressure condition on both faces or on none"):
        g.objective(10, 4, 5)
    g = L.Lattice(mask=mask, dims=3, ni=8, nj=6, nk=12, nsubs=2, tau=(0.8, 0.75), zface_pressure=1,
                  zface_conc=(ZGRAD, DIRI), c_top=1.0, flux_plane=9)
    g.init_rest()
    with pytest.raises(L.LbmError, match="Dirichlet solute inlet"):
        g.objective(10, 4, 5)

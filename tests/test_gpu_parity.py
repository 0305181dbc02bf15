"""HIP product path against the FP64 CPU oracle, through the C ABI.

FP64 storage + FP64 collision is required to match the oracle to 1e-12
(normwise, per field, over fluid nodes); the kernels keep the reference's
operation order without fused multiply-adds, so the expected difference is
zero and the tests also report whether the match is bitwise."""
import sys

import numpy as np
import pytest

import _pkg
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relnorm(a, b):
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 0.0
    return d / s if s > 0 else d


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build_pair(mask, **kw):
    L = _pkg.load()
    o = Oracle(**kw)
    o.set_solids(mask)
    g = L.Lattice(mask=mask, **kw)
    return o, g


def random_state(rng, Q, N, nsubs):
    f = rng.uniform(-1e-3, 1e-3, (Q, N))
    gg = rng.uniform(0.0, 0.1, (Q, N)) if nsubs == 2 else None
    return f, gg


def load_both(o, g, f, gg):
    o.set_f(0, f)
    g.upload_f(0, f)
    if gg is not None:
        o.set_f(1, gg)
        g.upload_f(1, gg)


def compare(o, g, mask, nsubs, label, tol=TOL, compare_f=True):
    fl = np.asarray(mask).reshape(-1) != 1
    worst = 0.0
    bitwise = True
    for s in range(nsubs):
        mo, mg = o.mv(s), g.mv(s)
        for c in range(1 + o.dims):
            r = relnorm(mg[c][fl], mo[c][fl])
            worst = max(worst, r)
            bitwise &= bool(np.array_equal(mg[c][fl], mo[c][fl]))
        if compare_f:
            fo, fg = o.get_f(s), g.download_f(s)
            r = relnorm(fg[:, fl], fo[:, fl])
            worst = max(worst, r)
            bitwise &= bool(np.array_equal(fg[:, fl], fo[:, fl]))
    say(f"{label}: max normwise diff {worst:.3e} bitwise={bitwise}")
    assert worst <= tol, f"{label}: {worst}"
    return worst, bitwise


@pytest.mark.parametrize("dims,shape,nsubs", [(2, (12, 17), 1), (2, (9, 10), 2), (3, (7, 6, 9), 1),
                                               (3, (8, 5, 6), 2)])
@pytest.mark.parametrize("nsteps", [1, 2, 7])
def test_random_state_fp64(dims, shape, nsubs, nsteps):
    say(f"case dims={dims} shape={shape} nsubs={nsubs} nsteps={nsteps}")
    rng = np.random.default_rng(1000 + dims * 10 + nsubs + nsteps)
    mask = (rng.random(shape) < 0.3).astype(np.uint8)
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=nsubs,
              tau=(0.8, 0.7), gaccel=(1e-4, -2e-4, 3e-4, 0, 0, 0))
    o, g = build_pair(mask, **kw)
    Q = 9 if dims == 2 else 19
    f, gg = random_state(rng, Q, mask.size, nsubs)
    load_both(o, g, f, gg)
    o.run(nsteps)
    g.run(nsteps)
    compare(o, g, mask, nsubs, f"random {dims}D nsubs={nsubs} steps={nsteps}")


def test_c1_d2q9_discs_full_length():
    """configuration 1: 128x128 D2Q9, tau=1, walls at y=0,127, discs to
    porosity 0.70, g_x = 1e-6, rest start, 10,000 steps"""
    L = _pkg.load()
    say("case C1 128x128 10000 steps")
    mask, phi = L.discs2d(128, 128, 101, 4.0, 10.0, 0.70, walls=True)
    kw = dict(dims=2, ni=128, nj=128, tau=(1.0, 1.0), gaccel=(1e-6, 0, 0, 0, 0, 0))
    o, g = build_pair(mask, **kw)
    o.init_rest()
    g.init_rest()
    o.run(10000)
    g.run(10000)
    compare(o, g, mask, 1, "C1 10000 steps")
    ro, rg = o.reduce_ref(0), g.reduce_ref(0)
    say("C1 compute.c set oracle", ro[:3], "gpu", rg[:3])
    assert relnorm(rg, ro) <= 1e-12


@pytest.mark.parametrize("phase_kw", [
    dict(pressure_n_out=(1, 0), pressure_s_in=(1, 0), rho_in=1.002, rho_out=0.999),
    dict(pressure_s_in=(2, 0), pressure_n_out=(1, 0), rho_out=1.0, bc_ramp_start=3, bc_ramp_stop=15,
         pfadd=0.5, pfmul=0.5),
])
def test_2d_reference_bcs_with_solute(phase_kw):
    say(f"case 2D BCs {sorted(phase_kw)}")
    rng = np.random.default_rng(77)
    ni, nj = 24, 20
    mask = (rng.random((nj, ni)) < 0.15).astype(np.uint8)
    mask[0] = 0
    mask[-1] = 0
    mask[1] = 0
    mask[-2] = 0
    kw = dict(dims=2, ni=ni, nj=nj, nsubs=2, tau=(0.9, 0.75), zeroconcgrad_n=1, zeroconcgrad_s=1,
              pressure_s_series=[np.array([2.002, 2.006, 1.998, 2.0]), None], **phase_kw)
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 9, mask.size, 2)
    load_both(o, g, f, gg)
    for t in range(40):
        o.step(True)
        g.step(True)
    compare(o, g, mask, 2, "2D Zou-He + zero gradient")


@pytest.mark.parametrize("nsubs", [1, 2])
def test_2d_reference_bcs_at_walled_corners(nsubs):
    """side walls reaching the boundary rows and solids next to them: the
    odd-phase unknowns whose wrapped source is solid go to the boundary
    node's own slot, which the interior node beside it pulls (the
    reference's set_f1d_d fallback)"""
    say(f"case 2D BC corners nsubs={nsubs}")
    rng = np.random.default_rng(88 + nsubs)
    ni, nj = 11, 12
    mask = np.zeros((nj, ni), np.uint8)
    mask[:, 0] = 1
    mask[:, -1] = 1
    mask[1, 4] = 1
    mask[nj - 2, 7] = 1
    mask[1, 2] = 1
    mask[nj - 2, ni - 3] = 1
    mask[5, 5] = 1
    kw = dict(dims=2, ni=ni, nj=nj, nsubs=nsubs, tau=(0.8, 0.75), pressure_s_in=(1, 0), pressure_n_out=(1, 0),
              rho_in=1.002, rho_out=0.999, zeroconcgrad_n=nsubs == 2, zeroconcgrad_s=nsubs == 2)
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 9, mask.size, nsubs)
    load_both(o, g, f, gg)
    for t in range(31):
        o.step(True)
        g.step(True)
    _, bitwise = compare(o, g, mask, nsubs, f"2D BC corners nsubs={nsubs}")
    assert bitwise


@pytest.mark.parametrize("halo", [0, 1])
def test_3d_zface_pressure_solute_flux(halo):
    say(f"case 3D z-face BCs halo={halo}")
    L = _pkg.load()
    ni, nj, nk = 20, 18, 32
    mask, _ = L.spheres3d(ni, nj, nk, 404, 3.0, 6.0, 0.55, zbuf=3)
    rng = np.random.default_rng(5)
    cmap = rng.uniform(0.8, 1.2, (nj, ni))
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, zface_solute=1, c_bot_map=cmap, flux_plane=nk - 2, halo_mode=halo)
    o, g = build_pair(mask, **kw)
    o.init_rest()
    g.init_rest()
    for t in range(1, 61):
        o.step(True)
        g.step(True)
        fo, fg = o.last_flux(), g.last_flux()
        assert abs(fg[0] - fo[0]) <= 1e-12 * max(abs(fo[0]), 1e-300) + 1e-300, (t, fo, fg)
        assert abs(fg[1] - fo[1]) <= 1e-12 * abs(fo[1]) + 1e-300
    compare(o, g, mask, 2, f"3D z-face halo={halo}")
    so, sg = o.reduce_sums(), g.reduce_sums()
    assert relnorm(sg, so) <= 1e-12


@pytest.mark.parametrize("nsubs", [1, 2])
def test_row_band_block_order(nsubs):
    """a wide, shallow lattice where one XCD's band of a plane holds at least
    3 blocks of 512 nodes (nf / nk / 8 >= 1536), so the odd half-step runs in
    the row-band block order: 24 steps with z-face conditions equal the
    oracle bit for bit"""
    say(f"case row-band block order nsubs={nsubs}")
    L = _pkg.load()
    ni, nj, nk = 256, 160, 10
    mask, _ = L.spheres3d(ni, nj, nk, 606, 3.0, 6.0, 0.45, zbuf=2)
    nf = int((mask.reshape(-1) != 1).sum())
    assert nf / nk / 8 >= 3 * 512, nf
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=nsubs, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, gaccel=(1e-6, -2e-6, 0, 0, 0, 0))
    if nsubs == 2:
        kw.update(zface_solute=1, c_bot=1.0, flux_plane=nk - 2)
    o, g = build_pair(mask, **kw)
    f, gg = random_state(np.random.default_rng(66), 19, ni * nj * nk, nsubs)
    load_both(o, g, f, gg)
    o.run(24)
    g.run(24)
    _, bitwise = compare(o, g, mask, nsubs, f"row-band order nsubs={nsubs}")
    assert bitwise
    g.close()


def test_row_band_block_order_fp16():
    """FP16 storage runs its odd half-step in 256-node blocks, with a
    row-band map of its own (nf / nk / 8 >= 3 * 256 here): every node is
    updated exactly once per half-step (a skipped or repeated block would be
    off by far more than FP16 rounding), so 24 steps stay within the storage's
    error of the FP64 oracle"""
    say("case row-band block order fp16")
    L = _pkg.load()
    ni, nj, nk = 256, 160, 10
    mask, _ = L.spheres3d(ni, nj, nk, 606, 3.0, 6.0, 0.45, zbuf=2)
    nf = int((mask.reshape(-1) != 1).sum())
    assert nf / nk / 8 >= 3 * 256, nf
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, gaccel=(1e-6, -2e-6, 0, 0, 0, 0), zface_solute=1, c_bot=1.0, flux_plane=nk - 2)
    o = Oracle(**kw)
    o.set_solids(mask)
    g = L.Lattice(mask=mask, storage=2, compute=1, **kw)
    f, gg = random_state(np.random.default_rng(67), 19, ni * nj * nk, 2)
    load_both(o, g, f, gg)
    o.run(24)
    g.run(24)
    fl = mask.reshape(-1) != 1
    for s in range(2):
        for c in range(4):
            r = relnorm(g.mv(s)[c][fl], o.mv(s)[c][fl])
            say(f"fp16 band order subs {s} field {c}: {r:.2e}")
            assert r <= 5e-3, (s, c, r)
    g.close()


@pytest.mark.parametrize("compute,tol", [(0, 1e-6), (1, 1e-5)])
def test_row_band_block_order_fp32_storage(compute, tol):
    """FP32 storage with the FP64 collision runs both half-steps in small
    workgroups (one-wave odd blocks in the row-band map, two-wave even
    blocks), with the FP32 collision in 512 / 256-thread ones: on a lattice
    where the map is on, 24 steps from a random state stay within the
    storage's error of the FP64 oracle (a skipped or repeated block would be
    off by far more)"""
    say(f"case row-band block order fp32 storage compute={compute}")
    L = _pkg.load()
    ni, nj, nk = 256, 160, 10
    mask, _ = L.spheres3d(ni, nj, nk, 606, 3.0, 6.0, 0.45, zbuf=2)
    nf = int((mask.reshape(-1) != 1).sum())
    assert nf / nk / 8 >= 3 * 512, nf
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, gaccel=(1e-6, -2e-6, 0, 0, 0, 0), zface_solute=1, c_bot=1.0, flux_plane=nk - 2)
    o = Oracle(**kw)
    o.set_solids(mask)
    g = L.Lattice(mask=mask, storage=1, compute=compute, **kw)
    f, gg = random_state(np.random.default_rng(71), 19, ni * nj * nk, 2)
    load_both(o, g, f, gg)
    o.run(24)
    g.run(24)
    fl = mask.reshape(-1) != 1
    for s in range(2):
        for c in range(4):
            r = relnorm(g.mv(s)[c][fl], o.mv(s)[c][fl])
            say(f"fp32 storage compute={compute} subs {s} field {c}: {r:.2e}")
            assert r <= tol, (s, c, r)
    g.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_pest_partial_macro_write_and_upload(dims):
    """LBM_MACRO_PEST writes the flow velocity and the solute concentration
    only; the rest of the buffer keeps what lbm_upload_macro put there"""
    say(f"case PEST macro write dims={dims}")
    rng = np.random.default_rng(31 + dims)
    shape = (11, 13) if dims == 2 else (6, 7, 9)
    mask = (rng.random(shape) < 0.25).astype(np.uint8)
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=2,
              tau=(0.8, 0.7), gaccel=(1e-4, -2e-4, 3e-4, 0, 0, 0), sink=1e-3)
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 9 if dims == 2 else 19, mask.size, 2)
    load_both(o, g, f, gg)
    o.run(3)
    g.run(3)
    compare(o, g, mask, 2, "before PEST steps")
    fl = mask.reshape(-1) != 1
    N = mask.size
    sentinel = [rng.uniform(-1, 1, N) for _ in range(8)]
    g.set_mv(0, *sentinel[0:4])
    g.set_mv(1, *sentinel[4:8])
    back = g.mv(0) + g.mv(1)
    for c in range(1 + dims):
        assert np.array_equal(back[c][fl], sentinel[c][fl])
        assert np.array_equal(back[4 + c][fl], sentinel[4 + c][fl])
        assert not back[c][~fl].any()
    for t in range(2):  # one odd and one even half-step
        o.step(2)
        g.step(2)
        mo0, mo1, mg0, mg1 = o.mv(0), o.mv(1), g.mv(0), g.mv(1)
        assert np.array_equal(mg0[0][fl], sentinel[0][fl])
        for c in range(1, 1 + dims):
            assert np.array_equal(mg0[c][fl], mo0[c][fl])
            assert np.array_equal(mg1[c][fl], sentinel[4 + c][fl])
        assert np.array_equal(mg1[0][fl], mo1[0][fl])
    for s in range(2):
        assert np.array_equal(g.download_f(s)[:, fl], o.get_f(s)[:, fl])
    # single-node transfers (lbm_sample_macro) equal the full download;
    # solid nodes give 0
    nodes = np.array([0, N - 1, int(np.argmax(~fl)), int(np.argmax(fl)), N // 2])
    for s in range(2):
        full = g.mv(s)
        got = g.sample_mv(s, nodes)
        for c in range(1 + dims):
            assert np.array_equal(got[:, c], full[c][nodes])


@pytest.mark.parametrize("case", ["flux_on_face", "solute_starts_late", "flow_only_faces"])
def test_3d_zface_boundary_paths(case):
    """the z-face conditions are formed inside the half-step when they can
    be; the separate boundary launch remains for a flux plane on a face and
    for steps where the solute rests (its condition still writes there), and
    both give the oracle's values bit for bit"""
    say(f"case 3D z-face path {case}")
    L = _pkg.load()
    ni, nj, nk = 14, 12, 20
    mask, _ = L.spheres3d(ni, nj, nk, 405, 2.5, 5.0, 0.6, zbuf=2)
    rng = np.random.default_rng(8)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.002,
              rho_top=0.999, zface_solute=1, c_bot_map=rng.uniform(0.8, 1.2, (nj, ni)), flux_plane=nk - 2)
    if case == "flux_on_face":
        kw["flux_plane"] = 0
    if case == "solute_starts_late":
        kw.update(sigma_t_on=7, sigma_t_off=25)
    if case == "flow_only_faces":
        kw.update(zface_solute=0, gaccel=(0, 0, 1e-5, 0, 0, 0))
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 19, mask.size, 2)
    load_both(o, g, f, gg)
    for t in range(30):
        o.step(True)
        g.step(True)
        if kw["flux_plane"] >= 0 and (kw.get("sigma_t_on", 0) <= t + 1 <= kw.get("sigma_t_off", 10**9)):
            fo, fg = o.last_flux(), g.last_flux()
            assert abs(fg[0] - fo[0]) <= 1e-12 * abs(fo[0]) + 1e-300, (t, fo, fg)
            assert abs(fg[1] - fo[1]) <= 1e-12 * abs(fo[1]) + 1e-300, (t, fo, fg)
    _, bitwise = compare(o, g, mask, 2, f"z-face path {case}")
    assert bitwise


class DeviceBuffer:
    """device memory from the HIP runtime the library already loaded"""

    def __init__(self, nbytes):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        assert self.hip.hipMemset(self.ptr, 0, ctypes.c_size_t(nbytes)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def numpy(self, dtype=np.float64):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.ct.c_void_p), self.ptr,
                                  self.ct.c_size_t(self.nbytes), 2) == 0
        return out

    def __del__(self):
        try:
            self.hip.hipFree(self.ptr)
        except Exception:
            pass


def test_halo_pack_buffers_match_oracle():
    """lbm_halo_pack fills device buffers with the reference's message
    layout [subs][pair][LX*LY] (fluid positions compared); unpacking the
    buffers of a periodic self-exchange reproduces lbm_halo_periodic"""
    say("case halo pack buffers")
    ni, nj, nk = 10, 8, 12
    rng = np.random.default_rng(12)
    mask = (rng.random((nk, nj, ni)) < 0.3).astype(np.uint8)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.7), halo_mode=1)
    o, g = build_pair(mask, **kw)
    f, gg = random_state(rng, 19, mask.size, 2)
    load_both(o, g, f, gg)
    nbytes = g.halo_bytes()
    E = ni * nj
    assert nbytes == 2 * 5 * E * 8
    pos, neg = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    first = mask[0].reshape(-1) != 1
    last = mask[-1].reshape(-1) != 1
    g.halo_pack(1, pos.ptr.value, neg.ptr.value)
    po, no = o.halo_pack(1)
    pg, ng = pos.numpy().reshape(2, 5, E), neg.numpy().reshape(2, 5, E)
    po, no = po.reshape(2, 5, E), no.reshape(2, 5, E)
    assert np.array_equal(pg[:, :, last], po[:, :, last])
    assert np.array_equal(ng[:, :, first], no[:, :, first])
    # explicit pack/unpack through the buffers == the built-in periodic exchange
    g.halo_unpack(1, pos.ptr.value, neg.ptr.value)
    g.bcs_1(1)
    g.stream_collide_stream(1, True)
    g.halo_pack(2, pos.ptr.value, neg.ptr.value)
    g.halo_unpack(2, pos.ptr.value, neg.ptr.value)
    g.bcs_2(2)
    g.collide(2, True)
    g.time = 2
    o.run(2)
    compare(o, g, mask, 2, "explicit halo buffers, 2 steps")


def test_halo_literal_unpack_matches_oracle():
    """halo_literal = 1 (the phase-2 unpack of k_bound_HtD_2, storing every
    received position) on a random lattice with solids on both slab faces:
    10 steps of the GPU path equal the oracle's literal mode bit for bit, and
    they differ from the default (masked) exchange"""
    say("case halo literal unpack")
    ni, nj, nk = 10, 8, 12
    rng = np.random.default_rng(21)
    mask = (rng.random((nk, nj, ni)) < 0.3).astype(np.uint8)
    f, gg = random_state(rng, 19, mask.size, 2)
    runs = {}
    for lit in (1, 0):
        kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.7), halo_mode=1, halo_literal=lit,
                  gaccel=(1e-5, 0, 2e-5, 0, 0, 0))
        o, g = build_pair(mask, **kw)
        load_both(o, g, f, gg)
        o.run(10)
        g.run(10)
        _, bitwise = compare(o, g, mask, 2, f"halo literal={lit}, 10 steps")
        assert bitwise
        runs[lit] = g.download_f(0)
        g.close()
    fl = mask.reshape(-1) != 1
    assert not np.array_equal(runs[1][:, fl], runs[0][:, fl])


@pytest.mark.parametrize("storage,compute,tol", [(0, 1, 1e-5), (1, 1, 1e-5), (1, 0, 1e-6), (2, 1, 1e-2)])
def test_reduced_precision_error_vs_fp64(storage, compute, tol):
    """storage FP64/FP32/FP16 with FP32 collision against the FP64 oracle
    (bounds are stated error levels, not parity)"""
    say(f"case precision storage={storage}")
    L = _pkg.load()
    ni, nj, nk = 24, 20, 40
    mask, _ = L.spheres3d(ni, nj, nk, 303, 3.0, 6.0, 0.5, zbuf=3)
    kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, zface_solute=1, c_bot=1.0)
    o = Oracle(**kw)
    o.set_solids(mask)
    o.init_rest()
    g = L.Lattice(mask=mask, storage=storage, compute=compute, store_shift_c=0.5 if storage == 2 else 0.0,
                  **kw)
    g.init_rest()
    o.run(300)
    g.run(300)
    fl = mask.reshape(-1) != 1
    mo, mg = o.mv(0), g.mv(0)
    er = relnorm(mg[0][fl] - 1, mo[0][fl] - 1)
    eu = max(relnorm(mg[c][fl], mo[c][fl]) for c in (1, 2, 3))
    ec = relnorm(g.mv(1)[0][fl], o.mv(1)[0][fl])
    say(f"storage={storage} compute={compute}: drho err {er:.2e} u err {eu:.2e} c err {ec:.2e}")
    assert eu <= tol and ec <= tol


def test_fp32_collision_permeability_within_1e_6():
    """FP32-shifted storage with the FP32 collision: the permeability of a
    pressure-driven sphere pack stays within 1e-6 of the FP64 run (the
    collision closes its mass and momentum exactly, and the boundary density
    offsets are formed in FP64)"""
    say("case FP32 collision permeability")
    L = _pkg.load()
    n, steps = 48, 1500
    mask, _ = L.spheres3d(n, n, n, 77, 3.0, 6.0, 0.45, zbuf=4)
    kw = dict(dims=3, ni=n, nj=n, nk=n, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
              rho_top=1.0, zface_solute=1, c_bot=1.0)
    k = {}
    for st, ct in ((0, 0), (1, 1)):
        g = L.Lattice(mask=mask, storage=st, compute=ct, **kw)
        g.init_rest()
        g.run(steps)
        s = g.reduce_sums()
        nu = (0.8 - 0.5) / 3
        k[st] = nu * (s[3] / n ** 3) / ((1.001 - 1.0) / 3 / (n - 1))
        g.close()
    err = abs(k[1] - k[0]) / abs(k[0])
    say(f"permeability FP64 {k[0]:.10e} FP32 {k[1]:.10e} relative {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("compute", [0, 1])
def test_2d_reference_bcs_fp32_storage(compute):
    """D2Q9 with the reference's N/S pressure rows and zero-gradient solute
    rows in FP32-shifted storage (FP64 or FP32 collision) against the FP64
    oracle: the FP32 paths of the 2-D boundary kernels form the density
    offset in FP64, so the mean flow stays within 1e-6"""
    say(f"case 2D BCs FP32 storage compute={compute}")
    L = _pkg.load()
    ni, nj = 40, 64
    mask, _ = L.discs2d(ni, nj, 31, 2.0, 4.0, 0.75, walls=False)
    mask = mask.copy()
    mask[:3] = 0
    mask[-3:] = 0
    kw = dict(dims=2, ni=ni, nj=nj, nsubs=2, tau=(0.8, 0.75), pressure_s_in=(1, 0), pressure_n_out=(1, 0),
              rho_in=1.001, rho_out=1.0, zeroconcgrad_n=1, zeroconcgrad_s=1)
    o = Oracle(**kw)
    o.set_solids(mask)
    g = L.Lattice(mask=mask, storage=1, compute=compute, **kw)
    c0 = np.zeros(mask.size)
    c0[: ni * 8] = 1.0
    o.init_rest(c0)
    g.init_rest(c0)
    o.run(2000)
    g.run(2000)
    fl = mask.reshape(-1) != 1
    mo, mg = o.mv(0), g.mv(0)
    ey = relnorm(mg[2][fl], mo[2][fl])
    em = abs(mg[2][fl].sum() / mo[2][fl].sum() - 1)
    ec = relnorm(g.mv(1)[0][fl], o.mv(1)[0][fl])
    say(f"compute={compute}: u_y {ey:.2e}, mean u_y {em:.2e}, c {ec:.2e}")
    assert em <= 1e-6 and ey <= 1e-5 and ec <= 1e-5


@pytest.mark.parametrize("case", ["2d_pressure_rows_solute", "3d_zface_solute_walsh", "2d_ramp_not_captured",
                                  "3d_zface_solute_fp32", "3d_zface_solute_fp16", "3d_solute_window"])
def test_captured_step_chunks_equal_stepwise_launches(case):
    """lbm_run launches its steps in captured HIP graphs of 64 steps where
    the launches repeat; three runs (the graphs replayed several times, with
    eager steps, a parity change and synchronisations between them) equal
    the same steps launched one by one, bit for bit, after every run.
    3d_solute_window: the solute is active in steps 150..360 only; runs
    wholly before, wholly inside and wholly after that window replay graphs
    captured on their own side of it (a graph captured before the window
    must not be replayed inside it, nor one captured inside after it)."""
    say(f"case captured chunks {case}")
    L = _pkg.load()
    rng = np.random.default_rng(515)
    if case.startswith("2d"):
        ni, nj = 40, 32
        mask = (rng.random((nj, ni)) < 0.15).astype(np.uint8)
        mask[:2] = 0
        mask[-2:] = 0
        kw = dict(dims=2, ni=ni, nj=nj, nsubs=2, tau=(0.9, 0.75), zeroconcgrad_n=1, zeroconcgrad_s=1,
                  pressure_s_in=(1, 0), pressure_n_out=(1, 0), rho_in=1.002, rho_out=0.999)
        if case == "2d_ramp_not_captured":
            kw.update(bc_ramp_start=3, bc_ramp_stop=500)
        Q, N, ns = 9, ni * nj, None
    else:
        ni, nj, nk = 16, 12, 20
        mask, _ = L.spheres3d(ni, nj, nk, 616, 2.0, 3.5, 0.6, zbuf=3)
        kw = dict(dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(0.8, 0.75), zface_pressure=1, rho_bot=1.001,
                  rho_top=1.0, zface_solute=1, c_bot=1.0, flux_plane=nk - 2, gaccel=(1e-6, 0, 2e-6, 0, 0, 0))
        Q, N = 19, ni * nj * nk
        ns = np.where(rng.random(N) < 0.1, 0.3, 0.0)
        if case.endswith("fp32"):
            kw.update(storage=1, compute=1)
            ns = None
        elif case.endswith("fp16"):
            kw.update(storage=2, compute=1, store_shift_c=0.5)
            ns = None
        elif case == "3d_solute_window":
            kw.update(sigma_t_on=150, sigma_t_off=360)
            ns = None
    f, gg = random_state(rng, Q, N, 2)
    lats = []
    for _ in range(2):
        lat = L.Lattice(mask=mask, **kw)
        if ns is not None:
            lat.set_ns(ns)
        lat.upload_f(0, f)
        lat.upload_f(1, gg)
        lats.append(lat)
    a, b = lats
    runs = (140, 10, 150, 60, 150) if case == "3d_solute_window" else (150, 1, 200, 1, 129)
    for n in runs:
        if n == 1:
            a.step()
        else:
            a.run(n, write_last=True)
        for q in range(n):
            b.step(n > 1 and q == n - 1)
        for s in range(2):
            assert np.array_equal(a.download_f(s), b.download_f(s)), (n, s)
        if n > 1:
            for s in range(2):
                assert all(np.array_equal(x, y) for x, y in zip(a.mv(s), b.mv(s))), (n, s)
    say(f"{case}: equal after {sum(runs)} steps")
    a.close()
    b.close()

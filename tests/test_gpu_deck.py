"""A whole input deck through lbm_deck_main (the reference's main loop on
the HIP library) against the FP64 oracle stepped over the same deck: the
per-frame .dat files must hold the oracle's macroscopic fields bit for bit,
the text files and frames.dat the same values at their print precision."""
import os
import struct
import sys

import numpy as np
import pytest

import _pkg
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_in")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def write_deck(path, edits):
    """the reference's sample deck with some entries replaced"""
    lines = open(os.path.join(GOLD, "params.in")).read().split("\n")
    index = {ln.split()[0]: q for q, ln in enumerate(lines) if ln.strip()}
    for lab, val in edits.items():
        lines[index[lab]] = f"{lab} {val}"
    open(path, "w").write("\n".join(lines))


def write_bmp24(path, mask):
    h, w = mask.shape
    row = (3 * w + 3) // 4 * 4
    data = bytearray()
    for j in range(h):
        r = b"".join(b"\x00\x00\x00" if mask[j, i] else b"\xff\xff\xff" for i in range(w))
        data += r + b"\x00" * (row - 3 * w)
    hdr = struct.pack("<2sIHHI", b"BM", 54 + len(data), 0, 0, 54)
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(data), 2835, 2835, 0, 0)
    open(path, "wb").write(hdr + info + data)


def read_txt(path, LX, LY):
    rows = [r for r in open(path).read().split("\n") if r.strip()]
    assert len(rows) == LY
    return np.array([[float(v) for v in r.split()] for r in rows]).reshape(-1)


def frames_rows(path):
    out = []
    for ln in open(path).read().split("\n")[3:]:
        if ln.strip():
            out.append([float(v) for v in ln[:64].split() + [ln[64:76]] + ln[76:].split()])
    return out


@pytest.mark.parametrize("nsubs", [1, 2])
def test_deck_run_matches_oracle(tmp_path, nsubs):
    say(f"case deck main nsubs={nsubs}")
    io = _pkg.load_submodule("deck")
    LX, LY, FR, NF = 48, 32, 7, 3
    rng = np.random.default_rng(60 + nsubs)
    mask = (rng.random((LY, LX)) < 0.2).astype(np.uint8)
    mask[:2] = 0
    mask[-2:] = 0
    ind = tmp_path / "in"
    ind.mkdir()
    write_bmp24(str(ind / f"{LX}x{LY}.bmp"), mask)
    ser_n = np.array([0.004, 0.006, 0.002, 0.005, 0.003])
    ser_s = np.array([-0.002, -0.004, -0.001])
    (ind / "pressure_n_in0.in").write_text(" ".join(f"{v:.3f}" for v in ser_n) + "\n")
    (ind / "pressure_s_in0.in").write_text("\n".join(f"{v:.3f}" for v in ser_s) + "\n")
    edits = {"LX": LX, "LY": LY, "BX": 16, "BY": 8, "NumFrames": NF, "FrameRate": FR,
             "tau[0]": "0.8", "gforce_x[0]": "0.00001", "tau[1]": "0.7", "pfmul": "1.0", "pfadd": "1.0",
             "bc_ramp_start": 2, "bc_ramp_stop": 9, "sigma_t_on": 0, "sigma_t_off": -1,
             "x1": 10, "x2": 20, "y1": 8, "y2": 20, "rho_B": "1.01", "rho_sigma": "1.0", "sink": "0.001"}
    write_deck(str(ind / "params.in"), edits)
    # PEST observation points: odd, even and frame-end times, two points at
    # one time, one solid point and one outside the lattice (dropped)
    sy, sx = np.argwhere(mask == 1)[0]
    obs = [(3, 5, 6), (4, 30, 20), (4, 12, 14), (7, 40, 3), (9, int(sx), int(sy)), (11, 60, 4), (16, 15, 10)]
    (ind / "timestep_file.in").write_text("".join(f"{t}\n" for t, _, _ in obs))
    (ind / "x_coord_file.in").write_text(" ".join(str(x) for _, x, _ in obs) + "\n")
    (ind / "y_coord_file.in").write_text(" ".join(str(y) for _, _, y in obs) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    rc = io.deck_main("--dims", "2", "--subs", str(nsubs), "--in", str(ind), "--out", str(out))
    assert rc == 0

    cube = np.zeros((LY, LX), bool)
    cube[8:21, 10:21] = True
    kw = dict(dims=2, ni=LX, nj=LY, nsubs=nsubs, tau=(0.8, 0.7), rho_A=(1.0, 1.01),
              gaccel=(1e-5, 0, 0, 0, 0, 0), sink=0.001, sigma_t_on=0, sigma_t_off=NF * FR,
              pressure_n_in=(2, 0), pressure_s_in=(2, 0), rho_in=1.00001, rho_out=0.99999,
              pfadd=1.0, pfmul=1.0, bc_ramp_start=2, bc_ramp_stop=9,
              zeroconcgrad_n=nsubs == 2, zeroconcgrad_s=nsubs == 2,
              pressure_n_series=[ser_n, None], pressure_s_series=[ser_s, None])
    o = Oracle(**kw)
    o.set_solids(mask)
    c0 = np.where(cube, 1.0, 0.0).reshape(-1)
    o.init_rest(c0 if nsubs == 2 else None)
    # init_problem's fields: flow rho_A (rho_B in the cube for one component),
    # solute c0; zero velocity.  Solid nodes keep them in every frame.
    rho0 = [np.where(cube, 1.01, 1.0).reshape(-1) if nsubs == 1 else np.ones(LX * LY), c0]
    solid = mask.reshape(-1) == 1
    fluid = ~solid

    def name(prefix, frame, s, suffix):
        return io.filename(str(out), 2, LX, LY, nsubs, prefix, frame, s, suffix)

    expected_rows = [[] for _ in range(nsubs)]
    for s in range(nsubs):
        z = np.zeros(LX * LY)
        expected_rows[s].append(io.frame_stats(rho0[s], z, z, z, mask))
        assert np.max(np.abs(read_txt(name("rho", 0, s, ".txt"), LX, LY) - rho0[s])) <= 1e-15
    obs_times = {t for t, _, _ in obs}
    conc, vel, order = [], [], []
    time = 0
    for frame in range(1, NF + 1):
        for _ in range(FR):
            time += 1
            o.step(1 if time % FR == 0 else (2 if time in obs_times else 0))
            for q, (t, x, y) in enumerate(obs):
                if t != time or x >= LX:
                    continue
                n = y * LX + x
                if solid[n]:
                    c, ux, uy = rho0[1][n] if nsubs == 2 else 0.0, 0.0, 0.0
                else:
                    c = o.mv(1)[0][n] if nsubs == 2 else 0.0
                    ux, uy = o.mv(0)[1][n], o.mv(0)[2][n]
                conc.append(c)
                vel.append(np.sqrt(ux * ux + uy * uy))
                order.append(q)
        for s in range(nsubs):
            mo = o.mv(s)
            fields = {"rho": mo[0], "u_x": mo[1], "u_y": mo[2]}
            for key, want in fields.items():
                got = np.fromfile(name(key, frame, s, ".dat"), np.float64)
                assert np.array_equal(got[fluid], want[fluid]), (frame, s, key)
                assert not got[solid].any()
            full = [np.where(solid, rho0[s], mo[0]), np.where(solid, 0.0, mo[1]), np.where(solid, 0.0, mo[2])]
            for key, want in zip(("rho", "ux", "uy"), full):
                got = read_txt(name(key, frame, s, ".txt"), LX, LY)
                assert np.max(np.abs(got - want)) <= 1e-15, (frame, s, key)
            z = np.zeros(LX * LY)
            expected_rows[s].append(io.frame_stats(full[0], full[1], full[2], z, mask))
        say(f"frame {frame} matches")
    for s in range(nsubs):
        rows = frames_rows(name("frames", -1, s, ".dat"))
        assert [int(r[0]) for r in rows] == [0, FR, 2 * FR, 3 * FR]
        for r, st in zip(rows, expected_rows[s]):
            mn, mx, av = st[6:9], st[3:6], st[9:12]
            want = [r[0], 0, 0, 0, 0] + list(mn) + list(mx) + list(av)
            assert np.allclose(r[1:14], want[1:14], rtol=0, atol=6e-8)
            assert np.allclose(r[17:20], [st[1], st[0], st[2]], rtol=0, atol=6e-8)
    got_c = [float(v) for v in (out / "conc_data_proc0000.dat").read_text().split()]
    got_v = [float(v) for v in (out / "vel_data_proc0000.dat").read_text().split()]
    got_o = [int(v) for v in (out / "conc_order_proc0000.dat").read_text().split()]
    assert got_o == order == [0, 1, 2, 3, 4, 6]
    assert np.allclose(got_c, conc, rtol=0, atol=1e-17) and np.allclose(got_v, vel, rtol=0, atol=1e-17)
    assert max(vel) > 0 and (nsubs == 1 or max(conc) > 0)
    if nsubs == 2:
        # the solute left the cube and reached the fluid around it
        assert o.mv(1)[0][fluid & ~cube.reshape(-1)].max() > 0


def test_walsh_deck_matches_oracle(tmp_path):
    """--walsh: the image's black pixels carry params.in's ns and every node
    is computed (a WALSH_NS_ON + COMPUTE_ON_SOLIDS build); the .dat files
    keep the image's solids as zeros, the text files hold every node"""
    say("case deck main --walsh")
    io = _pkg.load_submodule("deck")
    LX, LY, FR, NF = 32, 24, 5, 2
    rng = np.random.default_rng(71)
    mask = (rng.random((LY, LX)) < 0.25).astype(np.uint8)
    mask[:2] = 0
    mask[-2:] = 0
    ind = tmp_path / "in"
    ind.mkdir()
    write_bmp24(str(ind / f"{LX}x{LY}.bmp"), mask)
    edits = {"LX": LX, "LY": LY, "BX": 16, "BY": 8, "NumFrames": NF, "FrameRate": FR, "tau[0]": "0.8",
             "gforce_x[0]": "0.00001", "tau[1]": "0.7", "ns": "0.35", "sigma_t_on": 0, "sigma_t_off": -1,
             "x1": 10, "x2": 20, "y1": 8, "y2": 16, "rho_sigma": "1.0", "pressure_n_in[0]": 0,
             "pressure_s_in[0]": 0}
    write_deck(str(ind / "params.in"), edits)
    out = tmp_path / "out"
    out.mkdir()
    assert io.deck_main("--dims", "2", "--subs", "2", "--in", str(ind), "--out", str(out), "--walsh") == 0
    assert io.deck_main("--dims", "3", "--in", str(ind), "--out", str(out), "--walsh") == 1

    cube = np.zeros((LY, LX), bool)
    cube[8:17, 10:21] = True
    deck = io.read_deck(str(ind / "params.in"), 2, 2)
    kw = dict(dims=2, ni=LX, nj=LY, nsubs=2, tau=(0.8, 0.7), rho_A=(1.0, 1.0), gaccel=(1e-5, 0, 0, 0, 0, 0),
              sink=deck.sink, sigma_t_on=0, sigma_t_off=NF * FR,
              zeroconcgrad_n=deck.zeroconcgrad_n, zeroconcgrad_s=deck.zeroconcgrad_s)
    o = Oracle(**kw)
    o.set_solids(np.zeros_like(mask))
    o.set_ns(np.where(mask == 1, 0.35, 0.0))
    o.init_rest(np.where(cube, 1.0, 0.0).reshape(-1))
    solid = mask.reshape(-1) == 1
    for frame in range(1, NF + 1):
        o.run(FR)
        for s in range(2):
            mo = o.mv(s)
            for key, want in zip(("rho", "u_x", "u_y"), mo):
                got = np.fromfile(io.filename(str(out), 2, LX, LY, 2, key, frame, s, ".dat"), np.float64)
                assert np.array_equal(got[~solid], want[~solid]), (frame, s, key)
                assert not got[solid].any()
            for key, want in zip(("rho", "ux", "uy"), mo):
                got = read_txt(io.filename(str(out), 2, LX, LY, 2, key, frame, s, ".txt"), LX, LY)
                assert np.max(np.abs(got - want)) <= 1e-15, (frame, s, key)
        say(f"walsh frame {frame} matches")
    # the partially solid pixels moved: they are computed, not held
    assert np.abs(o.mv(0)[1][solid]).max() > 0

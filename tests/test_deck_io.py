"""Input-deck readers and output writers of the host-format layer
(include/lbm_io.h) on the CPU, pinned by the input files the reference ships
(tests/golden/ref_in: its params.in deck and solids images) and by
independent decodes of the formats."""
import os
import struct

import numpy as np
import pytest

import _pkg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_in")


@pytest.fixture(scope="module")
def io():
    return _pkg.load_submodule("deck")


def deck_lines():
    """label and value text of each line of the reference deck"""
    out = []
    with open(os.path.join(GOLD, "params.in")) as f:
        for line in f:
            lab, val = line.split()
            out.append((lab, val))
    return out


# deck labels whose value belongs to the second component (stored only
# when the build has two components)
def second_component(label):
    solute = ("rho_sigma", "u_sigma", "sigma_", "bc_sigma_slip", "constcon", "constflx", "zeroconcgrad")
    return label.endswith("[1]") or label.startswith(solute)


def field_of(deck, label):
    base = label.split("[")[0]
    idx = int(label[label.index("[") + 1]) if "[" in label else None
    alias = {"characteristic_length": "length_scale", "rho_A": "rho_A", "rho_B": "rho_B"}
    if base.startswith("gforce_"):
        return deck.gforce[idx]["xyz".index(base[-1])]
    name = alias.get(base, base)
    v = getattr(deck, name)
    if idx is not None:
        return v[idx]
    if name in ("rho_A", "rho_B"):
        return v[0]
    return v


@pytest.mark.parametrize("nsubs", [1, 2])
def test_reference_deck_every_entry(io, nsubs):
    d = io.read_deck(os.path.join(GOLD, "params.in"), dims=2, nsubs=nsubs)
    lines = deck_lines()
    assert len(lines) == 164
    derived = {"LZ", "sigma_t_off", "cut"}
    for lab, val in lines:
        got = field_of(d, lab)
        if lab in derived:
            continue
        want = float(val) if ("." in val or "e" in val.lower()) else int(val)
        if nsubs == 1 and second_component(lab):
            assert got == 0, lab
        else:
            assert got == want, (lab, got, want)
    # derived values (src/params.h:380-650)
    assert d.LZ == 1 and d.NumNodes == 1088 * 1224
    assert d.NumTimeSteps == 8 * 1512
    assert d.cut == 0.3  # -1. in the deck
    assert (d.x0, d.y0, d.z0, d.r0) == (17, 17, 77, 10)
    if nsubs == 2:
        assert d.sigma_t_on == 657000 and d.sigma_t_off == d.NumTimeSteps
        assert d.rho_A[1] == d.rho_B[0] and d.rho_B[1] == d.rho_A[0]
        assert d.zeroconcgrad_n == 1 and d.zeroconcgrad_s == 1
    else:
        assert d.rho_A[1] == 0.0


def test_deck_params_mapping(io):
    d = io.read_deck(os.path.join(GOLD, "params.in"), dims=2, nsubs=2)
    p = io.deck_params(d, dims=2, nsubs=2)
    assert (p.dims, p.ni, p.nj, p.nk, p.nsubs) == (2, 1088, 1224, 1, 2)
    assert tuple(p.tau) == (0.6, 1.0)
    assert p.sink == 1e-5 and p.pfmul == 0.1 and p.pfadd == 1.0
    assert p.rho_in == 1.00001 and p.rho_out == 0.99999
    assert tuple(p.pressure_n_in) == (2, 0) and tuple(p.pressure_s_in) == (2, 0)
    assert (p.bc_ramp_start, p.bc_ramp_stop) == (0, 200000)
    assert p.flux_plane == -1 and p.storage == 0 and p.compute == 0


def labelled_deck(tmp_path, edits, name="params.in"):
    """the reference deck with the values of some labels replaced"""
    lines = []
    for lab, val in deck_lines():
        lines.append(f"{lab} {edits.pop(lab, val)}")
    assert not edits, edits
    path = tmp_path / name
    path.write_text("\n".join(lines) + "\n")
    return str(path)


# the reference deck's own 2-D boundary flags, cleared for 3-D decks
NO_2D = {"pressure_n_in[0]": 0, "pressure_s_in[0]": 0, "zeroconcgrad_n": 0, "zeroconcgrad_s": 0}


def test_deck_params_3d_faces(io, tmp_path):
    """3-D decks: the t/b pressure and velocity flags become z-face
    conditions (t = z = LZ-1, b = z = 0) with rho_in/rho_out and
    uz_in/uz_out; constcon_{t,b} a Dirichlet face with rho_sigma_in/out,
    zeroconcgrad_{t,b} a zero-gradient face"""
    edits = dict(NO_2D, **{"pressure_b_out[0]": 1, "velocity_t_in[0]": 1, "constcon_t_in": 1,
                           "zeroconcgrad_b": 1, "rho_sigma_in": "0.75"})
    d = io.read_deck(labelled_deck(tmp_path, edits), dims=3, nsubs=2)
    p = io.deck_params(d, dims=3, nsubs=2)
    assert tuple(p.zface_flow) == (2, 3)  # bottom pressure, top velocity
    assert p.rho_bot == 0.99999 and tuple(p.zface_uz) == (0.0, -0.01)
    assert tuple(p.zface_conc) == (3, 2) and p.c_top == 0.75
    assert p.zface_pressure == 0 and p.zface_solute == 0
    assert tuple(p.pressure_n_in) == (0, 0) and p.zeroconcgrad_n == 0
    # the other combination: velocity out at the bottom, pressure in at the top
    edits = dict(NO_2D, **{"velocity_b_out[0]": 1, "pressure_t_in[0]": 1, "constcon_b_in": 1})
    d = io.read_deck(labelled_deck(tmp_path, edits), dims=3, nsubs=2)
    p = io.deck_params(d, dims=3, nsubs=2)
    assert tuple(p.zface_flow) == (3, 2) and p.zface_uz[0] == -0.01 and p.rho_top == 1.00001
    assert tuple(p.zface_conc) == (2, 1) and p.c_bot == 0.0  # rho_sigma_in is 0 in the deck
    # no flags: both faces explicitly without a condition (periodic)
    d = io.read_deck(labelled_deck(tmp_path, dict(NO_2D)), dims=3, nsubs=1)
    p = io.deck_params(d, dims=3, nsubs=1)
    assert tuple(p.zface_flow) == (1, 1) and tuple(p.zface_conc) == (1, 1)


def test_deck_params_2d_velocity_rows(io, tmp_path):
    edits = {"pressure_s_in[0]": 0, "velocity_s_in[0]": 1, "uy_in": "0.002"}
    d = io.read_deck(labelled_deck(tmp_path, edits), dims=2, nsubs=2)
    p = io.deck_params(d, dims=2, nsubs=2)
    assert p.velocity_s_in == 1 and p.uy_in == 0.002 and tuple(p.pressure_n_in) == (2, 0)
    assert tuple(p.pressure_s_in) == (0, 0)


@pytest.mark.parametrize("dims,edits,match", [
    (3, {}, r"pressure_n_in\[0\] = 2 is set"),  # the reference deck's 2-D rows in 3-D
    (2, {"zeroconcgrad_full": 1}, "zeroconcgrad_full = 1 is set"),
    (2, {"pressure_e_in[0]": 1}, r"pressure_e_in\[0\] = 1 is set"),
    (2, {"velocity_w_out[1]": 1}, r"velocity_w_out\[1\] = 1 is set"),
    (2, {"velocity_n_in[1]": 1}, r"velocity_n_in\[1\] = 1 is set"),
    (2, {"pressure_t_in[0]": 1}, r"pressure_t_in\[0\] = 1 is set"),
    (2, {"constflx_s_in": 1}, "constflx_s_in = 1 is set"),
    (2, {"GZL": 1}, "GZL = 1 is set"),
    (2, {"bc_slip_north": 1}, "bc_slip_north = 1 is set"),
    (3, dict(NO_2D, **{"pressure_t_in[1]": 1}), r"pressure_t_in\[1\] = 1 is set"),
    (3, dict(NO_2D, **{"zeroconcgrad_s": 0, "constcon_n_in": 1}), "constcon_n_in = 1 is set"),
    (3, dict(NO_2D, **{"pressure_t_in[0]": 1, "velocity_t_out[0]": 1}), "top .t. face has more than one flow"),
    (3, dict(NO_2D, **{"constcon_b_in": 1, "zeroconcgrad_b": 1}), "bottom .b. face has more than one solute"),
    (2, {"velocity_n_in[0]": 1}, "north row has both a pressure and a velocity condition"),
])
def test_deck_params_rejects_flags_without_condition(io, tmp_path, dims, edits, match):
    """a set boundary flag this build cannot apply is an error naming it,
    not a silent no-op as in the reference's GPU build"""
    d = io.read_deck(labelled_deck(tmp_path, dict(edits)), dims=dims, nsubs=2)
    with pytest.raises(Exception, match=match):
        io.deck_params(d, dims=dims, nsubs=2)


def edited_deck(tmp_path, edits):
    lines = [f"{lab} {val}" for lab, val in deck_lines()]
    for k, v in edits.items():
        lines[k] = v
    path = tmp_path / "params.in"
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def test_deck_integer_written_as_real_shifts_entries(io, tmp_path):
    """An integer entry written "1512.5" leaves ".5" in the stream; the next
    entry has no label then (skip_label sees '.') and scans ".5", and every
    later entry takes its predecessor's value (src/lattice.h:203-213)"""
    path = edited_deck(tmp_path, {8: "FrameRate 1512.5"})
    d = io.read_deck(path, dims=2, nsubs=1)
    assert d.FrameRate == 1512
    assert d.tau[0] == 0.5
    assert d.gforce[0][0] == 0.6  # the tau[0] value
    assert d.end_grav[0] == 0  # a real's text scanned with %d stops at '.'


def test_deck_ends_early_is_an_error(io, tmp_path):
    path = tmp_path / "params.in"
    path.write_text("\n".join(f"{a} {b}" for a, b in deck_lines()[:40]) + "\n")
    with pytest.raises(Exception, match="end of file"):
        io.read_deck(str(path))


def test_deck_defaults_for_negative_bounds(io, tmp_path):
    edits = {143: "x0 -1", 144: "y0 -1", 145: "z0 -1", 146: "r0 -1",
             148: "x1 -1", 149: "x2 -1", 150: "y1 -1", 151: "y2 -1", 152: "z1 -1", 153: "z2 -1",
             154: "rel_x1 -1.", 156: "rel_y1 0.25", 157: "rel_y2 0.8"}
    d = io.read_deck(edited_deck(tmp_path, edits), dims=3, nsubs=1)
    LX, LY, LZ = 1088, 1224, 1
    assert (d.x0, d.y0, d.z0) == (LX // 2, LY // 2, LZ // 2)
    assert d.r0 == min(LX, LY, LZ)
    assert d.x1 == LX // 4  # rel_x1 < 0
    assert d.x2 == round(0.6 * LX)  # rel_x2 = 0.6
    assert d.y1 == round(0.25 * LY) and d.y2 == round(0.8 * LY)
    assert d.z1 == round(0.4 * LZ) and d.z2 == round(0.6 * LZ)


def decode_bmp24(path):
    """independent BMP decode: bottom-up rows in file order, 4-byte row
    padding, black = solid"""
    raw = open(path, "rb").read()
    w, h = struct.unpack_from("<ii", raw, 18)
    bits = struct.unpack_from("<H", raw, 28)[0]
    off = struct.unpack_from("<I", raw, 10)[0]
    assert bits == 24
    row = (3 * w + 3) // 4 * 4
    px = np.frombuffer(raw, np.uint8, count=row * h, offset=off).reshape(h, row)[:, :3 * w]
    px = px.reshape(h, w, 3)
    return (px == 0).all(axis=2).astype(np.uint8)


@pytest.mark.parametrize("name,L", [("3x3.bmp", 3), ("9x9.bmp", 9)])
def test_reference_bmp_solids(io, name, L):
    m = io.read_solids_bmp(os.path.join(GOLD, name), L, L)
    want = decode_bmp24(os.path.join(GOLD, name))
    assert np.array_equal(m, want)
    if name == "9x9.bmp":
        assert 0 < m.sum() < 81  # the image holds a black diagonal band


def test_bmp_size_mismatch_is_an_error(io):
    with pytest.raises(Exception, match="width"):
        io.read_solids_bmp(os.path.join(GOLD, "9x9.bmp"), 8, 9)
    with pytest.raises(Exception, match="height"):
        io.read_solids_bmp(os.path.join(GOLD, "9x9.bmp"), 9, 8)


def write_bmp24(path, mask):
    """a 24-bit BMP whose black pixels are mask == 1, row j = 0 first"""
    h, w = mask.shape
    row = (3 * w + 3) // 4 * 4
    data = bytearray()
    for j in range(h):
        r = bytearray()
        for i in range(w):
            r += b"\x00\x00\x00" if mask[j, i] else b"\xff\xff\xff"
        data += r + b"\x00" * (row - 3 * w)
    hdr = struct.pack("<2sIHHI", b"BM", 54 + len(data), 0, 0, 54)
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, len(data), 2835, 2835, 0, 0)
    with open(path, "wb") as f:
        f.write(hdr + info + data)


@pytest.mark.parametrize("shape", [(5, 7), (6, 13), (4, 4)])
def test_bmp_round_trip_every_padding(io, tmp_path, shape):
    rng = np.random.default_rng(sum(shape))
    mask = (rng.random(shape) < 0.4).astype(np.uint8)
    p = str(tmp_path / "s.bmp")
    write_bmp24(p, mask)
    assert np.array_equal(io.read_solids_bmp(p, shape[1], shape[0]), mask)


@pytest.mark.parametrize("name,L", [("3x3x3.raw", 3), ("9x9x9.raw", 9)])
def test_reference_raw_solids(io, name, L):
    m = io.read_solids_raw(os.path.join(GOLD, name), L, L, L)
    want = np.fromfile(os.path.join(GOLD, name), np.uint8).reshape(L, L, L)
    assert np.array_equal(m, want)
    if name == "9x9x9.raw":
        # the file marks its obstacles with 255, which the GPU kernels of the
        # reference (and these) treat as fluid: only a byte of 1 is solid
        assert set(np.unique(m)) == {0, 255}


def test_raw_too_short_is_an_error(io):
    with pytest.raises(Exception, match="read"):
        io.read_solids_raw(os.path.join(GOLD, "3x3x3.raw"), 3, 3, 4)


def test_series_counting(io, tmp_path):
    a = tmp_path / "a.in"
    a.write_text("1.0 2.0\n3.5\n")
    assert np.array_equal(io.read_series(str(a)), [1.0, 2.0, 3.5])
    b = tmp_path / "b.in"
    b.write_text("1.0 2.0 3.5")  # no newline: the last value is not counted
    assert np.array_equal(io.read_series(str(b)), [1.0, 2.0])


def test_gen_filename(io):
    f = io.filename
    assert f("./out", 2, 1088, 1224, 1, "rho", 3, 0, ".txt") == "./out/cuda_rho1088x1224_frame0003.txt"
    assert f("./out", 2, 64, 32, 2, "frames", -1, 1, ".dat") == "./out/cuda_frames64x32_subs01.dat"
    assert f("./out", 3, 64, 32, 1, "u_x", 12, 0, ".raw") == "./out/cuda_u_x64x32x32_frame0012.raw"
    assert f("o", 3, 8, 9, 2, "rho", 0, 1, ".dat", num_procs=4, proc_id=2) == \
        "o/cuda_rho8x9x9_frame0000_subs01_proc0002.dat"


def test_frame_stats_against_numpy(io):
    rng = np.random.default_rng(4)
    N = 1000
    rho = rng.uniform(0.9, 1.1, N)
    u = [rng.uniform(-0.1, 0.1, N) for _ in range(3)]
    mask = (rng.random(N) < 0.3).astype(np.uint8)
    mask[mask == 1] = rng.choice([1, 255], int(mask.sum())).astype(np.uint8)
    st = io.frame_stats(rho, *u, mask)
    f = mask == 0
    def seq_mean(a):  # node order, one rounding per addition, as compute.c sums
        s = 0.0
        for v in a:
            s += float(v)
        return s / a.size

    assert st[0] == np.abs(rho[f]).max() and st[1] == np.abs(rho[f]).min()
    assert st[2] == seq_mean(rho[f])
    for c in range(3):
        assert st[3 + c] == np.abs(u[c][f]).max()
        assert st[6 + c] == np.abs(u[c][f]).min()
        assert st[9 + c] == seq_mean(u[c][f])
    st2 = io.frame_stats(rho, u[0], u[1], None, mask)
    assert st2[5] == 0 and st2[8] == 0 and st2[11] == 0


def test_frames_dat_rows(io, tmp_path):
    p = str(tmp_path / "frames.dat")
    st0 = np.array([1.0, 1.0, 1.0] + [0.0] * 9)
    # max|rho|, min|rho|, ave rho, max|u| (3), min|u| (3), ave u (3)
    st1 = np.array([1.002, 0.998, 1.0001, 0.01, 0.02, 0.0, 1234.5, 0.001, 0.0, 0.005, -0.004, 0.0])
    io.frames_dat(p, 0, st0)
    io.frames_dat(p, 20, st1, flux=(0.1, 0.2, 0.25, 0.3))
    lines = open(p).read().split("\n")
    assert lines[0] == ""
    assert lines[1].split() == ["time", "|j|", "j_x", "j_y", "j_z", "min_u[x]", "min_u[y]", "min_u[z]",
                                "max_u[x]", "max_u[y]", "max_u[z]", "ave_u[x]", "ave_u[y]", "ave_u[z]",
                                "max/ave[x]", "max/ave[y]", "max/ave[z]", "min_rho", "max_rho", "ave_rho",
                                "max/ave"]
    assert lines[1].startswith("        time          |j| ")
    assert lines[2] == "------------ " * 21
    r0 = lines[3].split()
    assert r0[0] == "0" and float(r0[-1]) == 1.0 and r0[14:17] == ["1.0000000"] * 3
    row = lines[4]
    # 12-wide fields with one blank between, except that j_z runs straight
    # into min_u[x]
    assert row[:13] == "          20 "
    assert row[52:76] == "   0.30000001234.5000000"
    vals = [float(v) for v in row[:64].split() + [row[64:76]] + row[76:].split()]
    want = [20, 0.1, 0.2, 0.25, 0.3, 1234.5, 0.001, 0.0, 0.01, 0.02, 0.0, 0.005, -0.004, 0.0,
            0.01 / 0.005, 0.02 / -0.004, 1.0, 0.998, 1.002, 1.0001, 1.002 / 1.0001]
    assert np.allclose(vals, want, rtol=0, atol=5e-8)
    st2 = st1.copy()
    st2[6] = 0.0
    st2[9] = 1e-9  # a ratio beyond 9999 prints inf
    io.frames_dat(p, 22, st2)
    assert open(p).read().split("\n")[5].split()[14] == "inf"


def test_txt_dat_raw_writers(io, tmp_path):
    LX, LY, LZ = 4, 3, 2
    rng = np.random.default_rng(9)
    a = rng.uniform(-1, 1, LX * LY * LZ)
    mask = np.zeros(a.size, np.uint8)
    mask[[1, 7, 20]] = 1
    mask[5] = 255
    t = str(tmp_path / "a.txt")
    io.write_txt(t, a, LX, LY, LZ)
    text = open(t).read()
    planes = text.split("\n\n")
    assert len(planes) == LZ + 1 and planes[-1] == ""
    rows = [r for p in planes[:-1] for r in p.split("\n")]
    assert len(rows) == LY * LZ and all(r.startswith(" ") for r in rows)
    got = np.array([float(v) for r in rows for v in r.split()])
    assert np.max(np.abs(got - a)) <= 1e-16
    d = str(tmp_path / "a.dat")
    io.write_dat(d, a, mask)
    back = np.fromfile(d, np.float64)
    assert np.array_equal(back[mask == 0], a[mask == 0]) and not back[mask != 0].any()
    r = str(tmp_path / "a.raw")
    lo, hi = 0.2, 0.9
    io.write_raw8(r, a, mask, [hi], [lo])
    b = np.fromfile(r, np.uint8)
    x = np.floor(1 + 254 * (a - lo) / (hi - lo)).astype(np.int64) & 0xFF  # values below lo wrap
    assert np.array_equal(b[mask == 0], x[mask == 0]) and not b[mask != 0].any()
    # speed image of write_raw_u
    ux, uy, uz = (rng.uniform(-0.1, 0.1, a.size) for _ in range(3))
    io.write_raw8(r, ux, mask, [0.1, 0.1, 0.1], [0.0, 0.0, 0.0], uy=uy, uz=uz)
    s = np.sqrt(ux * ux + uy * uy + uz * uz)
    x = np.floor(1 + 254 * s / np.sqrt(0.03)).astype(np.int64) & 0xFF
    assert np.array_equal(np.fromfile(r, np.uint8)[mask == 0], x[mask == 0])
    # equal extremes divide by zero: the byte of 0x80000000
    io.write_raw8(r, a, mask, [0.5], [0.5])
    assert not np.fromfile(r, np.uint8).any()


def test_deck_main_argument_errors(io, tmp_path):
    assert io.deck_main("--dims", "4") == 1
    assert io.deck_main("--in", str(tmp_path / "missing")) == 1
    assert io.deck_main("--help") == 0
    # the reference's Walsh build sets ns from the 2-D image only
    assert io.deck_main("--dims", "3", "--walsh", "--in", str(tmp_path)) == 1


def test_observations_read_filter_and_order(io, tmp_path):
    t, x, y = tmp_path / "t.in", tmp_path / "x.in", tmp_path / "y.in"
    t.write_text("5\n10\n10\n20\n31\n")
    x.write_text("1 100 3 4 9\n")
    y.write_text("2 2 5 -1 7\n")
    ot, ox, oy, order = io.read_observations(t, x, y, 10, 8)
    # index 1 (x = 100) and index 3 (y = -1) lie outside the 10 x 8 lattice
    assert list(ot) == [5, 10, 31] and list(ox) == [1, 3, 9] and list(oy) == [2, 5, 7]
    assert list(order) == [0, 2, 4]
    t.write_text("5 10 20")  # no newline after the last time: it is not counted
    ot, ox, oy, order = io.read_observations(t, x, y, 10, 8)
    assert list(ot) == [5] and list(order) == [0]  # the second point is outside
    t.write_text("5\n10\n7\n")
    with pytest.raises(Exception, match="time order"):
        io.read_observations(t, x, y, 10, 8)


def test_observations_written(io, tmp_path):
    conc = np.array([0.25, 1.0 / 3.0, 0.0])
    vel = np.array([1e-5, 2.5e-3, 0.0])
    io.write_observations(tmp_path, conc, vel, [0, 2, 4])
    c = (tmp_path / "conc_data_proc0000.dat").read_text().split("\n")
    v = (tmp_path / "vel_data_proc0000.dat").read_text().split("\n")
    o = (tmp_path / "conc_order_proc0000.dat").read_text().split("\n")
    assert c[:3] == [" 0.25000000000000000", " 0.33333333333333331", " 0.00000000000000000"]
    assert [float(q) for q in v[:3]] == [1e-5, 2.5e-3, 0.0] and len(v[0]) == 20
    assert o == ["0", "2", "4", ""]

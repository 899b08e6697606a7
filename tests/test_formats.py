"""The configuration formats (lpmd2_amd/csrc/formats.cpp through
lpmd2_amd.snapshots) against the oracle's restatement (oracle/formats.py),
and the oracle against the reference's own files.

Fixtures (tests/golden/formats/, copied from lpmd/examples by
tests/golden/gen_format_fixtures.py): argon108.lpmd (LPMD 1.0), proyectil.lpmd
and zirconia-color.lpmd (LPMD 2.0 text), ZrO2-3conf.zlp (the first three
frames of ZrO2-100conf.zlp, a legacy ZLP 1 stream).  The reference's format
plugins derive from CellReader, whose runtime-context dependency cannot be
built here, so no reference-written output of the current writer exists: the
writer is pinned by the header line the reference's own 2.0 files carry, by
round trips through the reader the fixtures pin, and by zlib itself.

These run on the host only (no GPU)."""
import os
import zlib

import numpy as np
import pytest

from oracle import formats as F

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "formats")


def S():
    import lpmd2_amd.snapshots as snapshots

    try:
        snapshots.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    return snapshots


def fixture(name):
    return os.path.join(GOLD, name)


# ---------------------------------------------------------------- oracle pins
def test_legacy_zlp_frames_decode_to_the_zirconia_structure():
    """Each legacy ZLP 1 frame is a width byte (4: written by a 32-bit
    build), a network-order size and deflate data continuing one stream; the
    frames decode to 324 atoms, 108 Zr and 216 O, in the cell of the
    independent text file zirconia-color.lpmd of the same structure (15.58,
    15.99, 16.15), every scaled position inside [0, 1]."""
    raw = open(fixture("ZrO2-3conf.zlp"), "rb").read()
    assert raw[:5] == b"Z4L2P" and raw[5] == 1 and raw[10] == 4
    frames = F.read_lpmd(fixture("ZrO2-3conf.zlp"))
    text = F.read_lpmd(fixture("zirconia-color.lpmd"))[0]
    assert len(frames) == 3
    for fr in frames:
        assert fr["z"].shape == (324,) and fr["level"] == 0
        assert np.bincount(fr["z"], minlength=41)[[8, 40]].tolist() == [216, 108]
        np.testing.assert_allclose(np.diag(fr["cell"]), np.diag(text["cell"]), rtol=1e-14)
        assert fr["s"].min() >= 0.0 and fr["s"].max() <= 1.0
    # consecutive frames of one trajectory: the same atoms, moved a little
    d = frames[1]["s"] - frames[0]["s"]
    d -= np.round(d)
    # (a column read out of place would move atoms by a quarter cell on average)
    assert 0.0 < np.abs(d).mean() < 0.02 and np.abs(d).max() < 0.1
    assert np.array_equal(frames[0]["z"], frames[1]["z"])


def test_lpmd_one_argon_is_an_fcc_lattice():
    """argon108.lpmd (LPMD 1.0: no HDR line, columns from the third line's
    word count): 108 Ar atoms whose scaled positions are the 3 x 3 x 3 fcc
    lattice of the 17.1191 A cell."""
    fr = F.read_lpmd(fixture("argon108.lpmd"))
    assert len(fr) == 1
    fr = fr[0]
    assert fr["level"] == 0 and np.all(fr["z"] == 18)
    np.testing.assert_array_equal(np.diag(fr["cell"]), [17.1191] * 3)
    u = (fr["s"] * 3.0) % 1.0
    u = np.minimum(u, 1.0 - u)
    # every coordinate on the lattice or half a lattice step off it
    assert np.all(np.minimum(np.abs(u), np.abs(u - 0.5)) < 1e-6)
    # 108 distinct sites on the half-lattice grid (6 points an edge)
    assert len({tuple(np.round(r).astype(int) % 6) for r in fr["s"] * 6.0}) == 108


def test_writer_header_is_the_reference_files_header():
    """The LPMD 2.0 files the reference ships carry "HDR SYM X Y Z VX VY VZ "
    (each column followed by a space); the writer's header line at level 1
    is the same line."""
    line2 = open(fixture("proyectil.lpmd")).read().split("\n")[1]
    assert F.lpmd_header(False, 1).split("\n")[1] == line2
    assert F.lpmd_header(True, 1).startswith("LPMD 2.0 Z\n")


def test_zlp_writer_stream_decodes_with_zlib():
    """Frames of the oracle's zlp writer, decoded by zlib on their own
    framing, are the text frames of the plain writer."""
    rng = np.random.default_rng(4)
    cell = np.diag([10.0, 11.0, 12.0])
    z = rng.choice([18, 29], 50)
    x = rng.uniform(0, 1, (50, 3)) * [10.0, 11.0, 12.0]
    w = F.LpmdWriter(True, 0, blocksize=64)
    w.frame(cell, z, x)
    w.frame(cell, z, x + 0.5)
    b = w.bytes()
    head = F.lpmd_header(True, 0).encode()
    assert b.startswith(head)
    pos, dec, texts = len(head), zlib.decompressobj(), []
    while pos < len(b):
        assert b[pos] == 8 and b[pos + 5:pos + 9] == b"\0\0\0\0"
        size = int.from_bytes(b[pos + 1:pos + 5], "big")
        texts.append(dec.decompress(b[pos + 9:pos + 9 + size]).decode())
        pos += 9 + size
    assert texts == [F.lpmd_frame_text(cell, z, x), F.lpmd_frame_text(cell, z, x + 0.5)]


# ---------------------------------------------------------------- engine vs oracle
@pytest.mark.parametrize("name", ["argon108.lpmd", "proyectil.lpmd", "zirconia-color.lpmd", "ZrO2-3conf.zlp"])
def test_reader_matches_oracle_on_reference_files(name):
    s = S()
    got = s.read_lpmd(fixture(name))
    exp = F.read_lpmd(fixture(name))
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        for k in ("cell", "z", "s", "v", "a"):
            assert np.array_equal(g[k], e[k]), k
        assert g["level"] == e["level"]


def state(n=300, seed=1):
    rng = np.random.default_rng(seed)
    L = np.array([23.4, 19.7, 31.1])
    z = rng.choice([8, 18, 29, 40], n).astype(np.int32)
    return (np.diag(L), z, rng.uniform(0, 1, (n, 3)) * L, rng.normal(0, 0.01, (n, 3)), rng.normal(0, 0.1, (n, 3)))


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("ext", ["lpmd", "zlp"])
def test_writer_bytes_match_oracle_and_read_back(tmp_path, ext, level):
    """The engine's writer gives the oracle's bytes (zlp included: the same
    deflate calls on the same zlib), and its reader gets the state back to the
    six decimals the format keeps."""
    s = S()
    cell, z, x, v, a = state()
    p = tmp_path / f"t.{ext}"
    with s.LpmdWriter(p, f"level {level} blocksize 512") as w:
        for k in range(3):
            w.write(cell, z, x + 0.01 * k, v, a)
    ow = F.LpmdWriter(ext == "zlp", level, blocksize=512)
    for k in range(3):
        ow.frame(cell, z, x + 0.01 * k, v, a)
    assert open(p, "rb").read() == ow.bytes()
    back = s.read_lpmd(p)
    assert len(back) == 3
    for k, fr in enumerate(back):
        np.testing.assert_array_equal(fr["z"], z)
        assert np.abs(s.cartesian(fr["s"], fr["cell"]) - (x + 0.01 * k)).max() <= 5e-7 * 31.1 + 1e-12
        if level >= 1:
            assert np.abs(fr["v"] - v).max() <= 5e-7
        if level >= 2:
            assert np.abs(fr["a"] - a).max() <= 5e-7


def test_reference_file_round_trip(tmp_path):
    """proyectil.lpmd read, written at level 1 and read again: the same atoms,
    cell and velocities, positions within the written precision."""
    s = S()
    fr = s.read_lpmd(fixture("proyectil.lpmd"))[0]
    x = s.cartesian(fr["s"], fr["cell"])
    p = tmp_path / "p.lpmd"
    with s.LpmdWriter(p, "level 1") as w:
        w.write(fr["cell"], fr["z"], x, fr["v"])
    assert open(p).read().split("\n")[1] == open(fixture("proyectil.lpmd")).read().split("\n")[1]
    g = s.read_lpmd(p)[0]
    np.testing.assert_array_equal(g["z"], fr["z"])
    np.testing.assert_array_equal(g["cell"], fr["cell"])
    assert np.abs(g["s"] - fr["s"]).max() <= 5e-7
    assert np.abs(g["v"] - fr["v"]).max() <= 5e-7


@pytest.mark.parametrize("args,level", [("", 0), ("level 1 precout 3", 1), ("level 2 precout 5 inside true", 2)])
def test_xyz_matches_oracle(tmp_path, args, level):
    s = S()
    cell, z, x, v, a = state(seed=3)
    x = x - 6.0  # some atoms outside the cell
    p = tmp_path / "t.xyz"
    s.write_xyz(p, cell, z, x, v, a, args)
    s.write_xyz(p, cell, z, x + 1.0, v, a, args, append=True)
    kw = dict(zip(args.split()[0::2], args.split()[1::2]))
    prec = int(kw.get("precout", 8))
    inside = kw.get("inside") == "true"
    exp = (F.xyz_text(cell, z, x, v, a, level, prec, inside) + F.xyz_text(cell, z, x + 1.0, v, a, level, prec, inside))
    assert open(p).read() == exp


def test_dump_matches_oracle_and_restores(tmp_path):
    """Dump's text (six significant digits), then Restore: the step, cell,
    atoms and fixed-atom tags come back; values to the written precision."""
    s = S()
    cell, z, x, v, a = state(seed=5)
    n = z.shape[0]
    fl = np.zeros(n, np.int32)
    fl[3] = 1 | 4
    fl[7] = 2 | 4
    fl[9] = 4
    fl[11] = 1 | 2 | 4
    p = tmp_path / "last.dump"
    s.write_dump(p, 1234, cell, z, x, v, a, fl)
    tags = [[] for _ in range(n)]
    tags[3] = [("fixedvel", "true")]
    tags[7] = [("fixedpos", "true")]
    tags[9] = [("fixedpos", "false")]
    tags[11] = [("fixedvel", "true"), ("fixedpos", "true")]
    assert open(p).read() == F.dump_text(1234, cell, z, x, v, a, tags)
    r = s.read_dump(p)
    assert r["step"] == 1234
    np.testing.assert_array_equal(r["z"], z)
    np.testing.assert_array_equal(r["flags"], fl)
    np.testing.assert_array_equal(r["cell"], cell)
    for k, ref in (("x", x), ("v", v), ("a", a)):
        assert np.all(np.abs(r[k] - ref) <= 5e-6 * np.abs(ref) + 1e-300)


def test_format_errors(tmp_path):
    s = S()
    bad = tmp_path / "bad.lpmd"
    bad.write_text("XYZ nonsense\n")
    with pytest.raises(s.LpmdxError):
        s.read_lpmd(bad)
    v3 = tmp_path / "v3.lpmd"
    v3.write_text("LPMD 3.0\nHDR SYM X Y Z\n")
    with pytest.raises(s.LpmdxError, match="not supported"):
        s.read_lpmd(v3)
    cell, z, x, v, a = state(n=4)
    with pytest.raises(s.LpmdxError, match="orthogonal"):
        with s.LpmdWriter(tmp_path / "o.lpmd") as w:
            w.write(cell + np.eye(3)[::-1] * 0.5, z, x)
    with pytest.raises(s.LpmdxError, match="colours"):
        s.LpmdWriter(tmp_path / "c.lpmd", "extra RGB")
    # an unknown symbol reads as Z = 0 (ElemNum), six-number cells as
    # NonOrthogonalCell(a, b, c, alpha, beta, gamma)
    odd = tmp_path / "odd.lpmd"
    odd.write_text("LPMD 2.0\nHDR SYM X Y Z\n2\n10 11 12 90 90 90\nQq 0.1 0.2 0.3\nAr 0.5 0.5 0.5\n")
    fr = s.read_lpmd(odd)[0]
    assert fr["z"].tolist() == [0, 18]
    np.testing.assert_allclose(fr["cell"], F.cell_from_lengths(10, 11, 12, 90, 90, 90), rtol=0, atol=0)

"""Format converters (-m mtx-bel | tsv-bel | tsv-mtx | bel-mtx | txt-bel):
hand-built inputs and the bytes the reference's MtB_Writer writes for them
(FIleReader.cuh:459-912), read off its code."""
import os
import struct
import subprocess

import pytest

import beep_amd as B

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "beep_amd", "bin", "beep")


def bel(*recs):
    return b"".join(struct.pack("<3Q", *r) for r in recs)


def run(tmp_path, kind, text, binary=False):
    src = tmp_path / "in"
    dst = tmp_path / "out"
    src.write_bytes(text if binary else text.encode())
    B.convert(kind, src, dst)
    return dst.read_bytes()


def test_mtx_bel_symmetric_pattern(tmp_path):
    out = run(tmp_path, "mtx-bel", "%%MatrixMarket matrix coordinate pattern symmetric\n5 5 3\n1 0\n2 2\n4 1\n")
    # (1,0),(0,1),(2,2),(4,1),(1,4) sorted by (src, dst); records {dst, src, 0}; ids as written
    assert out == bel((1, 0, 0), (0, 1, 0), (4, 1, 0), (2, 2, 0), (1, 4, 0))


def test_mtx_bel_later_comment_resets_flags(tmp_path):
    out = run(tmp_path, "mtx-bel", "%%MatrixMarket matrix coordinate pattern symmetric\n% note\n3 3 2\n0 1\n1 2\n")
    assert out == bel((1, 0, 0), (2, 1, 0))


def test_mtx_bel_one_char_comment_keeps_flags(tmp_path):
    out = run(tmp_path, "mtx-bel", "%%MatrixMarket matrix coordinate pattern symmetric\n%\n3 3 1\n0 1\n")
    assert out == bel((1, 0, 0), (0, 1, 0))


def test_mtx_bel_stops_at_malformed_line_and_keeps_the_rest_read(tmp_path):
    out = run(tmp_path, "mtx-bel", "%%MatrixMarket matrix coordinate real general\n3 3 3\n2 0 1.5\nx y\n0 1\n")
    assert out == bel((0, 2, 0))


def test_mtx_bel_errors(tmp_path):
    with pytest.raises(B.BeepError):
        run(tmp_path, "mtx-bel", "%%MatrixMarket matrix coordinate general\n3 4 1\n0 1\n")
    with pytest.raises(B.BeepError):
        run(tmp_path, "mtx-bel", "%%MatrixMarket matrix coordinate general\nnot a size line\n")


def test_tsv_bel_until_first_bad_line(tmp_path):
    out = run(tmp_path, "tsv-bel", "3 1 7\n4 2 0\n5 x 1\n9 9 9\n")
    assert out == bel((3, 1, 0), (4, 2, 0))


def test_tsv_bel_keeps_64_bit_ids(tmp_path):
    out = run(tmp_path, "tsv-bel", "4294967297 5 0\n")
    assert out == bel((4294967297, 5, 0))


def test_tsv_mtx(tmp_path):
    out = run(tmp_path, "tsv-mtx", "3 1 7\n4 2 5\n0 1 4294967295\n")
    # edge (f2, f1), weight as int, n = first id of the last edge
    assert out.decode() == "%%MatrixMarket matrix coordinate general\n1 1 3\n1 3 7\n2 4 5\n1 0 -1\n"


def test_bel_mtx(tmp_path):
    out = run(tmp_path, "bel-mtx", bel((5, 1, 9), (6, 2, 3)), binary=True)
    assert out.decode() == "%%MatrixMarket matrix coordinate general\n2 2 2\n1 5 9\n2 6 3\n"


def test_empty_inputs_to_mtx_are_refused(tmp_path):
    with pytest.raises(B.BeepError):
        run(tmp_path, "bel-mtx", b"", binary=True)
    with pytest.raises(B.BeepError):
        run(tmp_path, "tsv-mtx", "")


def test_txt_bel_dedups_and_symmetrises(tmp_path):
    out = run(tmp_path, "txt-bel", "# SNAP header\n% pct\n\n1 2\n2 1\n3 3\n0 5 9\n")
    assert out == bel((5, 0, 0), (2, 1, 0), (1, 2, 0), (0, 5, 0))


def test_txt_bel_round_trips_through_the_loader(tmp_path):
    out_path = tmp_path / "g.bel"
    src = tmp_path / "g.txt"
    src.write_text("0 1\n1 2\n2 0\n2 3\n")
    B.convert("txt-bel", src, out_path)
    g = B.Graph.load(out_path)
    rp, col = g.csr()
    assert rp.tolist() == [0, 2, 4, 7, 8]
    assert col.tolist() == [1, 2, 0, 2, 0, 1, 3, 2]


def test_cli_conversion(tmp_path):
    src = tmp_path / "g.txt"
    dst = tmp_path / "g.bel"
    src.write_text("1 2\n")
    r = subprocess.run([CLI, "-m", "txt-bel", "-g", str(src), "-r", str(dst)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert dst.read_bytes() == bel((2, 1, 0), (1, 2, 0))


def test_cli_statistics_block(tmp_path):
    """-x on a 20-vertex star (centre 0): every figure worked out by hand from
    MatrixStats / PrintMtarixStruct (utils.cuh:110-183, 264-306) and the
    encoding counts of main.cu:140-170 (32-bit words: 19*50/32 = 29 for the
    centre plus 1 per leaf; the triangular layout needs 19 + 19)."""
    edges = [(0, i) for i in range(1, 20)]
    pairs = sorted(edges + [(b, a) for a, b in edges])
    g = B.Graph.from_edges([a for a, _ in pairs], [b for _, b in pairs])
    path = tmp_path / "star.bel"
    g.write_bel(path)
    r = subprocess.run([CLI, "-g", str(path), "-m", "kc", "-x", "1"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    i = lines.index("38,20,20,1,1.9,3.9,1,19,19,19,1")
    grid = lines[i + 1:i + 11]
    assert grid[0] == "52," * 10
    assert all(row == "52," + "-," * 9 for row in grid[1:])
    assert lines[i + 11] == "Now # of bytes we need to make this matrix binary encoded !!"
    assert lines[i + 12] == "n = 20, m = 38, elements = 48"
    assert lines[i + 13] == "n = 20, m = 38, elements = 38"

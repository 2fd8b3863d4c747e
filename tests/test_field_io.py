"""Field I/O (host side of libwhhd.so, no GPU): the reference's text and binary
dump formats (WaveletPotentials/Utils.cpp:812-990) written and read back, and the
PLY point cloud.  Expected text is spelled out for a small array, so the format
itself (row order, separators, notation) is pinned, not only the round trip."""
import numpy as np
import pytest

import waveletpotentials_amd as W


def small():
    # (nz, nx, ny) = (2, 2, 3): value at (x, y, z) = 100 z + 10 x + y
    z, x, y = np.meshgrid(np.arange(2), np.arange(2), np.arange(3), indexing="ij")
    return (100 * z + 10 * x + y).astype(np.float32)


def test_text_layout_and_round_trip(tmp_path):
    a = small()
    p = tmp_path / "a.txt"
    W.WriteDataTxt(p, a, digits=1)
    # one row of ny values per (x, z), rows in memory order (x fastest, then z)
    assert p.read_text() == ("0.0 1.0 2.0\n10.0 11.0 12.0\n"
                             "100.0 101.0 102.0\n110.0 111.0 112.0\n")
    rng = np.random.default_rng(1)
    b = rng.standard_normal((3, 4, 5)).astype(np.float32) * 1e-3
    W.WriteDataETxt(tmp_path / "b.txt", b)
    assert np.array_equal(W.ReadsDataTxt(tmp_path / "b.txt", (4, 5, 3)), b)
    with pytest.raises(W.WhhdError):
        W.ReadsDataTxt(tmp_path / "b.txt", (4, 5, 4))  # too few values


def test_sparse_text(tmp_path):
    a = small()
    a[0, 0, 0] = 5e-7                      # below the default threshold
    W.WriteDataTxt_sparse(tmp_path / "s.txt", a)
    lines = (tmp_path / "s.txt").read_text().splitlines()
    assert lines[0] == "0,1,0,1.00000e+00"
    assert lines[2] == "1,0,0,1.00000e+01"    # z outermost, then x, then y
    assert lines[-1] == "1,2,1,1.12000e+02"
    assert len(lines) == a.size - 1
    W.WriteVectorTxt_sparse(tmp_path / "v.txt", a, 0 * a, 0 * a + 3e-7, threshold=1e-6)
    vl = (tmp_path / "v.txt").read_text().splitlines()
    assert vl[0] == "0,1,0,1.00000e+00,0.00000e+00,3.00000e-07"
    assert len(vl) == a.size - 1


def test_binary_round_trip(tmp_path):
    a = np.random.default_rng(2).standard_normal((4, 3, 5)).astype(np.float32)
    W.WriteDataBinary(tmp_path / "a.f32", a)
    assert (tmp_path / "a.f32").stat().st_size == a.size * 4
    assert np.array_equal(W.ReadDataBinary(tmp_path / "a.f32", a.shape), a)
    assert np.array_equal(np.fromfile(tmp_path / "a.f32", np.float32).reshape(a.shape), a)


@pytest.mark.parametrize("binary", [True, False])
def test_ply_point_cloud(tmp_path, binary):
    a = small()
    h = (0.5, 0.25, 2.0)
    n = W.WritePly(tmp_path / "a.ply", a, h, threshold=50.0, binary=binary)
    assert n == 6                                   # the z = 1 cells
    raw = (tmp_path / "a.ply").read_bytes()
    head, body = raw.split(b"end_header\n")
    assert b"element vertex 6" in head
    assert (b"binary_little_endian" if binary else b"format ascii") in head
    pts = (np.frombuffer(body, "<f4") if binary else np.array(body.split(), float)).reshape(-1, 4)
    # first vertex: cell (x, y, z) = (0, 0, 1) at its centre, value 100
    assert np.allclose(pts[0], [0.25, 0.125, 3.0, 100.0])
    assert np.allclose(pts[:, 3], [100, 101, 102, 110, 111, 112])

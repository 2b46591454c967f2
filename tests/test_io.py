"""Host file formats: the reference's "%f" solution dump (writer and reader)
and the VTK writer.  The writer is pinned by the digest of the reference's own
dump (its main() built with PRI, 16^3, 8 iterations) of the same field."""
import hashlib
import json
import os

import numpy as np
import pytest

import fold3d

HERE = os.path.dirname(os.path.abspath(__file__))


def _py_dump(T, n):
    H = n + 2
    T3 = T.reshape(H, H, H)  # [k, j, i]
    return "".join("%f\n" % T3[k, j, i] for i in range(H) for j in range(H) for k in range(H))


def test_text_dump_matches_reference_main(golden, tmp_path):
    ref = json.load(open(os.path.join(HERE, "golden", "ref_dump_16.json")))
    T = golden["solve8_16"]
    path = tmp_path / "t.txt"
    fold3d.write_text(str(path), T, 16)
    data = path.read_bytes()
    assert data.count(b"\n") == ref["lines"] == 18 ** 3
    assert hashlib.md5(data).hexdigest() == ref["md5"]
    assert hashlib.sha256(data).hexdigest() == ref["sha256"]


def test_text_dump_order_and_format(golden, tmp_path):
    T = golden["solve8_24"]
    path = tmp_path / "t.txt"
    fold3d.write_text(str(path), T, 24)
    assert path.read_text() == _py_dump(T, 24)


def test_text_round_trip_and_box_grid(tmp_path):
    rng = np.random.default_rng(5)
    n3 = (5, 7, 3)
    T = rng.uniform(-20, 20, (n3[0] + 2) * (n3[1] + 2) * (n3[2] + 2))
    T[::7] = 1e16
    path = str(tmp_path / "b.txt")
    fold3d.write_text(path, T, n3)
    back = fold3d.read_text(path, n3)
    want = np.array([float("%f" % v) for v in T])
    assert np.array_equal(back, want)
    # order: i outermost, k innermost
    lines = open(path).read().split()
    H0, H1, H2 = (x + 2 for x in n3)
    assert float(lines[1]) == want[0 + H0 * (0 + H1 * 1)]
    assert float(lines[H2]) == want[0 + H0 * (1 + H1 * 0)]


def test_text_errors(tmp_path):
    T = np.zeros(27)
    with pytest.raises(fold3d.Fold3dError) as e:
        fold3d.write_text(str(tmp_path / "no" / "such" / "dir.txt"), T, 1)
    assert e.value.code == fold3d.FOLD3D_E_IO
    short = tmp_path / "short.txt"
    short.write_text("1.0\n2.0\n")
    with pytest.raises(fold3d.Fold3dError) as e:
        fold3d.read_text(str(short), 1)
    assert e.value.code == fold3d.FOLD3D_E_IO and "too few" in str(e.value)
    long_ = tmp_path / "long.txt"
    long_.write_text("0\n" * 28)
    with pytest.raises(fold3d.Fold3dError) as e:
        fold3d.read_text(str(long_), 1)
    assert "more values" in str(e.value)


def _parse_vtk(path):
    raw = open(path, "rb").read()
    head, lines = [], raw.split(b"\n")
    for ln in lines[:10]:
        head.append(ln.decode())
    body = b"\n".join(lines[10:])
    return head, body


@pytest.mark.parametrize("binary", [False, True])
def test_vtk_writer(tmp_path, binary):
    rng = np.random.default_rng(9)
    n3 = (4, 6, 5)
    h = (0.5, 0.25, 0.125)
    H0, H1, H2 = (x + 2 for x in n3)
    T = rng.normal(size=H0 * H1 * H2)
    path = str(tmp_path / "t.vtk")
    fold3d.write_vtk(path, T, n3, h, binary=binary)
    head, body = _parse_vtk(path)
    assert head[0] == "# vtk DataFile Version 3.0"
    assert head[2] == ("BINARY" if binary else "ASCII")
    assert head[3] == "DATASET STRUCTURED_POINTS"
    assert head[4] == "DIMENSIONS 4 6 5"
    assert head[5] == "ORIGIN 0 0 0"
    assert [float(x) for x in head[6].split()[1:]] == list(h)
    assert head[7] == "POINT_DATA 120"
    assert head[8] == "SCALARS T double 1" and head[9] == "LOOKUP_TABLE default"
    if binary:
        vals = np.frombuffer(body[:8 * 120], dtype=">f8")
    else:
        vals = np.array([float(x) for x in body.split()])
    T3 = T.reshape(H2, H1, H0)[1:-1, 1:-1, 1:-1]  # interior, x fastest
    assert np.array_equal(vals, T3.reshape(-1))

"""CI-vector restart files through the C ABI (context free, no GPU): the byte
layout of the reference's binary vector file (PsociVector.C:530-553, 627-643,
685-693: int64 nbasis, int64 nroots, int32 width, int32 prec, per root an
int64 root number and the coefficients, then the title), read-back with the
reference's completion of missing roots by unit vectors
(PsociGAvectors.C:539-659, 687-726), and the error paths."""
import struct

import numpy as np
import pytest

from psoci_amd import api


def reference_bytes(C, title=b"default", root_hi=0):
    """The file PsociGAvectors::dumpVectorScalable writes for C [n][k]: the
    root number is an int written as 8 bytes, so its upper half is whatever
    followed it in memory (root_hi stands for that)."""
    n, k = C.shape
    out = struct.pack("<qqii", n, k, 18, 14)
    for r in range(k):
        out += struct.pack("<q", (r + 1) | (root_hi << 32)) + C[:, r].astype("<f8").tobytes()
    return out + title


def test_writer_matches_reference_layout(tmp_path):
    C = np.random.default_rng(0).standard_normal((37, 3))
    p = tmp_path / "civec.dat"
    api.civec_write(str(p), C)
    assert p.read_bytes() == reference_bytes(C)
    api.civec_write(str(p), C, title="run 7")
    assert p.read_bytes() == reference_bytes(C, b"run 7")


def test_reader_takes_reference_files(tmp_path):
    C = np.random.default_rng(1).standard_normal((50, 4))
    p = tmp_path / "ref.dat"
    p.write_bytes(reference_bytes(C, b"PSOCI CI vectors", root_hi=0x7ffd1234))
    assert api.civec_info(str(p)) == (50, 4, "PSOCI CI vectors")
    got, nread = api.civec_read(str(p), 50, 4)
    assert nread == 4
    assert np.array_equal(got, C)


def test_roots_kept_and_completed(tmp_path):
    C = np.random.default_rng(2).standard_normal((20, 5))
    p = tmp_path / "v.dat"
    api.civec_write(str(p), C)
    # fewer sought than stored: the first k
    got, nread = api.civec_read(str(p), 20, 3)
    assert nread == 3 and np.array_equal(got, C[:, :3])
    # more sought than stored: unit vectors e_j complete columns j >= 5
    got, nread = api.civec_read(str(p), 20, 8)
    assert nread == 5 and np.array_equal(got[:, :5], C)
    assert np.array_equal(got[:, 5:], np.eye(20)[:, 5:8])
    # an empty file of vectors: all unit vectors (the reference's diagonal guess)
    api.civec_write(str(p), np.zeros((20, 0)))
    got, nread = api.civec_read(str(p), 20, 2)
    assert nread == 0 and np.array_equal(got, np.eye(20)[:, :2])


def test_errors(tmp_path):
    C = np.ones((10, 2))
    p = tmp_path / "v.dat"
    api.civec_write(str(p), C)
    with pytest.raises(api.PsociError, match="do not conform"):
        api.civec_read(str(p), 11, 2)
    data = p.read_bytes()
    q = tmp_path / "short.dat"
    q.write_bytes(data[:24 + 8 + 10 * 8 + 5])  # ends inside the second root
    with pytest.raises(api.PsociError, match="shorter than its header"):
        api.civec_read(str(q), 10, 2)
    q.write_bytes(data[:20])
    with pytest.raises(api.PsociError, match="ends inside the header"):
        api.civec_info(str(q))
    with pytest.raises(api.PsociError, match="cannot open"):
        api.civec_info(str(tmp_path / "missing.dat"))
    with pytest.raises(api.PsociError, match="bad vector arguments"):
        api.civec_read(str(p), 10, 11)  # k > n: no unit-vector completion
    # header sizes whose product would overflow the size check are refused
    # before any size arithmetic
    for nb, nr in ((1 << 62, 3), (5, 1 << 62), (1 << 41, 1), (-4, 2), (4, -2)):
        q.write_bytes(struct.pack("<qqii", nb, nr, 18, 14) + data[24:])
        with pytest.raises(api.PsociError, match="bad vector header"):
            api.civec_info(str(q))


def test_restart_filename():
    # PsociVector::Filename (PsociVector.C:185-208) for rank 0
    assert api.restart_filename("civec.dat") == "civec-0.dat"
    assert api.restart_filename("civec") == "civec-0"
    assert api.restart_filename("a.b.c", 3) == "a.b-3.c"

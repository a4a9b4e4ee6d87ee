"""SIFS-style integral files (psoci_sifs_read / psoci_sifs_write): round
trip, and the reference's processing of the records (PsociIntegrals.C:
505-532, 716-724, 79-82) on files assembled byte by byte from the layout
documented in sifs.cpp.  Host code only: no GPU needed."""
import os
import struct

import numpy as np
import pytest

from psoci_amd import api, inputs


def _frame(payload: bytes) -> bytes:
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def _headers(nbf, ecores, n1max=4, n2max=5):
    h1 = struct.pack("<6i", 1, 1, nbf, 5, len(ecores), 0)
    h2 = (b"t" * 80 + struct.pack("<i", nbf) + b"a   " + struct.pack("<5i", 1, n1max, n1max, n2max, n2max)
          + b"mo      " * nbf + struct.pack(f"<{len(ecores)}i", *([-1] * len(ecores)))
          + struct.pack(f"<{len(ecores)}d", *ecores))
    return _frame(h1) + _frame(h2)


def _rec(last, itypea, itypeb, vals, labs, nipv, fcore=0.0):
    num = len(vals)
    p = struct.pack("<6i", num, last, itypea, itypeb, 0, 0) + struct.pack("<d", fcore)
    p += struct.pack(f"<{num}d", *vals) + struct.pack(f"<{num * nipv}i", *[x for lab in labs for x in lab])
    return _frame(p)


def test_round_trip(tmp_path):
    nbf = 6
    rng = np.random.default_rng(4)
    n1 = nbf * (nbf + 1) // 2
    t1, v1, s1 = rng.standard_normal(n1), rng.standard_normal(n1), rng.standard_normal(n1)
    two = rng.standard_normal(n1 * (n1 + 1) // 2)
    path = str(tmp_path / "moints")
    api.sifs_write(path, nbf, two, t1=t1, v1=v1, s1=s1, ecore=-12.5, title="round trip")
    n, one, two_r, ecore = api.sifs_read(path)
    assert n == nbf and ecore == -12.5
    assert np.array_equal(one, (0.0 + t1) + v1)  # summed in record order; the overlap is skipped
    assert np.array_equal(two_r, two)


def test_workload_integrals_round_trip(tmp_path):
    cfg, ints, _ = inputs.workload("fci10")
    path = str(tmp_path / "fci10.sifs")
    inputs.write_sifs(path, ints)
    back = inputs.read_sifs(path)
    assert back.nbf == ints.nbf
    assert np.array_equal(back.valone, ints.valone)
    assert np.array_equal(back.valtwo, ints.valtwo)


def test_reference_record_processing(tmp_path):
    nbf = 3
    n1 = 6
    b = _headers(nbf, [7.25, 99.0])
    # overlap (skipped), kinetic in two records (labels in either order),
    # potential, a type the reference ignores (itypea 1), then spin-orbit
    # x (overwrites; smaller label first -> negated) and z
    b += _rec(1, 0, 0, [5.0, 5.0, 5.0], [(1, 1), (2, 2), (3, 3)], 2)
    b += _rec(0, 0, 1, [1.0, 2.0, 3.0, 4.0], [(1, 1), (2, 1), (2, 2), (1, 3)], 2, fcore=3.0)
    b += _rec(1, 0, 1, [5.0, 6.0], [(3, 2), (3, 3)], 2)
    b += _rec(1, 0, 2, [0.5, 0.25], [(1, 1), (3, 1)], 2)
    b += _rec(1, 1, 0, [100.0], [(2, 1)], 2)
    b += _rec(1, 2, 0, [0.125], [(1, 2)], 2)
    b += _rec(2, 2, 2, [0.0625], [(3, 2)], 2)
    # two-electron: (21|11) written twice under permuted labels, the last wins
    b += _rec(0, 3, 0, [0.7, 0.8], [(2, 1, 1, 1), (1, 1, 1, 1)], 4)
    b += _rec(2, 3, 0, [0.9, 0.3], [(1, 1, 1, 2), (2, 2, 1, 1)], 4)
    path = tmp_path / "hand.sifs"
    path.write_bytes(b)
    n, one, two, ecore = api.sifs_read(str(path))
    assert n == 3 and ecore == 7.25
    P = inputs.pair_index
    want = np.zeros(n1)
    want[P(0, 0)] = 1.0 + 0.5
    want[P(1, 0)] = -0.125          # kinetic 2.0 overwritten by spin-orbit x, labels (1, 2): negated
    want[P(1, 1)] = 3.0
    want[P(2, 0)] = 4.0 + 0.25
    want[P(2, 1)] = 0.0625          # kinetic 5.0 overwritten by spin-orbit z, labels (3, 2)
    want[P(2, 2)] = 6.0
    assert np.array_equal(one, want)
    C = inputs.canonical_index
    wtwo = np.zeros(len(two))
    wtwo[C(1, 0, 0, 0)] = 0.9
    wtwo[C(0, 0, 0, 0)] = 0.8
    wtwo[C(1, 1, 0, 0)] = 0.3
    assert np.array_equal(two, wtwo)


def test_errors(tmp_path):
    with pytest.raises(api.PsociError) as e:
        api.sifs_read(str(tmp_path / "missing"))
    assert e.value.code == api.ERR_ARG and "missing" in str(e.value)
    path = tmp_path / "short"
    path.write_bytes(_headers(2, [0.0])[:-10])
    with pytest.raises(api.PsociError) as e:
        api.sifs_read(str(path))
    assert e.value.code == api.ERR_ARG
    # a record larger than the declared capacity
    path = tmp_path / "big"
    path.write_bytes(_headers(2, [0.0], n1max=1) + _rec(2, 0, 1, [1.0, 2.0], [(1, 1), (2, 2)], 2))
    with pytest.raises(api.PsociError) as e:
        api.sifs_read(str(path))
    assert "capacity" in str(e.value)

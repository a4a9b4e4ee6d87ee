"""Damaged input files through the library's readers (context free, no GPU):
seeded truncations, byte flips, insertions and deletions of valid
configuration, SIFS-style and CI-vector files.  Every read either succeeds
or returns an error status with a message; none may crash the process or
read out of bounds (the same mutations were run against a host build with
AddressSanitizer and UndefinedBehaviorSanitizer)."""
import random

import numpy as np
import pytest

from psoci_amd import api, inputs


def mutate(rng, good, alphabet=None):
    b = bytearray(good)
    m = rng.randrange(4)
    if m == 0:
        return bytes(b[:rng.randrange(len(b) + 1)])
    if m == 1:
        for _ in range(rng.randrange(1, 8)):
            b[rng.randrange(len(b))] = rng.choice(alphabet) if alphabet else rng.randrange(256)
        return bytes(b)
    p = rng.randrange(len(b))
    if m == 2:
        ins = bytes(rng.choice(alphabet) if alphabet else rng.randrange(256) for _ in range(rng.randrange(1, 20)))
        return bytes(b[:p] + ins + b[p:])
    return bytes(b[:p] + b[p + rng.randrange(1, 64):])


def survives(call):
    try:
        call()
        return "ok"
    except (api.PsociError, inputs.ConfigFileError) as e:
        assert str(e)
        return "error"


@pytest.mark.parametrize("seed", [1])
def test_damaged_files(tmp_path, seed):
    rng = random.Random(seed)
    cfg = inputs.mrcisd_configs(4, 6, 4)
    cpath = tmp_path / "good.cfg"
    inputs.write_configs(str(cpath), cfg)
    ints = inputs.make_integrals(6, 3)
    spath = tmp_path / "good.sifs"
    inputs.write_sifs(str(spath), ints)
    vpath = tmp_path / "good.civ"
    api.civec_write(str(vpath), np.random.default_rng(seed).standard_normal((40, 3)))
    goods = {"cfg": cpath.read_bytes(), "sifs": spath.read_bytes(), "civ": vpath.read_bytes()}
    bad = tmp_path / "bad"
    seen = {}
    for _ in range(120):
        for kind, good in goods.items():
            bad.write_bytes(mutate(rng, good, b"0123456789 \t\n\r-+x\x00" if kind == "cfg" else None))
            if kind == "cfg":
                r = survives(lambda: api.read_configs_file(str(bad), sorted=rng.random() < 0.5))
            elif kind == "sifs":
                r = survives(lambda: api.sifs_read(str(bad)))
            else:
                r = survives(lambda: api.civec_read(str(bad), 40, 2))
            seen[(kind, r)] = seen.get((kind, r), 0) + 1
    for kind in goods:
        assert seen.get((kind, "error"), 0) > 0

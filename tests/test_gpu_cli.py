"""The compiled C99 driver (psoci_amd/cli/psoci_run, a caller of
include/psoci.h with no Python in between) on the 246-SEF FCI space: files
in, eigenvalues out, checked against LAPACK on the committed golden
Hamiltonian; the stored-H checkpoint, the vector restart and the direct
mixed-precision path through the same program."""
import os
import re
import subprocess

import numpy as np
import pytest

from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(api.LIB_PATH), "cli", "psoci_run")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fci6_h.npz")


def run(*args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    ev = [float(m.group(2)) for m in re.finditer(r"Eigenvalue (\d+) Total energy \(au\) \S+ Electronic \(au\) (\S+)",
                                                  r.stdout)]
    tot = [float(m.group(1)) for m in re.finditer(r"Total energy \(au\) (\S+)", r.stdout)]
    return np.array(ev), np.array(tot), r.stdout


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    cfg, ints, _ = inputs.workload("fci6")
    ints = inputs.Integrals(nbf=ints.nbf, valone=ints.valone, valtwo=ints.valtwo, ecore=-7.25)
    inputs.write_configs(str(d / "fci6.configs"), cfg)
    inputs.write_sifs(str(d / "fci6.sifs"), ints)
    g = np.load(GOLDEN)
    return d, g["evals"]


def test_cli_stored_fp64_matches_golden(files):
    d, w = files
    ev, tot, out = run("-c", str(d / "fci6.configs"), "-m", str(d / "fci6.sifs"), "-n", "4", "-F", "1e-9",
                       "-v", str(d / "vec.bin"), "-W", "--hamiltonian-file", str(d / "h.bin"))
    np.testing.assert_allclose(ev, w[:4], atol=1e-9, rtol=0)
    np.testing.assert_allclose(tot, w[:4] - 7.25, atol=1e-9, rtol=0)
    # the checkpoint and the vectors written above: load instead of build,
    # restart from the vectors (converged already)
    ev2, _, out2 = run("-c", str(d / "fci6.configs"), "-m", str(d / "fci6.sifs"), "-n", "4", "-F", "1e-9",
                       "-v", str(d / "vec.bin"), "-r", "-H", "--hamiltonian-file", str(d / "h.bin"))
    assert "Hamiltonian loaded" in out2 and "restart: 4 vectors" in out2
    np.testing.assert_allclose(ev2, w[:4], atol=1e-9, rtol=0)
    it = int(re.search(r"Number of total iterations is (\d+)", out2).group(1))
    assert it <= 2


def test_cli_direct_mixed(files):
    d, w = files
    ev, _, _ = run("-c", str(d / "fci6.configs"), "-m", str(d / "fci6.sifs"), "-n", "4", "-F", "1e-8", "--direct",
                   "--prec", "mixed", "--guard", "2")
    np.testing.assert_allclose(ev, w[:4], atol=1e-9, rtol=0)

"""Builds the product library libpsoci_amd.so in-tree with hipcc for gfx950.

Usage: python -m psoci_amd.build   (or psoci_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsoci_amd.so")
SOURCES = ["space.cpp", "api.cpp", "davidson.cpp", "sifs.cpp", "configs.cpp", "restart.cpp", "nos.cpp", "hbuild.hip", "sigma.hip", "linalg.hip", "direct.hip", "rdm.hip", "analysis.hip", "noalloc.hip", "transition.hip", "ras.cpp", "strsig.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no fused multiply-add unless written explicitly: the element arithmetic
    # must keep the reference's rounding sequence
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _obj(src: str) -> str:
    return os.path.join(CSRC, "_obj", src + ".o")


def build(verbose: bool = False, jobs: int = 4) -> str:
    os.makedirs(os.path.join(CSRC, "_obj"), exist_ok=True)
    headers = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".hpp", ".h"))]
    headers.append(os.path.join(HERE, "..", "include", "psoci.h"))
    newest_header = max(os.path.getmtime(h) for h in headers)
    procs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = _obj(s)
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmd = [HIPCC, *FLAGS, *lang, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait() if procs[0][1].poll() is None else None
            for _, p in procs:
                if p.poll() is None:
                    p.wait()
                    break
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((s, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *[_obj(s) for s in SOURCES]]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    del newest_header
    build_cli(verbose)
    return LIB


CLI_DIR = os.path.join(HERE, "cli")
INCLUDE = os.path.join(HERE, "..", "include")


def build_cli(verbose: bool = False) -> None:
    """The C99 programs over the ABI: psoci_run (the command-line driver) and
    abi_layout (struct layout probe), built with gcc against include/psoci.h;
    psoci_run links libpsoci_amd.so from its parent directory."""
    cc = os.environ.get("CC", "gcc")
    common = [cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-O2", "-I", INCLUDE]
    jobs = [
        common + [os.path.join(CLI_DIR, "psoci_run.c"), "-o", os.path.join(CLI_DIR, "psoci_run"), "-L", HERE,
                  "-lpsoci_amd", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"],
        common + [os.path.join(CLI_DIR, "abi_layout.c"), "-o", os.path.join(CLI_DIR, "abi_layout")],
    ]
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

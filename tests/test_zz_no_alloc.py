"""psoci_sigma_noalloc (the allocation-free sigma) through the C ABI: its
output equals psoci_sigma's on the same device buffers, bit for bit (the
same kernels; only the staging memory differs), on two calls in a row, the
second with a smaller block; blocks whose staging exceeds the scratch and
operators that are not built yet are refused with an error."""
import ctypes

import numpy as np
import pytest

from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


def noalloc_fn():
    L = api.load()
    f = L.psoci_sigma_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return f


def call(H, mode, prec, X, Y):
    import torch
    torch.cuda.synchronize()
    return noalloc_fn()(H._h, mode, prec, X.shape[1], ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(Y.data_ptr()))


def run_pair(H, mode, k, prec, seed):
    import torch
    dt = torch.float32 if prec == api.FP32 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    X = torch.randn(H.n, k, device="cuda", dtype=torch.float64, generator=g).to(dt).contiguous()
    Ya = torch.full_like(X, float("nan"))
    Yb = torch.full_like(X, float("nan"))
    H.sigma_device(X, Ya, mode, prec)
    rc = call(H, mode, prec, X, Yb)
    assert rc == api.OK, H.L.psoci_last_error(H._h).decode()
    torch.cuda.synchronize()
    assert torch.equal(Ya, Yb), (mode, k, prec)


@pytest.mark.parametrize("mode,engine", [(api.STORED, 0), (api.DIRECT, 1), (api.DIRECT, 2)])
def test_noalloc_equals_sigma_fci10(fci10, mode, engine):
    cfg, ints, _ = fci10
    H = api.Hamiltonian(cfg, ints)
    H.build()
    if mode == api.DIRECT:
        assert H.direct_engine(engine) == engine  # builds the direct sigma's operator
    # staged blocks (12 -> width 16, then 3 -> width 4), both precisions,
    # then an in-place block of a kernel width
    run_pair(H, mode, 12, api.FP64, 1)
    run_pair(H, mode, 3, api.FP64, 2)
    run_pair(H, mode, 12, api.FP32, 3)
    run_pair(H, mode, 3, api.FP32, 4)
    run_pair(H, mode, 8, api.FP64, 5)
    H.close()


def test_noalloc_refusals():
    import torch
    cfg, ints, _ = inputs.workload("fci6")
    H = api.Hamiltonian(cfg, ints)
    X = torch.zeros(H.n, 4, device="cuda", dtype=torch.float64)
    Y = torch.zeros_like(X)
    # nothing built yet: building allocates, so the call refuses
    assert call(H, api.STORED, api.FP64, X, Y) == api.ERR_STATE
    assert call(H, api.DIRECT, api.FP64, X, Y) == api.ERR_STATE
    H.close()
    cfg, ints, _ = inputs.workload("mrci8x24")
    H = api.Hamiltonian(cfg, ints)
    H.direct_info()
    # 12 columns stage through 2 * n * 16 * 8 bytes (145 MB) > 64 MiB
    X = torch.zeros(H.n, 12, device="cuda", dtype=torch.float64)
    Y = torch.zeros_like(X)
    assert call(H, api.DIRECT, api.FP64, X, Y) == api.ERR_ARG
    assert "staging" in H.L.psoci_last_error(H._h).decode()
    # a compact block of a kernel width needs no staging at any size
    run_pair(H, api.DIRECT, 16, api.FP32, 6)
    H.close()

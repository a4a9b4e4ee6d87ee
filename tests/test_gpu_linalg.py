"""FP64 MFMA subspace projection (A^T B over n rows) against numpy, for the
shapes the Davidson uses: tall-skinny, ragged tiles, column offsets."""
import numpy as np
import pytest

from psoci_amd import api, inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    cfg, ints, _ = inputs.workload("fci6")
    H = api.Hamiltonian(cfg, ints)
    yield H
    H.close()


@pytest.mark.parametrize("n,m,b,ld", [(246, 4, 4, 20), (246, 20, 3, 20), (31650, 37, 17, 40), (100003, 80, 16, 96),
                                      (5, 1, 1, 1)])
def test_project_matches_numpy(ctx, n, m, b, ld):
    import torch
    g = torch.Generator().manual_seed(n + m)
    A = torch.randn(n, ld, generator=g, dtype=torch.float64)
    B = torch.randn(n, ld, generator=g, dtype=torch.float64)
    ref = A[:, :m].numpy().T @ B[:, ld - b:].numpy()
    Ad = A.cuda()
    Bd = B.cuda()
    out = ctx.project(Ad[:, :m], Bd[:, ld - b:])
    assert np.abs(out - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), np.abs(out - ref).max()


@pytest.mark.parametrize("n,m,k,lda,ldo,alpha", [(246, 4, 4, 20, 4, 0.0), (246, 20, 3, 20, 7, 1.0),
                                                 (31650, 37, 17, 40, 17, 0.0), (100003, 120, 24, 120, 24, 0.0),
                                                 (100003, 120, 48, 120, 48, -0.5), (5, 1, 1, 1, 1, 2.0),
                                                 (4097, 130, 129, 130, 140, 1.0)])
def test_combine_matches_numpy(ctx, n, m, k, lda, ldo, alpha):
    """The MFMA basis update Out = alpha Out + A Y: ragged row tiles, t
    chunks and column tiles (k up to 129: two column slices)."""
    import torch
    g = torch.Generator().manual_seed(n + m + k)
    A = torch.randn(n, lda, generator=g, dtype=torch.float64)
    O = torch.randn(n, ldo, generator=g, dtype=torch.float64)
    Y = np.random.default_rng(m).standard_normal((m, k))
    ref = O.numpy().copy()
    ref[:, :k] = alpha * ref[:, :k] + A[:, :m].numpy() @ Y
    Od = O.cuda()
    ctx.combine(A.cuda()[:, :m], Y, Od[:, :k], alpha)
    out = Od.cpu().numpy()
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(out - ref).max() <= 1e-12 * scale * max(1.0, m / 16)

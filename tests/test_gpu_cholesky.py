"""The batched FP64 Cholesky factorisation (chol.hip: right-looking panels
of 128 columns with look-ahead on a second stream), as the mixed path
factors |A+|, against numpy's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(a):
    """L of each column-major matrix in a (batch, n, n); the strict upper
    triangle of the input holds garbage that the factorisation must not read."""
    import torch
    from gammcor_amd import _lib
    batch, n, _ = a.shape
    # a row-major (n, n) view of a column-major matrix is its transpose: the
    # column-major strict upper triangle is the view's strict lower one
    ain = a.copy()
    for b in range(batch):
        ain[b][np.tril_indices(n, -1)] = 1e6
    d = torch.from_numpy(ain).cuda()
    info = np.zeros(batch, dtype=np.int32)
    _lib.call("gmc_potrf_dev", n, batch, d.data_ptr(), info.ctypes.data, None)
    return d.cpu().numpy().transpose(0, 2, 1), info


@pytest.mark.parametrize("n,batch", [(1, 1), (127, 2), (128, 1), (129, 3), (300, 2), (517, 2)])
def test_potrf_matches_numpy(cuda, n, batch):
    rng = np.random.default_rng(n + 7 * batch)
    a = np.empty((batch, n, n))
    for b in range(batch):
        x = rng.standard_normal((n, n))
        a[b] = x @ x.T / n + (0.1 + b) * np.eye(n)
    print(f"potrf n={n} batch={batch}", flush=True)
    l, info = _run(a)
    assert (info == 0).all(), info
    for b in range(batch):
        ref = np.linalg.cholesky(a[b])
        assert np.max(np.abs(l[b] - ref)) < 1e-12 * np.max(np.abs(ref)), np.max(np.abs(l[b] - ref))
        assert not np.triu(l[b], 1).any()


def test_potrf_reports_the_first_nonpositive_pivot(cuda):
    """A pivot that is not positive in a later panel: info names its 1-based
    column, and the other problem of the batch is factored as alone."""
    n = 300
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n, n))
    good = x @ x.T / n + 0.5 * np.eye(n)
    bad = np.diag(np.linspace(1.0, 2.0, n))
    bad[200, 200] = -1.0
    print("potrf n=300 batch=2 with a negative pivot at column 201", flush=True)
    l, info = _run(np.stack([good, bad]))
    assert list(info) == [0, 201]
    ref = np.linalg.cholesky(good)
    assert np.max(np.abs(l[0] - ref)) < 1e-12 * np.max(np.abs(ref))

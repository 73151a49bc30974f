"""gmc_syevd_noalloc_dev, the allocation-free form of the batched eigensolver
(its workspace in one static device array per GPU), against gmc_syevd_dev
through the C ABI: two calls in a row, the second smaller (the array keeps
the first call's contents), in FP64 and FP32, and a problem set larger than
the array rejected with a message.

Tolerances: eigenvalues within 1e-12 (FP64) / 2e-5 (FP32) of the matrix
norm; eigenvectors of the separated spectra used here agree with the
allocating entry's up to sign to 1e-10 / 1e-3 (max |V1^T V2| - I), with
residual and orthogonality at the same levels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _matrices(n, batch, seed, dtype, cuda):
    import torch
    rng = np.random.default_rng(seed)
    out = np.empty((batch, n, n))
    for b in range(batch):
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        # distinct eigenvalues at least 1/n apart, both signs
        d = np.sort(rng.permutation(np.linspace(-1.0, 2.0, n)) + rng.uniform(0, 0.2 / n, n))
        out[b] = (q * d) @ q.T
        out[b] = 0.5 * (out[b] + out[b].T)
    return torch.from_numpy(out).to(device=cuda, dtype=dtype)


def _solve(name, prec, a, n, batch):
    import torch
    from gammcor_amd import _lib
    v = a.clone()     # column-major n x n per problem: the symmetric matrix is its own transpose
    w = torch.zeros(batch, n, dtype=a.dtype, device=a.device)
    _lib.call(name, prec, n, batch, v.data_ptr(), w.data_ptr(), None)
    # eigenvectors are columns of the column-major result: rows of v[b]
    return v.transpose(1, 2).contiguous(), w


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_allocating_entry(cuda, prec):
    import torch
    from gammcor_amd import host
    code = host.GMC_FP64 if prec == "fp64" else host.GMC_FP32
    dtype = torch.float64 if prec == "fp64" else torch.float32
    tol_w, tol_v = (1e-12, 1e-10) if prec == "fp64" else (2e-5, 1e-3)
    for n, batch, seed in ((300, 4, 1), (130, 2, 2)):
        print(f"no-alloc syevd {prec}: n={n} batch={batch}", flush=True)
        a = _matrices(n, batch, seed, dtype, cuda)
        v1, w1 = _solve("gmc_syevd_dev", code, a, n, batch)
        v2, w2 = _solve("gmc_syevd_noalloc_dev", code, a, n, batch)
        a64, v1, v2 = a.double(), v1.double(), v2.double()
        w1, w2 = w1.double(), w2.double()
        nrm = torch.linalg.matrix_norm(a64, ord=2).max().item()
        assert (w2 - w1).abs().max().item() < tol_w * nrm
        eye = torch.eye(n, dtype=torch.float64, device=cuda)
        overlap = torch.bmm(v1.transpose(1, 2), v2).abs()
        assert (overlap - eye).abs().max().item() < tol_v
        res = torch.bmm(a64, v2) - v2 * w2[:, None, :]
        assert res.abs().max().item() < tol_v * nrm
        assert (torch.bmm(v2.transpose(1, 2), v2) - eye).abs().max().item() < tol_v


def test_noalloc_rejects_a_problem_set_beyond_the_array(cuda):
    import torch
    from gammcor_amd import host, _lib
    n = 6000        # FP64 workspace about 4 n^2 8 B = 1.15 GB > 256 MiB
    a = torch.zeros(n, n, dtype=torch.float64, device=cuda)
    w = torch.zeros(n, dtype=torch.float64, device=cuda)
    with pytest.raises(_lib.GmcError, match="device scratch"):
        _lib.call("gmc_syevd_noalloc_dev", host.GMC_FP64, n, 1, a.data_ptr(), w.data_ptr(), None)
    # the array is intact for the next call
    b = _matrices(64, 1, 3, torch.float64, cuda)
    v, wv = _solve("gmc_syevd_noalloc_dev", host.GMC_FP64, b, 64, 1)
    ref = torch.linalg.eigvalsh(b[0].cpu())
    assert (wv[0].cpu() - ref).abs().max().item() < 1e-12 * 3

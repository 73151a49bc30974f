"""Batched symmetric eigensolver (tridiagonalisation + divide & conquer +
back transformation) against LAPACK through numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mats(kind, n, batch, seed):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(batch):
        if kind == "rand":
            a = rng.standard_normal((n, n))
            a = a + a.T
        elif kind == "clustered":
            q, _ = np.linalg.qr(rng.standard_normal((n, n)))
            lam = np.repeat(rng.standard_normal((n + 3) // 4), 4)[:n] + 1e-9 * rng.standard_normal(n)
            a = (q * lam) @ q.T
        elif kind == "degenerate":
            q, _ = np.linalg.qr(rng.standard_normal((n, n)))
            lam = np.where(np.arange(n) % 3 == 0, 1.0, -2.0)
            a = (q * lam) @ q.T
        elif kind == "diag":
            a = np.diag(rng.standard_normal(n))
        elif kind == "graded":
            a = rng.standard_normal((n, n))
            a = a + a.T
            s = np.logspace(0, -6, n)
            a = a * s[:, None] * s[None, :]
        out.append(0.5 * (a + a.T))
    return np.stack(out)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind,n,batch", [("rand", 1, 1), ("rand", 2, 2), ("rand", 5, 3), ("rand", 33, 2),
                                          ("rand", 100, 1), ("rand", 257, 4), ("rand", 600, 2),
                                          ("clustered", 130, 2), ("degenerate", 97, 2), ("diag", 70, 1),
                                          ("graded", 120, 1), ("rand", 256, 3), ("rand", 452, 2),
                                          ("rand", 1024, 1), ("clustered", 516, 2),
                                          # orders padded to the vector width: diag(A, c I) inside
                                          ("clustered", 301, 2), ("degenerate", 263, 2), ("graded", 259, 1),
                                          ("diag", 258, 1), ("diag", 255, 1)])
def test_syevd(cuda, prec, kind, n, batch):
    import torch
    from gammcor_amd import _lib
    a = _mats(kind, n, batch, n + batch)
    dt = torch.float64 if prec == "f64" else torch.float32
    # column-major: a[b] symmetric so row-major data is the same matrix
    da = torch.from_numpy(a).to(cuda, dt).contiguous()
    dw = torch.zeros(batch, n, device=cuda, dtype=dt)
    print(f"syevd {prec} {kind} n={n} batch={batch}", flush=True)
    _lib.call("gmc_syevd_dev", 0 if prec == "f64" else 1, n, batch, da.data_ptr(), dw.data_ptr(), None)
    vecs = da.double().cpu().numpy().transpose(0, 2, 1)   # vecs[b][:, k] = eigenvector k
    w = dw.double().cpu().numpy()
    eps = 2.2e-16 if prec == "f64" else 1.2e-7
    for b in range(batch):
        ref = np.linalg.eigvalsh(a[b])
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(w[b] - ref)) < 200 * n ** 0.5 * eps * scale
        res = np.linalg.norm(a[b] @ vecs[b] - vecs[b] * w[b][None, :]) / scale
        orth = np.linalg.norm(vecs[b].T @ vecs[b] - np.eye(n))
        assert res < 200 * n * eps, res
        assert orth < 200 * n * eps, orth


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [10500, 10501])
def test_syevd_large_order_global_v(cuda, prec, n):
    """n large enough that the panel kernel keeps the Householder vector in
    global memory instead of LDS (2 n sizeof(T) > 160 KiB for FP64 at
    n = 10500; FP32 stays in LDS); the odd order is solved padded to the
    vector width.  Checked on the GPU with torch's FP64 products
    (independent of this library)."""
    import torch
    from gammcor_amd import _lib
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device=cuda).manual_seed(5)
    a = torch.randn(n, n, dtype=torch.float64, device=cuda, generator=g)
    a = 0.5 * (a + a.T)
    da = a.to(dt).clone()                 # .to() returns `a` itself for float64; the solver overwrites da
    dw = torch.zeros(1, n, device=cuda, dtype=dt)
    print(f"syevd {prec} n={n} batch=1 (global v)", flush=True)
    _lib.call("gmc_syevd_dev", 0 if prec == "f64" else 1, n, 1, da.data_ptr(), dw.data_ptr(), None)
    v = da.double().T.contiguous()        # row-major view of the column-major result: v[:, k] = vector k
    w = dw.double()[0]
    eps = 2.2e-16 if prec == "f64" else 1.2e-7
    scale = float(w.abs().max())
    res = torch.linalg.norm(a.to(v.dtype) @ v - v * w[None, :]) / scale
    orth = torch.linalg.norm(v.T @ v - torch.eye(n, dtype=v.dtype, device=cuda))
    assert float(res) < 200 * n * eps, float(res)
    assert float(orth) < 200 * n * eps, float(orth)
    assert bool(torch.all(w[1:] >= w[:-1]))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_syevd_erpa_batch_shape(cuda, prec):
    """The ERPA batch shape at N=240: n = 4524, 16 problems, 16 workgroups per
    problem.  In FP64 the own rows of the V and W' panels do not fit in LDS
    beside u and v and are read from the global panel; in FP32 they stay in
    LDS.  Checked with torch's FP64 batched products on the GPU."""
    import torch
    from gammcor_amd import _lib
    n, batch = 4524, 16
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device=cuda).manual_seed(7)
    a = torch.randn(batch, n, n, dtype=torch.float64, device=cuda, generator=g)
    a = (0.5 * (a + a.transpose(1, 2))).contiguous()
    da = a.to(dt).clone()
    dw = torch.zeros(batch, n, device=cuda, dtype=dt)
    print(f"syevd {prec} n={n} batch={batch} (ERPA batch shape)", flush=True)
    _lib.call("gmc_syevd_dev", 0 if prec == "f64" else 1, n, batch, da.data_ptr(), dw.data_ptr(), None)
    v = da.double().transpose(1, 2).contiguous()   # v[b][:, k] = eigenvector k
    del da
    w = dw.double()
    eps = 2.2e-16 if prec == "f64" else 1.2e-7
    eye = torch.eye(n, dtype=torch.float64, device=cuda)
    for b in range(batch):
        scale = float(w[b].abs().max())
        res = torch.linalg.norm(a[b] @ v[b] - v[b] * w[b][None, :]) / scale
        orth = torch.linalg.norm(v[b].T.contiguous() @ v[b] - eye)
        assert float(res) < 200 * n * eps, (b, float(res))
        assert float(orth) < 200 * n * eps, (b, float(orth))
        assert bool(torch.all(w[b, 1:] >= w[b, :-1]))

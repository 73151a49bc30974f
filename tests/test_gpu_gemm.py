"""MFMA GEMM building block against a float64 torch reference."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k,batch", [(64, 64, 64, 1), (37, 130, 71, 3), (240, 240, 240, 2), (5, 3, 200, 1)])
def test_gemm(cuda, prec, ta, tb, m, n, k, batch):
    import torch
    from gammcor_amd import _lib
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator().manual_seed(m * 7 + n * 3 + k)
    # column-major (rows x cols) matrices are held as row-major (cols x rows) tensors
    ra, ca = (k, m) if ta else (m, k)
    rb, cb = (n, k) if tb else (k, n)
    a = torch.randn(batch, ca, ra, generator=g, dtype=torch.float64)
    b = torch.randn(batch, cb, rb, generator=g, dtype=torch.float64)
    c = torch.randn(batch, n, m, generator=g, dtype=torch.float64)
    ref_a = a.transpose(1, 2)
    ref_b = b.transpose(1, 2)
    opa = ref_a.transpose(1, 2) if ta else ref_a
    opb = ref_b.transpose(1, 2) if tb else ref_b
    alpha, beta = 1.5, -0.5
    ref = alpha * opa @ opb + beta * c.transpose(1, 2)
    da, db, dc = a.to(cuda, dt).contiguous(), b.to(cuda, dt).contiguous(), c.to(cuda, dt).contiguous()
    lda, ldb = ra, rb
    _lib.call("gmc_gemm_dev", 0 if prec == "f64" else 1, ta, tb, m, n, k, alpha, da.data_ptr(), lda,
              da[0].numel(), db.data_ptr(), ldb, db[0].numel(), beta, dc.data_ptr(), m, dc[0].numel(), batch, None)
    out = dc.double().cpu().transpose(1, 2)
    tol = 1e-12 if prec == "f64" else 2e-4
    assert torch.allclose(out, ref, rtol=tol, atol=tol * np.sqrt(k)), float((out - ref).abs().max())


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ta,tb", [(0, 1), (0, 0), (1, 1)])
@pytest.mark.parametrize("n,k,batch", [(1, 1, 1), (33, 7, 2), (100, 64, 3), (257, 64, 2), (130, 65, 1),
                                       (64, 200, 1), (31, 3, 5)])
def test_gemm_lower(cuda, prec, ta, tb, n, k, batch):
    """The lower-triangle form (C square, i >= j written): FP32 with op(B) =
    B^T and k <= 64 takes the one-wave-per-32x32-tile kernel, the other
    shapes the 64x64 LDS tile.  Leading dimensions exceed the row counts;
    the strict upper triangle keeps its values."""
    import torch
    from gammcor_amd import _lib
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator().manual_seed(n * 11 + k * 5 + batch)
    pad = 3
    ra, ca = (k, n) if ta else (n, k)
    rb, cb = (n, k) if tb else (k, n)
    lda, ldb, ldc = ra + pad, rb + pad, n + pad
    a = torch.randn(batch, ca, lda, generator=g, dtype=torch.float64)
    b = torch.randn(batch, cb, ldb, generator=g, dtype=torch.float64)
    c = torch.randn(batch, n, ldc, generator=g, dtype=torch.float64)
    ref_a = a[:, :, :ra].transpose(1, 2)
    ref_b = b[:, :, :rb].transpose(1, 2)
    opa = ref_a.transpose(1, 2) if ta else ref_a
    opb = ref_b.transpose(1, 2) if tb else ref_b
    alpha, beta = -1.0, 1.0
    full = alpha * opa @ opb + beta * c[:, :, :n].transpose(1, 2)
    low = torch.tril(torch.ones(n, n, dtype=torch.bool))
    ref = torch.where(low, full, c[:, :, :n].transpose(1, 2))
    da, db, dc = a.to(cuda, dt).contiguous(), b.to(cuda, dt).contiguous(), c.to(cuda, dt).contiguous()
    _lib.call("gmc_gemm_lower_dev", 0 if prec == "f64" else 1, ta, tb, n, k, alpha, da.data_ptr(), lda,
              da[0].numel(), db.data_ptr(), ldb, db[0].numel(), beta, dc.data_ptr(), ldc, dc[0].numel(), batch, None)
    out = dc.double().cpu()
    tol = 1e-12 if prec == "f64" else 2e-4
    got = out[:, :, :n].transpose(1, 2)
    assert torch.allclose(got, ref, rtol=tol, atol=tol * np.sqrt(k)), float((got - ref).abs().max())
    # the padding rows beyond n in every column are never touched
    assert torch.equal(out[:, :, n:], c.to(dt).double()[:, :, n:])

"""The arithmetic of the mixed path's refinement update X E, emulated on the
CPU with numpy's IEEE half (round to nearest even, like __float2half):

- the hi + 2^-11 lo split of a value v (|v| <= 1, lo scaled by 2^11 into
  half's normal range) represents v to 2^-22 relative, and to 1.5e-11
  absolute where v is below half's normal range;
- the sum of the three split products (hi hi, hi lo, lo hi) with FP32
  accumulation is as accurate as an FP32 product;
- a product of plain half inputs errs by about 1.6e-4 max|E| (RMS, in the
  eigenvector coordinates X^T dP where the step's own remaining error is
  measured), which for a correction of at least 5e-3 stays below a twentieth
  of the max|E|^2 a step leaves (largest element: a fifth)."""
import numpy as np


def _split(v32):
    h = v32.astype(np.float16)
    lo = ((v32 - h.astype(np.float32)) * np.float32(2048.0)).astype(np.float16)
    return h, lo


def _prod(a16, b16):
    # half x half products are exact in FP32; FP32 accumulation
    return a16.astype(np.float32) @ b16.astype(np.float32)


def test_split_represents_values_to_2e22_relative():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-1e-4, 1e-4, 20000),
                        rng.uniform(-1e-7, 1e-7, 2000), [0.0, 1.0, -1.0, 6.1e-5, 5.96e-8]]).astype(np.float32)
    h, lo = _split(v)
    assert np.all(np.isfinite(h)) and np.all(np.isfinite(lo))
    rec = h.astype(np.float64) + lo.astype(np.float64) / 2048.0
    err = np.abs(rec - v.astype(np.float64))
    assert np.all(err <= np.maximum(2.0 ** -22 * np.abs(v), 1.5e-11))
    tiny = np.abs(v) < 6.1e-5
    assert np.max(err[tiny]) <= 1.5e-11


def test_three_split_products_match_fp32():
    rng = np.random.default_rng(1)
    n = 384
    x, _ = np.linalg.qr(rng.standard_normal((n, n)))          # eigenvector-like, |x| <= 1
    e = rng.standard_normal((n, n)) * 3e-4                     # a late step's corrections
    emax = np.abs(e - np.diag(np.diag(e))).max()
    den = max(emax, 1e-4)
    exact = x @ e
    xh, xl = _split(x.astype(np.float32))
    eh, el = _split((e / den).astype(np.float32))
    p = den * _prod(xh, eh) + (den / 2048.0) * (_prod(xh, el) + _prod(xl, eh))
    p32 = x.astype(np.float32) @ e.astype(np.float32)
    err_split = np.max(np.abs(p - exact)) / np.max(np.abs(exact))
    err_fp32 = np.max(np.abs(p32 - exact)) / np.max(np.abs(exact))
    assert err_split < 4.0 * err_fp32 + 1e-7
    assert err_split < 1e-6


def test_plain_half_product_against_the_step_error():
    rng = np.random.default_rng(2)
    n = 384
    x, _ = np.linalg.qr(rng.standard_normal((n, n)))
    for emax in (0.125, 2.6e-2, 5e-3):
        e = rng.uniform(-1, 1, (n, n)) * emax
        exact = x @ e
        p = emax * _prod(x.astype(np.float16), (e / emax).astype(np.float16))
        # the update's error in the eigenvector coordinates, where the error
        # a step leaves (about max|E|^2 per element) is measured
        d = x.T @ (p - exact)
        rms = np.sqrt(np.mean(d * d))
        assert rms < 2.5e-4 * emax
        assert rms < 0.05 * emax * emax
        assert np.max(np.abs(d)) < 0.2 * emax * emax

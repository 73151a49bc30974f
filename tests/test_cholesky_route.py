"""The mixed path's Cholesky route for the symmetric ERPA problem, emulated in
numpy on reference-fixture systems whose A+ is indefinite.

The reference (ERPASYMM1, interpa.f:1693-1740) takes S = |A+|^(1/2) from the
eigenvalues of A+ with their absolute values, M = S A- S, and Y = S Z.  Any
factor L with L L^T = |A+| gives the same Y: L = S Q with Q orthogonal, so
L^T A- L = Q^T M Q has the eigenvectors Q^T Z and L Q^T Z = S Z.  The route:

- the FP32 eigenvectors of A+; the K smallest columns (every negative
  eigenvalue and at least 8 more) refined in FP64 by column-restricted
  Ogita-Aishima steps against the full FP32 basis, which converge linearly
  (each step gains the FP32 basis' accuracy, about 1e-4);
- |A+| = A+ - 2 sum_{lambda < 0} lambda x x^T, its Cholesky factor L;
- M' = L^T A- L, its eigenvectors Z', Y = L Z', then the reference's
  normalisation and filter.

W(alpha) of the route matches the reference's fixtures to 1e-9 Eh; a |A+|
with the reflection's factor 2 dropped does not."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SMALL_OM2 = 1e-6


def gold(name):
    with open(os.path.join(GOLD, f"ref_{name}.json")) as f:
        return json.load(f)


def partial_refine(a, x32, w32, k, steps=8, tol=1e-11):
    """Column-restricted Ogita-Aishima: the first k columns of the FP32
    eigenvector matrix x32 refined in FP64; the other columns and their FP32
    eigenvalues serve as the fixed basis.  Stops when max |E| over the
    columns with a negative eigenvalue is below tol (the rounding floor of E
    is about 1e-13 here).  Returns (X_J, lambda_J, history of
    max |E| off the diagonal)."""
    n = a.shape[0]
    x = x32.astype(np.float64).copy()
    lam_c = w32.astype(np.float64)
    hist = []
    for _ in range(steps):
        xj = x[:, :k]
        s = x.T @ (a @ xj)
        g = x.T @ xj
        lam_j = np.diag(s[:k]) / np.diag(g[:k])
        lam = lam_c.copy()
        lam[:k] = lam_j
        gap = lam_j[None, :] - lam[:, None]
        delta = 64 * np.finfo(np.float64).eps * np.max(np.abs(lam))
        with np.errstate(divide="ignore", invalid="ignore"):
            e = (s - lam_j[None, :] * g) / gap
        close = (np.abs(gap) <= delta) | ~(np.abs(e) < 0.125)
        e[close] = -0.5 * g[close]
        off = np.ones_like(e, dtype=bool)
        off[np.arange(k), np.arange(k)] = False
        off &= ~(lam_j >= 0.0)[None, :]     # only the columns |A+| takes
        hist.append(float(np.max(np.abs(e[off]), initial=0.0)))
        e[np.arange(k), np.arange(k)] = 0.5 * (1.0 - np.diag(g[:k]))
        x[:, :k] = xj + x @ e
        if hist[-1] < tol:
            break
    xj = x[:, :k]
    lam_j = np.einsum("ik,ij,jk->k", xj, a, xj) / np.einsum("ik,ik->k", xj, xj)
    return xj, lam_j, hist


def erpa_cholesky(ap, am, factor=2.0):
    """(Y, eig, nneg) of ERPASYMM1 by the Cholesky route; `factor` is the
    reflection's 2 (a planted slip changes it)."""
    ap = 0.5 * (ap + ap.T)
    am = 0.5 * (am + am.T)
    n = ap.shape[0]
    w32, v32 = np.linalg.eigh(ap.astype(np.float32))
    nneg32 = int(np.sum(w32 < 0))
    k = min(n, ((nneg32 + 8 + 15) // 16) * 16)
    xj, lam_j, hist = partial_refine(ap, v32, w32, k)
    d = -factor * np.minimum(lam_j, 0.0)
    aabs = ap + (xj * d[None, :]) @ xj.T
    L = np.linalg.cholesky(aabs)
    m = L.T @ am @ L
    m = 0.5 * (m + m.T)
    om2, z = np.linalg.eigh(m)
    y = L @ z
    eig = om2.copy()
    pos = om2 > SMALL_OM2
    eig[pos] = np.sqrt(om2[pos])
    nrm = 2.0 / eig[pos] * np.einsum("ik,ij,jk->k", y[:, pos], am, y[:, pos])
    scale = np.where(nrm > 0.0, 1.0 / np.sqrt(np.where(nrm > 0.0, nrm, 1.0)), 0.0)
    y[:, pos] *= scale[None, :]
    return y, eig, int(np.sum(lam_j < 0.0)), hist


def _walpha(name, factor=2.0):
    import accas_oracle as O
    from gammcor_amd import synth
    fx = gold(name)
    s = synth.system_from_meta(fx["system"])
    sd = O.system_dict(s)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    indn = O.select_pairs(s.occ, s.nele)
    g = O.pair_integrals(tw, indn)
    out = []
    for a in fx["alphas"]:
        ap, am, _, ci = O.ab_cas(s.nbasis, s.ninact, s.nact, s.occ, s.xone, tw, s.gamma_act, s.igem, indn, a)
        y, eig, nneg, hist = erpa_cholesky(ap, am, factor)
        w, _ = O.aceneerpa(y, eig, tw, s.occ, ci, s.igem, indn, g)
        _, _, nneg_ref = O.erpasymm1(ap, am)
        out.append((w, nneg, nneg_ref, hist))
    return fx, out


@pytest.mark.parametrize("name", ["n24_cas66", "n48_cas44"])
def test_cholesky_route_matches_reference_walpha(name):
    fx, out = _walpha(name)
    assert any(o[2] > 0 for o in out), "the fixture's A+ should be indefinite somewhere"
    for (w, nneg, nneg_ref, hist), w_ref in zip(out, fx["walpha"]):
        assert nneg == nneg_ref
        assert abs(w - w_ref) < 1e-9, (w, w_ref)
        # linear convergence of the restricted refinement from the FP32 start
        assert hist[-1] < 1e-11 and len(hist) <= 4, hist


def test_cholesky_route_planted_factor_is_caught():
    """|A+| with lambda x x^T instead of 2 lambda x x^T: either not positive
    definite (the Cholesky fails) or a W(alpha) away from the reference."""
    fx = gold("n48_cas44")
    try:
        _, out = _walpha("n48_cas44", factor=1.0)
    except np.linalg.LinAlgError:
        return
    assert max(abs(o[0] - w) for o, w in zip(out, fx["walpha"])) > 1e-6

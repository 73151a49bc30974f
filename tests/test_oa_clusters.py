"""Cluster refinement of the mixed path's Ogita-Aishima steps, emulated in
numpy (the arithmetic of oa_pairs_kernel, oa_cluster_kernel, oa_e_kernel,
oa_cluster_solve_kernel and oa_cluster_apply_kernel in accas.hip).

The ERPA matrix M has modes the integrand drops (omega^2 <= 1e-6), kept
modes just above that threshold and near-degenerate groups.  From an FP32
eigendecomposition, the step's quadratic contraction (constant max|l| / gap)
does not reach such pairs, and 2x2 rotations from one S do not resolve groups
of three or more.  Pairs the step cannot contract are joined into clusters
whose projected matrices are solved exactly (Ogita and Aishima, Japan
J. Indust. Appl. Math. 36 (2019)).  The checks: the weighted projector
f(M) = sum_k w(omega_k^2) x_k x_k^T the integrand depends on (ACEneERPA,
interpa.f:1748-1776) against numpy's FP64 eigh, the convergence measure the
path stops on, and planted slips that must fail."""
import numpy as np
import pytest

EPS = 2.220446049250313e-16
JOIN = 1.0 / 64
RHO = 1e-7
CMAX = 64


def weight(l):
    """oa_weight, wmode 2: 1 / (2 omega) for a kept mode, 0 otherwise."""
    w = np.zeros_like(l)
    ok = l > 1e-6
    om = np.sqrt(np.where(ok, l, 1.0))
    ok &= (om > 1e-3) & (om < 1e8)
    w[ok] = 0.5 / om[ok]
    return w


def m_like(n, seed, nnull=40, ngroups=30):
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(0.0, 50.0, n)) ** 2 / 50.0
    lam[:nnull] = rng.uniform(-1e-9, 1e-9, nnull)          # dropped modes
    lam[nnull:nnull + 5] = 1e-6 * (1 + rng.uniform(0.5, 3, 5))   # kept, next to them
    for _ in range(ngroups):
        i = rng.integers(nnull + 10, n - 5)
        m = rng.integers(2, 5)
        lam[i:i + m] = lam[i] + lam[i] * rng.choice([1e-5, 1e-6, 1e-7, 1e-8], m) * np.arange(m)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * lam) @ q.T
    return 0.5 * (a + a.T)


def jacobi(h, tol, max_sweeps=30):
    """oa_cluster_solve_kernel's cyclic Jacobi: round-robin order (m/2
    disjoint rotations per round, position 0 fixed), H <- J^T H J, Q <- Q J,
    rotations with |h_pq| <= tol skipped; eigenvalues ascending (ties by
    position) with their columns."""
    h = h.copy()
    m = h.shape[0]
    q = np.eye(m)
    mp = m + (m & 1)
    for _ in range(max_sweeps):
        any_rot = False
        for rd in range(mp - 1):
            rots = []
            for t in range(mp // 2):
                a = 0 if t == 0 else 1 + (t - 1 + rd) % (mp - 1)
                c = 1 + (mp - 2 - t + rd) % (mp - 1)
                p, r = min(a, c), max(a, c)
                if r >= m or abs(h[p, r]) <= tol:
                    continue
                tau = (h[r, r] - h[p, p]) / (2.0 * h[p, r])
                tt = (1.0 if tau >= 0 else -1.0) / (abs(tau) + np.sqrt(1.0 + tau * tau))
                cs = 1.0 / np.sqrt(1.0 + tt * tt)
                rots.append((p, r, cs, tt * cs))
                any_rot = True
            for p, r, cs, sn in rots:          # rows
                hp, hq = h[p, :].copy(), h[r, :].copy()
                h[p, :], h[r, :] = cs * hp - sn * hq, sn * hp + cs * hq
            for p, r, cs, sn in rots:          # columns
                hp, hq = h[:, p].copy(), h[:, r].copy()
                h[:, p], h[:, r] = cs * hp - sn * hq, sn * hp + cs * hq
                qp, qq = q[:, p].copy(), q[:, r].copy()
                q[:, p], q[:, r] = cs * qp - sn * qq, sn * qp + cs * qq
        if not any_rot:
            break
    d = np.diag(h)
    order = np.lexsort((np.arange(m), d))
    return d[order], q[:, order]


def oa_step(a, x, mode="clusters", slip=None):
    """One step; returns (X, measure, clusters, largest cluster).  mode
    "pairs" is the 2x2-rotation step that clusters replace."""
    n = a.shape[0]
    s = x.T @ a @ x
    g = x.T @ x
    s = np.tril(s) + np.tril(s, -1).T       # the kernels read the lower triangle
    g = np.tril(g) + np.tril(g, -1).T
    r = np.eye(n) - g
    lam = np.diag(s) / np.diag(g)
    delta = 64 * EPS * np.max(np.abs(lam))
    mx = np.max(np.abs(lam))
    w = weight(lam)
    num = s + lam[None, :] * r
    nsym = np.maximum(np.abs(num), np.abs(s + lam[:, None] * r))
    den = lam[None, :] - lam[:, None]
    sep = np.abs(den) > np.maximum(8 * nsym, delta)
    np.fill_diagonal(sep, False)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(sep, num / np.where(sep, den, 1.0), 0.5 * r)
        wm = np.maximum(w[:, None], w[None, :])
        used = wm > 0
        rel = np.where(used, np.abs(w[:, None] - w[None, :]) / np.where(used, wm, 1.0), 0.0)
        kap = mx / np.abs(np.where(den == 0, 1.0, den))
    if mode == "pairs":
        pred = np.where(sep & used, e * e * kap * rel, 0.0)
        close = (~sep) & used
        np.fill_diagonal(close, False)
        emask = used.copy()
        np.fill_diagonal(emask, False)
        meas = max(np.max(np.abs(e[emask]), initial=0.0), 50 * np.max(pred, initial=0.0))
        xn = x + x @ e
        for i, j in zip(*np.nonzero(np.tril(close, -1))):
            lo, hi = j, i
            c = s[hi, lo] - 0.5 * (lam[lo] + lam[hi]) * g[hi, lo]
            if c == 0.0:
                continue
            tau = (lam[hi] - lam[lo]) / (2 * c)
            t = (1 if tau >= 0 else -1) / (abs(tau) + np.sqrt(1 + tau * tau))
            cs = 1 / np.sqrt(1 + t * t)
            u, v = xn[:, lo].copy(), xn[:, hi].copy()
            xn[:, lo], xn[:, hi] = cs * u - t * cs * v, t * cs * u + cs * v
        return xn, meas, 0, 0
    # oa_pairs_kernel: every close pair, and the used pairs mixed by at least
    # JOIN whose mix matters (|E| rel >= RHO)
    es = nsym / np.abs(np.where(den == 0, 1.0, den))
    with np.errstate(invalid="ignore"):
        join = (~sep) | (used & (es >= JOIN) & (es * rel >= RHO))
    np.fill_diagonal(join, False)
    pred = np.where(sep & used, np.abs(e) * np.minimum(1.0, np.abs(e) * kap) * rel, 0.0)
    # oa_cluster_kernel: connected components, members ascending
    par = np.arange(n)

    def find(i):
        while par[i] != i:
            par[i] = par[par[i]]
            i = par[i]
        return i
    for i, j in zip(*np.nonzero(np.tril(join, -1))):
        ra, rb = find(j), find(i)
        if ra != rb:
            par[max(ra, rb)] = min(ra, rb)
    roots = np.array([find(i) for i in range(n)])
    clusters = [np.nonzero(roots == rt)[0] for rt in np.unique(roots)]
    clusters = [c for c in clusters if len(c) > 1]
    incl = np.zeros((n, n), bool)
    for c in clusters:
        incl[np.ix_(c, c)] = True
    if slip != "double_count":
        e = np.where(incl, 0.5 * r, e)
    emask = used & ~incl
    np.fill_diagonal(emask, False)
    meas = max(np.max(np.abs(e[emask]), initial=0.0), 50 * np.max(np.where(incl, 0.0, pred), initial=0.0))
    big = max((len(c) for c in clusters), default=0)
    if 2 * big > n:
        meas = np.inf
    # oa_cluster_solve_kernel (from S, G and the step's eigenvalues), then
    # the update and oa_cluster_apply_kernel
    qs = []
    for c in clusters:
        h = s[np.ix_(c, c)] - 0.5 * (lam[c][:, None] + lam[c][None, :]) * g[np.ix_(c, c)]
        if slip == "no_metric":
            h = s[np.ix_(c, c)].copy()
        np.fill_diagonal(h, lam[c])
        # up to CMAX columns the workgroup Jacobi, above it the batched eigensolver
        _, qv = jacobi(h, 0.01 * delta / 64.0) if len(c) <= CMAX else np.linalg.eigh(h)
        qs.append(qv.T if slip == "transposed" else qv)
    xn = x + x @ e
    for c, qv in zip(clusters, qs):
        xn[:, c] = xn[:, c] @ qv
    return xn, meas, len(clusters), big


def projector(a, x):
    lr = np.einsum("ij,ik,kj->j", x, a, x) / np.einsum("ij,ij->j", x, x)
    xn = x / np.linalg.norm(x, axis=0)
    return (xn * weight(lr)) @ xn.T


def run(n, seed, mode, steps, slip=None):
    a = m_like(n, seed)
    _, v32 = np.linalg.eigh(a.astype(np.float32))
    x = v32.astype(np.float64)
    w64, v64 = np.linalg.eigh(a)
    f_ref = (v64 * weight(w64)) @ v64.T
    hist = []
    for _ in range(steps):
        x, meas, ncl, big = oa_step(a, x, mode, slip)
        hist.append(meas)
    err = np.max(np.abs(projector(a, x) - f_ref)) / np.max(np.abs(f_ref))
    return hist, err, big


def test_jacobi_round_robin_matches_eigh():
    rng = np.random.default_rng(0)
    for m in (2, 3, 7, 16, 33, 64):
        h = rng.standard_normal((m, m))
        h = 0.5 * (h + h.T)
        h[0, 0] = h[1, 1] + 1e-13          # a near tie
        d, q = jacobi(h, 1e-18 * np.abs(h).max())
        ref = np.linalg.eigvalsh(h)
        assert np.max(np.abs(d - ref)) < 1e-13 * np.abs(ref).max()
        assert np.max(np.abs(q.T @ q - np.eye(m))) < 1e-13
        assert np.max(np.abs(h @ q - q * d[None, :])) < 1e-12 * np.abs(ref).max()
        # every pair once per sweep, m/2 disjoint pairs per round
        mp = m + (m & 1)
        seen = set()
        for rd in range(mp - 1):
            used = []
            for t in range(mp // 2):
                a = 0 if t == 0 else 1 + (t - 1 + rd) % (mp - 1)
                c = 1 + (mp - 2 - t + rd) % (mp - 1)
                used += [a, c]
                seen.add((min(a, c), max(a, c)))
            assert sorted(used) == list(range(mp))
        assert len(seen) == mp * (mp - 1) // 2


@pytest.mark.parametrize("seed", [3, 4])
def test_clusters_converge_where_pair_rotations_do_not(seed):
    hist_p, err_p, _ = run(300, seed, "pairs", 4)
    hist_c, err_c, big = run(300, seed, "clusters", 4)
    # the path stops at a measure below 5e-4
    assert hist_c[-1] < 5e-4, hist_c
    assert hist_p[-1] > 5e-4, hist_p
    # the projector the integrand uses, against FP64 (whose own error for
    # gaps of 1e-8 relative is about 1e-16 max|l| / gap)
    assert err_c < 1e-8, err_c
    assert err_p > 10 * err_c


@pytest.mark.parametrize("slip", ["transposed", "double_count", "no_metric"])
def test_planted_slips_fail(slip):
    hist, err, _ = run(300, 3, "clusters", 4, slip)
    assert not (hist[-1] < 5e-4 and err < 1e-8), (slip, hist, err)

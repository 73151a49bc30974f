"""ORACLE — test infrastructure only.

CPU restatement of the reference's AC-CAS path (pernalk/GAMMCOR, in-core
variant), used by tests/, by ``__graft_entry__.smoke()`` and by bench.py's
cpu_baseline leg as the checker.  The product path (gammcor_amd) never imports
it.  Each function cites the reference lines it follows.

Pinning (what independently fixes each function; see tests/test_oracle.py):
  gauleg          numpy.polynomial.legendre.leggauss; compiled reference
  select_pairs    compiled reference (ACCAS end to end) and the NDimX closed form
  two_no1         closed form from the Cholesky factors; brute-force einsum at
                  small N; compiled reference TwoNO1 checksums
  ab_cas          compiled reference AB_CAS (full matrices at small N);
                  alpha-affinity identity
  erpasymm1       scipy eigh; omega^2 = eig(A+ A-); the matrix-function
                  identity sum_K Y_K Y_K^T = 1/2 S (S A- S)^(-1/2) S
  aceneerpa       compiled reference ACEInteg W(alpha) values
  acecorr         compiled reference ECorr (8-point grid and ACCAS 5-point)
  ac0cas          compiled reference AC0CAS ECorr (N = 12 .. 48); planted
                  slips (block triangle, block structure, X/Y) fail it

Deliberate difference: TwoNO1 screens |value| <= 1e-12 terms (misc.f:175,
212, 247, 281); the restatement does not.  The seeded integrals are O(1e-2),
so the screen changes no value by more than rounding.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SMALL_E = 1.0e-3      # acfd.f:4976
BIG_E = 1.0e8
SMALL_OM2 = 1.0e-6    # interpa.f:1663


# ---------------------------------------------------------------- GauLeg

def gauleg(x1, x2, n):
    """Gauss-Legendre nodes/weights on [x1, x2] by Newton iteration on P_n
    (misc.f:828-865, Numerical Recipes 4.5)."""
    x = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1) // 2
    xm = 0.5 * (x2 + x1)
    xl = 0.5 * (x2 - x1)
    for i in range(1, m + 1):
        z = np.cos(3.141592654 * (i - 0.25) / (n + 0.5))
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
            if abs(z - z1) <= 3.0e-14:
                break
        x[i - 1] = xm - xl * z
        x[n - i] = xm + xl * z
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return x, w


# ---------------------------------------------------------------- pairs

def select_pairs(occ, nele, thr_sel_act=1e-8, thr_qvirt=1e-7, iflcore=1):
    """ERPA pair list IndN (1-based, p > q) of accas.f:30-99."""
    nb = len(occ)
    indaux = np.where(np.arange(nb) < nele, 0, 2)
    for i in range(nb):
        if occ[i] < 1.0 and occ[i] != 0.0:
            indaux[i] = 1
    pairs = []
    for i in range(nb):
        for j in range(i):
            s = indaux[i] + indaux[j]
            if s == 0 or s == 4:
                continue
            if indaux[i] == 1 and indaux[j] == 1 and abs(occ[i] - occ[j]) / occ[i] < thr_sel_act:
                continue
            if iflcore == 1 or (occ[i] != 1.0 and occ[j] != 1.0):
                if abs(occ[i] + occ[j] - 2.0) > 1e-10 and abs(occ[i] + occ[j]) > thr_qvirt:
                    pairs.append((i + 1, j + 1))
    return np.array(pairs, dtype=np.int32).reshape(-1, 2)


# ---------------------------------------------------------------- transform

def _pair(i, j):
    a = np.maximum(i, j)
    b = np.minimum(i, j)
    return a * (a + 1) // 2 + b


def naddr3(p, q, r, s):
    """0-based NAddr3 (misc.f:505-525)."""
    return _pair(_pair(p, q), _pair(r, s))


def two_no1(ao_packed, u, chunk=1024):
    """(ab|cd) = sum U[a,m]U[b,n]U[c,l]U[d,s] (mn|ls), 8-fold packed in and
    out (misc.f:166-317: first/second index at 189-254, third/fourth at
    263-311).  Done as two half-transforms of the T x T pair matrix."""
    n = u.shape[0]
    t = n * (n + 1) // 2
    il, jl = np.tril_indices(n)
    # full T x T pair matrix from the packed triangle
    v = np.zeros((t, t))
    ii, jj = np.tril_indices(t)
    v[ii, jj] = ao_packed
    v[jj, ii] = ao_packed

    def half(m):
        out = np.empty_like(m)
        for c0 in range(0, t, chunk):
            c1 = min(t, c0 + chunk)
            blk = np.zeros((n, n, c1 - c0))
            blk[il, jl, :] = m[:, c0:c1]
            blk[jl, il, :] = m[:, c0:c1]
            x = np.einsum("am,mnc->anc", u, blk, optimize=True)
            y = np.einsum("anc,bn->abc", x, u, optimize=True)
            out[:, c0:c1] = y[il, jl, :]
        return out

    h = half(v)          # rows ab (NO), cols kl (AO)
    h = half(h.T)        # rows cd (NO), cols ab (NO)
    return h[ii, jj]     # packed (ab|cd), ab >= cd  (h symmetric)


# ---------------------------------------------------------------- AB_CAS

_lib = None


def _abcas_lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle_abcas.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-C", _HERE, "liboracle_abcas.so"], check=True,
                           stdout=subprocess.DEVNULL)
        _lib = ctypes.CDLL(path)
        d = ctypes.POINTER(ctypes.c_double)
        i = ctypes.POINTER(ctypes.c_int)
        _lib.oracle_ab_cas.argtypes = [ctypes.c_int] * 3 + [d, d, d, d, d, i, ctypes.c_int, i,
                                                             ctypes.c_double, d, d, d, d]
        _lib.oracle_ab_cas.restype = ctypes.c_int
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def ab_cas(nbasis, ninact, nact, occ, xone, twono, gamma_act, igem, indn, alpha, ure=None):
    """A+B, A-B of the alpha-Hamiltonian (interpa.f:3659-4104).  Returns
    (abplus, abmin, ecasscf, cicoef) with column-major (Fortran) matrices."""
    lib = _abcas_lib()
    n = len(indn)
    if ure is None:
        ure = np.eye(nbasis)
    ure = np.asfortranarray(ure, dtype=np.float64)
    occ = np.ascontiguousarray(occ, dtype=np.float64)
    xone = np.ascontiguousarray(xone, dtype=np.float64)
    twono = np.ascontiguousarray(twono, dtype=np.float64)
    g = np.ascontiguousarray(gamma_act, dtype=np.float64)
    ig = np.ascontiguousarray(igem, dtype=np.int32)
    ind = np.ascontiguousarray(indn, dtype=np.int32)
    ap = np.zeros((n, n), order="F")
    am = np.zeros((n, n), order="F")
    et = np.zeros(1)
    ci = np.zeros(nbasis)
    rc = lib.oracle_ab_cas(nbasis, ninact, nact, _dp(occ), _dp(xone), _dp(ure), _dp(twono), _dp(g),
                           _ip(ig), n, _ip(ind), float(alpha), _dp(ap), _dp(am), _dp(et), _dp(ci))
    if rc != 0:
        raise MemoryError("oracle_ab_cas failed")
    return ap, am, float(et[0]), ci


# ---------------------------------------------------------------- ERPASYMM1

def erpasymm1(abplus, abmin):
    """Symmetric ERPA solve of interpa.f:1652-1781.  Returns (Y, eig) where
    Y[:, K] are the normalised Y vectors and eig[K] = omega_K when
    omega_K^2 > 1e-6, else omega_K^2 (ascending order)."""
    ap = 0.5 * (abplus + abplus.T)            # 1680-1687
    am = 0.5 * (abmin + abmin.T)
    lam, v = np.linalg.eigh(ap)               # Diag8 -> DSYEVD (misc.f:541-616)
    s = (v * np.sqrt(np.abs(lam))[None, :]) @ v.T    # 1695-1716, |lambda| semantics
    m = s @ (am @ s)                          # 1725-1726
    m = 0.5 * (m + m.T)
    om2, z = np.linalg.eigh(m)                # 1729
    y = s @ z                                 # 1734-1740
    eig = om2.copy()
    pos = om2 > SMALL_OM2                     # 1748-1776
    eig[pos] = np.sqrt(om2[pos])
    nrm = 2.0 / eig[pos] * np.einsum("ik,ij,jk->k", y[:, pos], am, y[:, pos])
    scale = np.where(nrm > 0.0, 1.0 / np.sqrt(np.where(nrm > 0.0, nrm, 1.0)), 0.0)
    y[:, pos] *= scale[None, :]
    return y, eig, int(np.sum(lam < 0.0))


# ---------------------------------------------------------------- ACEneERPA

def pair_integrals(twono, indn):
    """G[I,J] = (p r|q s) for I=(p,r), J=(q,s) = IndN columns (acfd.f:5030)."""
    p = indn[:, 0] - 1
    r = indn[:, 1] - 1
    return twono[naddr3(p[:, None], r[:, None], p[None, :], r[None, :])]


def aceneerpa(y, eig, twono, occ, cicoef, igem, indn, g=None):
    """W(alpha) of acfd.f:4968-5043 (returned without the caller's weight)."""
    p = indn[:, 0] - 1
    r = indn[:, 1] - 1
    keep = (eig > SMALL_E) & (eig < BIG_E)
    yk = y[:, keep]
    pmat = yk @ yk.T                                           # SumY
    c = cicoef
    cij = c[p] + c[r]
    aux = 2.0 * cij[:, None] * cij[None, :] * pmat
    aux[np.diag_indices_from(aux)] -= occ[p] * (1.0 - occ[r]) + occ[r] * (1.0 - occ[p])
    same = (igem[p] == igem[r])[:, None] & (igem[p][:, None] == igem[p][None, :]) & \
        (igem[p] == igem[r])[None, :]
    if g is None:
        g = pair_integrals(twono, indn)
    return float(np.sum(np.where(same, 0.0, aux * g))), int(np.sum(~keep))


# ---------------------------------------------------------------- drivers

def aceinteg(sysd, twono, alpha, indn, g=None):
    """One alpha point (acfd.f:4832-4966): AB_CAS, ERPASYMM1, ACEneERPA."""
    ap, am, ecas, ci = ab_cas(sysd["nbasis"], sysd["ninact"], sysd["nact"], sysd["occ"], sysd["xone"],
                              twono, sysd["gamma_act"], sysd["igem"], indn, alpha)
    y, eig, nneg = erpasymm1(ap, am)
    w, nskip = aceneerpa(y, eig, twono, sysd["occ"], ci, sysd["igem"], indn, g)
    return dict(w=w, eig=eig, ecasscf=ecas, nneg_aplus=nneg, nskip=nskip)


def acecorr(sysd, twono, alphas, weights):
    """ECorr = sum_i w_i W(alpha_i) (acfd.f:161-201)."""
    indn = select_pairs(sysd["occ"], sysd["nele"])
    g = pair_integrals(twono, indn)
    res = [aceinteg(sysd, twono, a, indn, g) for a in alphas]
    ecorr = float(sum(wi * r["w"] for wi, r in zip(weights, res)))
    return dict(ecorr=ecorr, walpha=[r["w"] for r in res], ecasscf=res[0]["ecasscf"],
                eig=[r["eig"] for r in res], ndimx=len(indn),
                nneg_aplus=[r["nneg_aplus"] for r in res])


# ---------------------------------------------------------------- AC0CAS

def ac0_block_order(indn, nbasis, ninact, nact):
    """Blocks of the alpha=0 problem in AC0CAS's order (acfd.f:2122-2416):
    the active-active pairs (q outer, p inner), the active-inactive pairs of
    each inactive q, the virtual-active pairs of each virtual p, then every
    virtual-inactive pair on its own.  Returns (block label per pair,
    position of each pair in that order), both indexed like indn."""
    p = indn[:, 0] - 1
    q = indn[:, 1] - 1
    nocc = ninact + nact
    index = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(p, q))}
    order, label = [], np.full(len(indn), -1, dtype=np.int64)
    nb = 0

    def block(members):
        nonlocal nb
        members = [m for m in members if m is not None]
        if members:
            for m in members:
                order.append(m)
                label[m] = nb
            nb += 1
    block([index.get((ip, iq)) for iq in range(ninact, nocc) for ip in range(iq + 1, nocc)])
    for iq in range(ninact):
        block([index.get((ip, iq)) for ip in range(ninact, nocc)])
    for ip in range(nocc, nbasis):
        block([index.get((ip, iq)) for iq in range(ninact, nocc)])
    for ip in range(nocc, nbasis):
        for iq in range(ninact):
            block([index.get((ip, iq))])
    if len(order) != len(indn) or len(set(order)) != len(indn):
        raise ValueError("AC0CAS: a selected pair lies in none of the alpha=0 blocks")
    pos = np.empty(len(indn), dtype=np.int64)
    pos[np.asarray(order)] = np.arange(len(order))
    return label, pos


def lower_in_order(a, pos):
    """a[i, j] where pos[i] >= pos[j], mirrored: the triangle that AB0ELEMENT
    and AB1_CAS compute in block order (acfd.f:2168, 2244, 2323, 4200-4216)."""
    keep = pos[:, None] >= pos[None, :]
    return np.where(keep, a, a.T)


def erpasymm0(ap, am):
    """interpa.f:1496-1650: Y and X of the symmetric ERPA problem of one
    block.  eig = omega with Y normalised by 2 omega^-1 Y^T A- Y = 1 and
    X = omega^-1 A- Y when omega^2 > 1e-6; otherwise eig = omega^2 and
    Y = X = 0."""
    ap = 0.5 * (ap + ap.T)
    am = 0.5 * (am + am.T)
    lam, v = np.linalg.eigh(ap)
    s = (v * np.sqrt(np.abs(lam))[None, :]) @ v.T
    m = s @ (am @ s)
    m = 0.5 * (m + m.T)
    om2, z = np.linalg.eigh(m)
    y = s @ z
    x = np.zeros_like(y)
    eig = om2.copy()
    for k in range(len(om2)):
        if om2[k] > SMALL_OM2:
            eig[k] = np.sqrt(om2[k])
            nrm = 2.0 / eig[k] * (y[:, k] @ am @ y[:, k])
            y[:, k] *= 1.0 / np.sqrt(nrm) if nrm > 0.0 else 0.0
            x[:, k] = am @ y[:, k] / eig[k]
        else:
            y[:, k] = 0.0
    return y, x, eig


def ac0cas_fofo(sysd, twono, indn=None):
    """AC0CAS_FOFO (abfofo.f90:1485-1750), the FOFO form of the linearised
    AC correction, restated on its own terms.  It differs from AC0CAS in
    which parts of the (not exactly symmetric) A(0) and A' it uses:

    - A(0) = ABPM0_FOFO (abfofo.f90:1105-1320), the alpha=0 matrices over
      all pairs; each block (active-active, active-inactive per inactive
      orbital, virtual-active per virtual orbital; blocks of create_AC0_blocks
      1105, 1374-1483) goes to ERPASYMM0 whole, which symmetrises it
      (interpa.f:1525-1533) - AC0CAS uses the triangle at or below the
      diagonal in block order instead;
    - a virtual-inactive pair keeps omega = A+(0)_ii, X = Y = 1/sqrt(2)
      (abfofo.f90:1598-1613, 1662-1681);
    - A' = AB_CAS_FOFO(alpha=1, AB1) (abfofo.f90:1621-1623) is dA/dalpha of
      the affine A(alpha), symmetrised by sq_symmetrize (abfofo.f90:268-272,
      tran.f90:1303-1319): (A' + A'^T) / 2;
    - X^T A'+ X, Y^T A'- Y (abpm_tran, tran.f90:620-729), D = 2 (Tp - Tm) /
      (omega_m + omega_n) over nonzero omegas, P = Y D Y^T (ABPM_BACKTRAN,
      abfofo.f90:2152-2258);
    - ECorr = sum over pairs pq, rs of (C_p+C_q)(C_r+C_s) P_pq,rs (rs|pq)
      from the FOFO records, less the terms whose four orbitals share a
      geminal (abfofo.f90:1697-1735), the geminals fixed by index ranges
      (1 inactive, 2 active, 3 virtual; abfofo.f90:1543-1552)."""
    if indn is None:
        indn = select_pairs(sysd["occ"], sysd["nele"])
    args = (sysd["nbasis"], sysd["ninact"], sysd["nact"], sysd["occ"], sysd["xone"], twono, sysd["gamma_act"],
            sysd["igem"], indn)
    a0p, a0m, ecas, ci = ab_cas(*args, 0.0)
    a1p, a1m, _, _ = ab_cas(*args, 1.0)
    a0p, a0m, a1p, a1m = (np.asarray(x) for x in (a0p, a0m, a1p, a1m))
    label, _ = ac0_block_order(indn, sysd["nbasis"], sysd["ninact"], sysd["nact"])
    n = len(indn)
    sym = lambda a: 0.5 * (a + a.T)
    dp = sym(a1p - a0p)
    dm = sym(a1m - a0m)
    y = np.zeros((n, n))
    x = np.zeros((n, n))
    eig = np.zeros(n)
    ni, nocc = sysd["ninact"], sysd["ninact"] + sysd["nact"]
    p = indn[:, 0] - 1
    q = indn[:, 1] - 1
    for b in range(label.max() + 1):
        m = np.nonzero(label == b)[0]
        if len(m) == 1 and p[m[0]] >= nocc and q[m[0]] < ni:
            eig[m[0]] = a0p[m[0], m[0]]
            y[m[0], m[0]] = x[m[0], m[0]] = 1.0 / np.sqrt(2.0)
            continue
        yb, xb, eb = erpasymm0(a0p[np.ix_(m, m)], a0m[np.ix_(m, m)])
        y[np.ix_(m, m)] = yb
        x[np.ix_(m, m)] = xb
        eig[m] = eb
    tp = x.T @ dp @ x
    tm = y.T @ dm @ y
    nz = eig != 0.0
    den = eig[:, None] + eig[None, :]
    d = np.where(nz[:, None] & nz[None, :], 2.0 * (tp - tm) / np.where(den != 0.0, den, 1.0), 0.0)
    pm = y @ d @ y.T
    cij = ci[p] + ci[q]
    gem = np.where(np.arange(sysd["nbasis"]) < ni, 1, np.where(np.arange(sysd["nbasis"]) < nocc, 2, 3))
    same = (gem[p] == gem[q])[:, None] & (gem[p][:, None] == gem[p][None, :]) & (gem[p] == gem[q])[None, :]
    g = pair_integrals(twono, indn)
    ecorr = float(np.sum(np.where(same, 0.0, cij[:, None] * cij[None, :] * pm * g)))
    return dict(ecorr=ecorr, ecasscf=ecas, eig=eig, ndimx=n)


def ac0cas(sysd, twono, indn=None):
    """The linearised AC correction of AC0CAS (acfd.f:1902-2614).

    A(alpha) = A(0) + alpha A', so A' = A(1) - A(0) is AB1_CAS's matrix
    (acfd.f:3834).  The alpha=0 problem is solved block by block
    (ERPASYMM0; a virtual-inactive block takes omega = A+ and
    Y = X = 1/sqrt(2), acfd.f:2362-2416); then with the block eigenvectors
    P = Y D Y^T, D_mn = 2 (X^T A'+ X - Y^T A'- Y)_mn / (omega_m + omega_n)
    over nonzero omegas (acfd.f:2437-2556), and
    ECorr = sum_IJ (C_s+C_q)(C_p+C_r) P_IJ (pr|qs) without the terms whose
    four orbitals share a geminal (acfd.f:2557-2592)."""
    if indn is None:
        indn = select_pairs(sysd["occ"], sysd["nele"])
    args = (sysd["nbasis"], sysd["ninact"], sysd["nact"], sysd["occ"], sysd["xone"], twono, sysd["gamma_act"],
            sysd["igem"], indn)
    a0p, a0m, ecas, ci = ab_cas(*args, 0.0)
    a1p, a1m, _, _ = ab_cas(*args, 1.0)
    label, pos = ac0_block_order(indn, sysd["nbasis"], sysd["ninact"], sysd["nact"])
    n = len(indn)
    ap0 = lower_in_order(np.asarray(a0p), pos)
    am0 = lower_in_order(np.asarray(a0m), pos)
    dp = lower_in_order(np.asarray(a1p) - np.asarray(a0p), pos)
    dm = lower_in_order(np.asarray(a1m) - np.asarray(a0m), pos)
    y = np.zeros((n, n))
    x = np.zeros((n, n))
    eig = np.zeros(n)
    nocc = sysd["ninact"] + sysd["nact"]
    p = indn[:, 0] - 1
    q = indn[:, 1] - 1
    for b in range(label.max() + 1):
        m = np.nonzero(label == b)[0]
        if len(m) == 1 and p[m[0]] >= nocc and q[m[0]] < sysd["ninact"]:
            eig[m[0]] = ap0[m[0], m[0]]
            y[m[0], m[0]] = x[m[0], m[0]] = 1.0 / np.sqrt(2.0)
            continue
        yb, xb, eb = erpasymm0(ap0[np.ix_(m, m)], am0[np.ix_(m, m)])
        y[np.ix_(m, m)] = yb
        x[np.ix_(m, m)] = xb
        eig[m] = eb
    tp = x.T @ dp @ x
    tm = y.T @ dm @ y
    nz = eig != 0.0
    den = eig[:, None] + eig[None, :]
    d = np.where(nz[:, None] & nz[None, :], 2.0 * (tp - tm) / np.where(den != 0.0, den, 1.0), 0.0)
    pm = y @ d @ y.T
    c = ci
    cij = c[p] + c[q]
    igem = np.asarray(sysd["igem"])
    same = (igem[p] == igem[q])[:, None] & (igem[p][:, None] == igem[p][None, :]) & (igem[p] == igem[q])[None, :]
    g = pair_integrals(twono, indn)
    ecorr = float(np.sum(np.where(same, 0.0, cij[:, None] * cij[None, :] * pm * g)))
    return dict(ecorr=ecorr, ecasscf=ecas, eig=eig, ndimx=n)


def erpasymmxy(ap, am, cicoef, indn):
    """ERPASYMMXY (interpa.f:1401-1446): ERPASYMM0's Y and X turned into the
    X, Y of the original ERPA problem, X' = X / (C_p + C_q),
    Y' = Y / (C_p - C_q), EigX = (X' - Y') / 2, EigY = (X' + Y') / 2.
    Returns (EigY, EigX, Eig)."""
    y0, x0, eig = erpasymm0(ap, am)
    p = indn[:, 0] - 1
    q = indn[:, 1] - 1
    xs = x0 / (cicoef[p] + cicoef[q])[:, None]
    ys = y0 / (cicoef[p] - cicoef[q])[:, None]
    return 0.5 * (xs + ys), 0.5 * (xs - ys), eig


def sapt_response(sysd, twono, indn=None):
    """One monomer's response for SAPT (sapt_main.f90:2604-2660): A+-B at
    alpha = 1, ERPASYMMXY, and the response vectors
    EigVecR = (C_p - C_q)(EigY - EigX) with the omegas (what writeresp
    stores, 4068-4080)."""
    if indn is None:
        indn = select_pairs(sysd["occ"], sysd["nele"])
    ap, am, ecas, c = ab_cas(sysd["nbasis"], sysd["ninact"], sysd["nact"], sysd["occ"], sysd["xone"], twono,
                             sysd["gamma_act"], sysd["igem"], indn, 1.0)
    ey, ex, eig = erpasymmxy(np.asarray(ap), np.asarray(am), c, indn)
    p = indn[:, 0] - 1
    q = indn[:, 1] - 1
    evec = (c[p] - c[q])[:, None] * (ey - ex)
    return dict(evec=evec, eig=eig, cicoef=c, indn=indn, ecasscf=ecas)


def e2disp(ra, rb, vab):
    """The coupled second-order dispersion energy (e2disp,
    sapt_pol.f90:1684-2034; the loops of its ICASSCF = 0, ISERPA = 0 branch,
    1876-1907, are the same sum as the CAS branch's e2disp_intermed1):
      tmp1(i, rs) = sum_pq (C_q+C_p)^A (C_s+C_r)^B (pq|rs) EVecA(pq, i)
      tmp2 = tmp1 EVecB,
      E2disp = -16 sum_ij tmp2_ij^2 / (omegaA_i + omegaB_j)
    over 1e-3 < |omega| < 1e8 on both sides (1950-1966).  vab[p,q,r,s] holds
    (pq|rs)_AB over A's and B's orbitals."""
    pa, qa = ra["indn"][:, 0] - 1, ra["indn"][:, 1] - 1
    pb, qb = rb["indn"][:, 0] - 1, rb["indn"][:, 1] - 1
    ca, cb = ra["cicoef"], rb["cicoef"]
    v = vab[pa[:, None], qa[:, None], pb[None, :], qb[None, :]]
    fact = (ca[qa] + ca[pa])[:, None] * (cb[qb] + cb[pb])[None, :] * v
    tmp2 = (ra["evec"].T @ fact) @ rb["evec"]
    oa, ob = ra["eig"], rb["eig"]
    conda = (np.abs(oa) > 1e-3) & (np.abs(oa) < 1e8)
    condb = (np.abs(ob) > 1e-3) & (np.abs(ob) < 1e8)
    t = tmp2[np.ix_(conda, condb)]
    return -16.0 * float(np.sum(t * t / (oa[conda][:, None] + ob[condb][None, :])))


def system_dict(sysm):
    return dict(nbasis=sysm.nbasis, ninact=sysm.ninact, nact=sysm.nact, occ=sysm.occ, xone=sysm.xone,
                gamma_act=sysm.gamma_act, igem=sysm.igem, nele=sysm.nele)

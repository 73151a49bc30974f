"""The oracle (CPU restatement) pinned against the compiled reference's golden
fixtures and against what the mathematics fixes; planted mistakes must fail."""
import glob
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def gold(name):
    with open(os.path.join(GOLD, f"ref_{name}.json")) as f:
        return json.load(f)


def system_of(fx):
    from gammcor_amd import synth
    return synth.system_from_meta(fx["system"])


SMALL_GRID = ["n12_cas44", "n16_cas22", "n20_cas44_noinact", "n24_cas66", "n48_cas44", "n48_cas44_plainci"]
# the reference's edge branches: a discarded nearly degenerate active pair,
# an active occupation of exactly 1.0 and of 1 - 1e-12, omegas on both
# sides of the 1e-3 filter
EDGE_GRID = ["n12_cas44_degen", "n12_cas44_occ1", "n12_cas44_occ1eps", "n12_cas44_omega_edge"]


# ---------------------------------------------------------------- GauLeg
@pytest.mark.parametrize("n", [1, 2, 5, 8, 16, 30])
def test_gauleg_matches_legendre(n):
    import accas_oracle as O
    x, w = O.gauleg(0.0, 1.0, n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, 0.5 * (xr + 1.0), atol=1e-14)
    assert np.allclose(w, 0.5 * wr, atol=1e-14)
    # exact for polynomials of degree 2n-1
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-13


def test_gauleg_matches_reference_nodes():
    import accas_oracle as O
    from gammcor_amd import host
    fx = gold("n48_cas44_accas")
    x, w = O.gauleg(0.0, 1.0, 5)
    assert np.allclose(x, fx["accas_alphas"], atol=1e-15)
    xh, wh = host.GauLeg(0.0, 1.0, 5)
    assert np.allclose(xh, x, atol=1e-15) and np.allclose(wh, w, atol=1e-15)


# ---------------------------------------------------------------- pairs
@pytest.mark.parametrize("name", SMALL_GRID)
def test_pair_selection(name):
    import accas_oracle as O
    from gammcor_amd import host
    fx = gold(name)
    s = system_of(fx)
    indn = O.select_pairs(s.occ, s.nele)
    assert len(indn) == fx["ndimx"]
    ni, na = s.ninact, s.nact
    nv = s.nbasis - ni - na
    assert len(indn) == ni * na + ni * nv + na * (na - 1) // 2 + na * nv   # closed form, no degeneracies
    assert np.array_equal(host.ACCAS_pairs(s.occ, s.nele), indn)


# ---------------------------------------------------------------- TwoNO1
def _factors_closed_form(chol, u):
    n = u.shape[0]
    il, jl = np.tril_indices(n)
    lp = np.zeros((chol.shape[0], n, n))
    lp[:, il, jl] = chol
    lp[:, jl, il] = chol
    b = np.einsum("am,Pmn,bn->Pab", u, lp, u, optimize=True)[:, il, jl]
    v = b.T @ b
    i, j = np.tril_indices(v.shape[0])
    return v[i, j]


def test_two_no1_brute_force_and_closed_form():
    import accas_oracle as O
    from gammcor_amd import synth
    n = 7
    chol = synth.cholesky_factors(n, 9, 0.3, 11)
    ao = synth.ao_integrals_packed(chol)
    u = synth.orbitals(n, 12).T
    out = O.two_no1(ao, u)
    # brute force 4-index einsum
    full = np.zeros((n, n, n, n))
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    full[p, q, r, s] = ao[O.naddr3(p, q, r, s)]
    no = np.einsum("am,bn,cl,ds,mnls->abcd", u, u, u, u, full, optimize=True)
    for a in range(n):
        for b in range(n):
            for c in range(n):
                for d in range(n):
                    assert abs(out[O.naddr3(a, b, c, d)] - no[a, b, c, d]) < 1e-13
    assert np.max(np.abs(out - _factors_closed_form(chol, u))) < 1e-13
    # orthogonal U preserves the Frobenius norm of the full 4-index tensor
    assert abs(np.sum(no ** 2) - np.sum(full ** 2)) < 1e-10


@pytest.mark.parametrize("name", ["n12_cas44", "n48_cas44"])
def test_two_no1_matches_reference(name):
    import accas_oracle as O
    fx = gold(name)
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    assert abs(tw.sum() - fx["twono_sum"]) < 1e-9
    assert abs((tw ** 2).sum() - fx["twono_sumsq"]) < 1e-8
    for i, j, k, l, v in fx["twono"]:
        assert abs(tw[O.naddr3(i - 1, j - 1, k - 1, l - 1)] - v) < 1e-13


# ---------------------------------------------------------------- AB_CAS
def test_ab_cas_matches_reference_matrices():
    import accas_oracle as O
    fx = gold("n12_cas44_ab")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    indn = O.select_pairs(s.occ, s.nele)
    # the dumpab fixture evaluates AB_CAS at the node of the 1-point rule, 0.5
    ap, am, ecas, _ = O.ab_cas(s.nbasis, s.ninact, s.nact, s.occ, s.xone, tw, s.gamma_act, s.igem, indn, 0.5)
    n = fx["ndimx"]
    ref_p = np.array(fx["abplus"]).reshape(n, n, order="F")
    ref_m = np.array(fx["abmin"]).reshape(n, n, order="F")
    assert np.max(np.abs(ap - ref_p)) < 1e-11
    assert np.max(np.abs(am - ref_m)) < 1e-11
    assert abs(ecas - fx["ecasscf"]) < 1e-11


def test_ab_cas_is_affine_in_alpha():
    import accas_oracle as O
    fx = gold("n12_cas44")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    indn = O.select_pairs(s.occ, s.nele)
    args = (s.nbasis, s.ninact, s.nact, s.occ, s.xone, tw, s.gamma_act, s.igem, indn)
    a0 = O.ab_cas(*args, 0.0)
    a1 = O.ab_cas(*args, 1.0)
    for al in (0.3, 0.77):
        aa = O.ab_cas(*args, al)
        for k in (0, 1):
            assert np.max(np.abs(aa[k] - ((1 - al) * a0[k] + al * a1[k]))) < 1e-12


def test_hf_reference_reduces_to_rpa():
    """nA = 0 (closed-shell determinant) at alpha = 1: with F = h + G,
    G_pq = sum_k [2(pq|kk) - (pk|qk)], the two orientations of AB_CAS give
    A(ai,bj) = d_ij F_ab - d_ab F_ij + 2(ai|bj) - (ab|ij) and
    A(ai,jb) = -[2(ai|bj) - (aj|bi)], the textbook singlet RPA blocks A and -B;
    with C = (-0, 1) the C scaling is the identity, so ABPLUS = A - B and
    ABMIN = A + B."""
    import accas_oracle as O
    from gammcor_amd import synth
    n, ni = 8, 3
    s = synth.make_system(n, ni, 0, 0, 12, 0.2, 3)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    occ = np.zeros(n)
    occ[:ni] = 1.0
    g = np.where(np.arange(n) < ni, 1, 2).astype(np.int32)
    indn = O.select_pairs(occ, ni)
    assert len(indn) == ni * (n - ni)
    ap, am, _, _ = O.ab_cas(n, ni, 0, occ, s.xone, tw, np.zeros((1, 1, 1, 1)), g, indn, 1.0)
    h = synth.unpack_sym(s.xone, n)

    def I(p, q, r, t):
        return tw[O.naddr3(p, q, r, t)]
    F = h.copy()
    for p in range(n):
        for q in range(n):
            F[p, q] += sum(2 * I(p, q, k, k) - I(p, k, q, k) for k in range(ni))
    for x, (a, i) in enumerate(indn - 1):
        for y, (b, j) in enumerate(indn - 1):
            A = F[a, b] * (i == j) - F[i, j] * (a == b) + 2 * I(a, i, b, j) - I(a, b, i, j)
            B = 2 * I(a, i, b, j) - I(a, j, b, i)
            assert abs(ap[x, y] - (A - B)) < 1e-11
            assert abs(am[x, y] - (A + B)) < 1e-11


# ---------------------------------------------------------------- ERPASYMM1
def test_erpasymm1_matrix_function_identity():
    """sum_K Y_K Y_K^T = 1/2 S (S A- S)^(-1/2) S for PD A+/A- (all omega kept),
    and omega^2 = eig(A+ A-)."""
    import accas_oracle as O
    rng = np.random.default_rng(4)
    n = 40
    x = rng.standard_normal((n, n))
    ap = x @ x.T / n + np.eye(n)
    y = rng.standard_normal((n, n))
    am = y @ y.T / n + 0.5 * np.eye(n)
    Y, eig, nneg = O.erpasymm1(ap, am)
    assert nneg == 0
    lam, v = np.linalg.eigh(ap)
    S = (v * np.sqrt(lam)) @ v.T
    m = S @ am @ S
    w2, z = np.linalg.eigh(m)
    minv = (z / np.sqrt(np.sqrt(w2))) @ (z / np.sqrt(np.sqrt(w2))).T
    assert np.allclose(Y @ Y.T, 0.5 * S @ minv @ S, atol=1e-11)
    om2 = np.sort(np.linalg.eigvals(ap @ am).real)
    assert np.allclose(np.sort(eig ** 2), om2, rtol=1e-10)
    # normalisation 2/omega Y^T A- Y = 1
    nrm = 2.0 / eig * np.einsum("ik,ij,jk->k", Y, am, Y)
    assert np.allclose(nrm, 1.0, atol=1e-11)


# ---------------------------------------------------------------- end to end
@pytest.mark.parametrize("name", SMALL_GRID + EDGE_GRID)
def test_oracle_matches_reference_walpha(name):
    import accas_oracle as O
    fx = gold(name)
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    r = O.acecorr(O.system_dict(s), tw, fx["alphas"], fx["weights"])
    assert r["ndimx"] == fx["ndimx"]
    assert abs(r["ecasscf"] - fx["ecasscf"]) < 1e-10
    assert np.max(np.abs(np.array(r["walpha"]) - fx["walpha"])) < 1e-9
    assert abs(r["ecorr"] - fx["ecorr"]) < 1e-9
    for e_ref, e in zip(fx["eig"], r["eig"]):
        keep = (e > 1e-3) & (e < 1e8)
        assert int(keep.sum()) == e_ref["npos"]
        assert abs(e[keep].sum() - e_ref["sum_kept"]) < 1e-9
        assert abs(e[0] - e_ref["first"]) < 1e-10 and abs(e[-1] - e_ref["last"]) < 1e-9


def test_oracle_matches_reference_accas_end_to_end():
    """The reference's own ACCAS: its pair selection, 5-point GauLeg and the
    printed integrand values and AC-Corr line (F15.8)."""
    import accas_oracle as O
    fx = gold("n48_cas44_accas")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    x, w = O.gauleg(0.0, 1.0, 5)
    r = O.acecorr(O.system_dict(s), tw, x, w)
    assert np.max(np.abs(np.array(r["walpha"]) - fx["accas_walpha"])) < 1e-9
    ecas, ecorr, etot = fx["accas_line"]
    assert abs(r["ecorr"] - ecorr) < 1e-8 and abs(r["ecasscf"] - ecas) < 1e-8


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLD, "ref_n160_cas88.json")) or not os.environ.get("GMC_SLOW"),
                    reason="NBasis=160 oracle run takes minutes; set GMC_SLOW=1")
def test_oracle_matches_reference_n160():
    import accas_oracle as O
    fx = gold("n160_cas88")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    r = O.acecorr(O.system_dict(s), tw, fx["alphas"], fx["weights"])
    assert abs(r["ecorr"] - fx["ecorr"]) < 1e-9


def test_edge_cases_take_the_reference_branches():
    """What each edge fixture exercises, read from the reference's output:
    the degenerate pair is dropped (NDimX 49 of 50), exactly 1.0 and
    1 - 1e-12 give different ECorr (different branches, interpa.f:3926-4051),
    and the scaled system keeps its smallest omega at three alpha nodes and
    drops it at the fourth, where the reference stores omega^2 < 1e-6."""
    base = gold("n12_cas44")
    assert gold("n12_cas44_degen")["ndimx"] == base["ndimx"] - 1
    o1, o1e = gold("n12_cas44_occ1"), gold("n12_cas44_occ1eps")
    assert o1["ndimx"] == o1e["ndimx"]
    assert abs(o1["ecorr"] - o1e["ecorr"]) > 1e-3
    om = gold("n12_cas44_omega_edge")
    npos = [e["npos"] for e in om["eig"]]
    assert npos[:3] == [om["ndimx"]] * 3 and npos[3] == om["ndimx"] - 1
    assert 0.0 < om["eig"][3]["first"] < 1e-6
    assert all(1e-3 < e["first"] < 1.01e-3 for e in om["eig"][:3])


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """Plausible slips in the oracle (a dropped diagonal term, alpha applied to
    an intra-geminal term, a transposed integral index) must break the match
    with the reference fixture."""
    import accas_oracle as O
    fx = gold("n12_cas44")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    ref_w = np.array(fx["walpha"])

    def run():
        r = O.acecorr(O.system_dict(s), tw, fx["alphas"], fx["weights"])
        return np.max(np.abs(np.array(r["walpha"]) - ref_w))

    assert run() < 1e-9
    orig_aux = O.aceneerpa

    def no_diag(y, eig, twono, occ, cicoef, igem, indn, g=None):
        occ2 = np.zeros_like(occ)      # kills -n_p(1-n_s) - n_s(1-n_p)
        return orig_aux(y, eig, twono, occ2, cicoef, igem, indn, g)
    monkeypatch.setattr(O, "aceneerpa", no_diag)
    assert run() > 1e-6
    monkeypatch.setattr(O, "aceneerpa", orig_aux)
    orig_pair = O.pair_integrals
    monkeypatch.setattr(O, "pair_integrals",
                        lambda twono, indn: twono[O.naddr3(indn[:, 0][:, None] - 1, indn[:, 0][None, :] - 1,
                                                           indn[:, 1][:, None] - 1, indn[:, 1][None, :] - 1)])
    assert run() > 1e-6
    monkeypatch.setattr(O, "pair_integrals", orig_pair)
    orig_ab = O.ab_cas
    monkeypatch.setattr(O, "ab_cas", lambda *a, **k: orig_ab(*a[:-1], min(1.0, a[-1] * 1.001), **k))
    assert run() > 1e-6
    monkeypatch.setattr(O, "ab_cas", orig_ab)


# ---------------------------------------------------------------- AC0CAS
AC0 = ["n12_cas44_ac0", "n16_cas22_ac0", "n20_cas44_noinact_ac0", "n24_cas66_ac0", "n48_cas44_ac0"]


@pytest.mark.parametrize("name", AC0)
def test_ac0_matches_reference(name):
    """The linearised AC correction against the reference's own AC0CAS
    (acfd.f:1902) on the same system."""
    import accas_oracle as O
    fx = gold(name)
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    r = O.ac0cas(O.system_dict(s), tw)
    assert r["ndimx"] == fx["ndimx"]
    assert abs(r["ecasscf"] - fx["ecasscf"]) < 1e-10
    assert abs(r["ecorr"] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"]))


def test_ac0_planted_mistakes_are_caught(monkeypatch):
    """The other triangle of the alpha=0 blocks and of A', all pairs in one
    alpha=0 block, and X and Y of the blocks swapped must each break the
    match with the reference's AC0CAS.  (A(1) in place of the slope
    A(1) - A(0) would not: in the alpha=0 eigenbasis X^T A+(0) X and
    Y^T A-(0) Y are both diag(omega)/2, so the A(0) part cancels in D.)"""
    import accas_oracle as O
    fx = gold("n24_cas66_ac0")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    sd = O.system_dict(s)

    def err():
        return abs(O.ac0cas(sd, tw)["ecorr"] - fx["ac0_ecorr"])

    assert err() < 1e-10
    orig = O.lower_in_order
    monkeypatch.setattr(O, "lower_in_order", lambda a, pos: orig(a, -pos))
    assert err() > 1e-6
    monkeypatch.setattr(O, "lower_in_order", orig)
    blocks = O.ac0_block_order
    monkeypatch.setattr(O, "ac0_block_order", lambda *a: (np.zeros(len(a[0]), dtype=np.int64), blocks(*a)[1]))
    assert err() > 1e-6
    monkeypatch.setattr(O, "ac0_block_order", blocks)
    e0 = O.erpasymm0
    monkeypatch.setattr(O, "erpasymm0", lambda ap, am: (lambda y, x, e: (x, y, e))(*e0(ap, am)))
    assert err() > 1e-6
    monkeypatch.setattr(O, "erpasymm0", e0)


def test_ac0_alpha0_part_cancels(monkeypatch):
    """In the alpha=0 eigenbasis X^T A+(0) X = Y^T A-(0) Y = diag(omega)/2
    (ERPASYMM0's normalisation 2 Y^T X = 1 with X = omega^-1 A- Y), so the
    first-order term is blind to A(0) inside the slope: A(1) in place of
    A(1) - A(0) leaves AC0's ECorr unchanged.  With X off by a factor the
    identity breaks."""
    import accas_oracle as O
    fx = gold("n16_cas22_ac0")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    sd = O.system_dict(s)
    orig = O.ab_cas

    def a1_plus_a0(*a, **k):
        # alpha = 1 returns A(1) + A(0): the slope A(1) - A(0) becomes A(1)
        ap, am, e, c = orig(*a, **k)
        if a[-1] == 1.0:
            bp, bm, _, _ = orig(*a[:-1], 0.0, **k)
            return ap + bp, am + bm, e, c
        return ap, am, e, c

    def shift(monkeypatch):
        base = O.ac0cas(sd, tw)["ecorr"]
        monkeypatch.setattr(O, "ab_cas", a1_plus_a0)
        shifted = O.ac0cas(sd, tw)["ecorr"]
        monkeypatch.setattr(O, "ab_cas", orig)
        return abs(shifted - base) / max(1.0, abs(base))

    assert shift(monkeypatch) < 1e-10
    e0 = O.erpasymm0
    monkeypatch.setattr(O, "erpasymm0", lambda ap, am: (lambda y, x, e: (y, 2.0 * x, e))(*e0(ap, am)))
    assert shift(monkeypatch) > 1e-6
    monkeypatch.setattr(O, "erpasymm0", e0)


# ---------------------------------------------------------------- AC0CAS_FOFO
AC0FOFO = ["n12_cas44_ac0fofo", "n16_cas22_ac0fofo", "n20_cas44_noinact_ac0fofo", "n24_cas66_ac0fofo",
           "n48_cas44_ac0fofo"]


@pytest.mark.parametrize("name", AC0FOFO)
def test_ac0fofo_matches_reference(name):
    """The FOFO form of the linearised AC correction against the
    reference's own AC0CAS_FOFO (abfofo.f90:1485) on the FFOO/FOFO records
    of the same system."""
    import accas_oracle as O
    fx = gold(name)
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    r = O.ac0cas_fofo(O.system_dict(s), tw)
    assert r["ndimx"] == fx["ndimx"]
    assert abs(r["ecasscf"] - fx["ecasscf"]) < 1e-10
    assert abs(r["ecorr"] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"]))


def test_ac0fofo_differs_from_ac0cas_by_the_triangle():
    """AC0CAS_FOFO and AC0CAS are not two routes to one number: AC0CAS takes
    the triangle at or below the diagonal (in block order) of the alpha=0
    blocks and of A', AC0CAS_FOFO their symmetrised whole.  The two
    reference routines differ by more than 1e-3 Eh on every system, and
    each restatement matches only its own."""
    import accas_oracle as O
    for name in ("n12_cas44", "n24_cas66"):
        ff, fa = gold(name + "_ac0fofo"), gold(name + "_ac0")
        s = system_of(ff)
        tw = O.two_no1(s.twoel_ao, s.uaomo)
        sd = O.system_dict(s)
        rf, ra = O.ac0cas_fofo(sd, tw)["ecorr"], O.ac0cas(sd, tw)["ecorr"]
        assert abs(ff["ac0_ecorr"] - fa["ac0_ecorr"]) > 1e-3
        assert abs(rf - ff["ac0_ecorr"]) < 1e-10 * max(1.0, abs(rf))
        assert abs(ra - ff["ac0_ecorr"]) > 1e-3


def test_ac0fofo_planted_mistakes_are_caught(monkeypatch):
    """X and Y of the blocks swapped, all pairs in one alpha=0 block, and
    the unsymmetrised A' (its upper triangle ignored by sq_symmetrize's
    mirror) each break the match with the reference's AC0CAS_FOFO."""
    import accas_oracle as O
    fx = gold("n24_cas66_ac0fofo")
    s = system_of(fx)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    sd = O.system_dict(s)

    def err():
        return abs(O.ac0cas_fofo(sd, tw)["ecorr"] - fx["ac0_ecorr"])

    assert err() < 1e-10 * max(1.0, abs(fx["ac0_ecorr"]))
    e0 = O.erpasymm0
    monkeypatch.setattr(O, "erpasymm0", lambda ap, am: (lambda y, x, e: (x, y, e))(*e0(ap, am)))
    assert err() > 1e-6
    monkeypatch.setattr(O, "erpasymm0", e0)
    blocks = O.ac0_block_order
    monkeypatch.setattr(O, "ac0_block_order", lambda *a: (np.zeros(len(a[0]), dtype=np.int64), blocks(*a)[1]))
    assert err() > 1e-6
    monkeypatch.setattr(O, "ac0_block_order", blocks)
    orig = O.ab_cas

    def skewed(*a, **k):
        # alpha = 1 with its strict upper triangle doubled: a symmetrised A'
        # moves, the lower triangle alone would not
        ap, am, e, c = orig(*a, **k)
        if a[-1] == 1.0:
            ap = np.asarray(ap) + np.triu(np.asarray(ap), 1)
            am = np.asarray(am) + np.triu(np.asarray(am), 1)
        return ap, am, e, c

    monkeypatch.setattr(O, "ab_cas", skewed)
    assert err() > 1e-6


# ---------------------------------------------------------------- SAPT E2disp
DIMERS = ["d12_cas44_e2disp", "d24_cas66_e2disp"]


def _dimer(name):
    import accas_oracle as O
    from gammcor_amd import synth
    fx = gold(name)
    a, b, vab = synth.dimer(**fx["dimer"])
    ra = O.sapt_response(O.system_dict(a), O.two_no1(a.twoel_ao, a.uaomo))
    rb = O.sapt_response(O.system_dict(b), O.two_no1(b.twoel_ao, b.uaomo))
    return fx, ra, rb, vab


@pytest.mark.parametrize("name", DIMERS)
def test_e2disp_matches_reference(name):
    """The monomer responses (AB_CAS at alpha = 1, ERPASYMMXY) and the
    coupled E2disp against the reference's own chain and e2disp
    (sapt_pol.f90:1684) on the same seeded dimer."""
    import accas_oracle as O
    fx, ra, rb, vab = _dimer(name)
    assert len(ra["indn"]) == fx["ndimx_A"] and len(rb["indn"]) == fx["ndimx_B"]
    assert abs(ra["ecasscf"] - fx["ecasscf_A"]) < 1e-10 and abs(rb["ecasscf"] - fx["ecasscf_B"]) < 1e-10
    e = O.e2disp(ra, rb, vab)
    assert abs(e - fx["e2disp"]) < 1e-11 * max(1.0, abs(fx["e2disp"]))


def test_e2disp_response_is_erpasymm0_y():
    """(C_p - C_q)(EigY - EigX) of ERPASYMMXY is ERPASYMM0's Y: the
    rescaling by 1/(C_p +- C_q) cancels."""
    import accas_oracle as O
    from gammcor_amd import synth
    fx = gold("d12_cas44_e2disp")
    a, _, _ = synth.dimer(**fx["dimer"])
    sd = O.system_dict(a)
    tw = O.two_no1(a.twoel_ao, a.uaomo)
    r = O.sapt_response(sd, tw)
    ap, am, _, _ = O.ab_cas(sd["nbasis"], sd["ninact"], sd["nact"], sd["occ"], sd["xone"], tw, sd["gamma_act"],
                            sd["igem"], r["indn"], 1.0)
    y0, _, eig = O.erpasymm0(np.asarray(ap), np.asarray(am))
    assert np.max(np.abs(r["evec"] - y0)) < 1e-13
    assert np.array_equal(r["eig"], eig)


def test_e2disp_planted_mistakes_are_caught(monkeypatch):
    """Without the (C_p + C_q)(C_r + C_s) factors, with EigY in place of the
    response vectors, or with (pr|qs) in place of (pq|rs)_AB the result
    leaves the reference's."""
    import accas_oracle as O
    fx, ra, rb, vab = _dimer("d12_cas44_e2disp")
    ref = fx["e2disp"]
    assert abs(O.e2disp(ra, rb, vab) - ref) < 1e-11
    flat = dict(ra, cicoef=np.full_like(ra["cicoef"], 0.5))
    assert abs(O.e2disp(flat, dict(rb, cicoef=np.full_like(rb["cicoef"], 0.5)), vab) - ref) > 1e-4
    orig = O.erpasymmxy
    monkeypatch.setattr(O, "erpasymmxy", lambda *a: (lambda y, x, e: (y, 0.0 * x, e))(*orig(*a)))
    from gammcor_amd import synth
    a, b, _ = synth.dimer(**fx["dimer"])
    ra2 = O.sapt_response(O.system_dict(a), O.two_no1(a.twoel_ao, a.uaomo))
    rb2 = O.sapt_response(O.system_dict(b), O.two_no1(b.twoel_ao, b.uaomo))
    assert abs(O.e2disp(ra2, rb2, vab) - ref) > 1e-4
    monkeypatch.setattr(O, "erpasymmxy", orig)
    assert abs(O.e2disp(ra, rb, np.ascontiguousarray(vab.transpose(0, 2, 1, 3))) - ref) > 1e-4


# ---------------------------------------------------------------- rdm2.dat
def test_rdm2_dat_round_trip():
    from gammcor_amd import synth, host
    g, d = synth.cas_rdms(4, 4, 7)
    w, v = np.linalg.eigh(g)
    gam = 0.5 * np.einsum("pqrs,pa,qb,rc,sd->abcd", d, v, v, v, v)
    text = synth.rdm2_dat_text(gam)
    back = synth.read_rdm2_dat(text, 4)
    assert np.max(np.abs(back - gam)) <= 5.1e-13        # F19.12 of 2*Gamma
    packed = host.pack_rdm2act(back)
    na = 4
    for i in range(na):
        for j in range(na):
            for k in range(na):
                for l in range(na):
                    ij, kl = i * na + j + 1, k * na + l + 1
                    mx, mn = max(ij, kl), min(ij, kl)
                    assert packed[mx * (mx - 1) // 2 + mn - 1] == back[i, j, k, l]
    # Gamma = D/2 with sum_pq D_pqpq = ne(ne-1): the trace is ne(ne-1)/2
    assert abs(np.einsum("pqpq->", gam) - 4 * 3 / 2) < 1e-12


REF_DRIVER = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_driver")


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref/ref_driver not built (make -C oracle ref)")
@pytest.mark.parametrize("name", ["n16_cas22", "n20_cas44_noinact"])
def test_reference_fofo_path_on_the_record_layout(name):
    """The compiled reference's own ACEInteg_FOFO (abfofo.f90:3664), fed the
    FFOO and FOFO record files in the layout the GPU entry point takes
    (written by the reference driver's fofo mode from the transformed
    integrals), reproduces the in-core fixture: the record order of both
    classes is the reference's."""
    import subprocess
    import sys
    import tempfile
    sys.path.insert(0, GOLD)
    import make_golden as G
    from gammcor_amd import synth
    fx = json.load(open(os.path.join(GOLD, f"ref_{name}.json")))
    m = fx["system"]
    s = synth.make_system(m["nbasis"], m["ninact"], m["nact"], m["nelact"], m["rank"], m["sigma"], m["seed"])
    alphas, weights = synth.gauss_legendre_01(fx["nalpha"])
    with tempfile.TemporaryDirectory() as d:
        synth.write_inputs(s, d, alphas, weights)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {REF_DRIVER} fofo"], cwd=d, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        # the FFOO records hold (kl|ij) in the order of synth.fofo_records
        from accas_oracle import two_no1
        tw = two_no1(s.twoel_ao, s.uaomo)
        jrec, krec = synth.fofo_records(tw, s.nbasis, s.ninact + s.nact)
        jfile = np.fromfile(os.path.join(d, "FFOO"), dtype=np.float64)
        kfile = np.fromfile(os.path.join(d, "FOFO"), dtype=np.float64)
    res = G.parse(r.stdout)
    assert np.max(np.abs(np.array(res["walpha"]) - np.array(fx["walpha"]))) < 1e-11
    assert abs(res["ecorr"] - fx["ecorr"]) < 1e-11
    assert jfile.shape == jrec.shape and kfile.shape == krec.shape
    assert np.max(np.abs(jfile - jrec)) < 1e-12 * max(1.0, np.max(np.abs(jrec)))
    assert np.max(np.abs(kfile - krec)) < 1e-12 * max(1.0, np.max(np.abs(krec)))

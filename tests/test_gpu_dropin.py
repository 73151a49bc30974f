"""The host-pointer drop-ins (the reference's TwoNO1, AB_CAS, ERPASYMM1,
ACEneERPA, ACEInteg and ACECORR interfaces) against the compiled reference's
fixtures and the oracle, on the same seeded inputs."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))


def gold(name):
    with open(os.path.join(GOLD, f"ref_{name}.json")) as f:
        return json.load(f)


def _system(meta):
    from gammcor_amd import synth
    return synth.make_system(meta["nbasis"], meta["ninact"], meta["nact"], meta["nelact"], meta["rank"],
                             meta["sigma"], meta["seed"])


def _inputs(s):
    """NO integrals by the host TwoNO1 drop-in, plus the COMMON-side arrays."""
    from gammcor_amd import host
    tno = s.twoel_ao.copy()
    host.TwoNO1(tno, s.uaomo, s.nbasis)
    indn = host.ACCAS_pairs(s.occ, s.nele)
    igem = host.geminals(s.nbasis, s.ninact, s.nact)
    rdm = host.pack_rdm2act(s.gamma_act) if s.nact > 0 else np.zeros(1)
    return tno, indn, igem, rdm


def test_twono1_host_matches_oracle(cuda):
    import accas_oracle as O
    from gammcor_amd import host
    s = _system(dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0))
    print("drop-in TwoNO1 n=12", flush=True)
    tno = s.twoel_ao.copy()
    host.TwoNO1(tno, s.uaomo, s.nbasis)
    ref = O.two_no1(s.twoel_ao, s.uaomo)
    assert np.max(np.abs(tno - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_ab_cas_host_matches_reference_dump(cuda):
    from gammcor_amd import host
    fx = gold("n12_cas44_ab")
    s = _system(fx["system"])
    tno, indn, igem, rdm = _inputs(s)
    print(f"drop-in AB_CAS n12 ndimx={len(indn)}", flush=True)
    # the dump was taken at the node of the 1-point rule, alpha = 0.5
    ap, am, etot, ci = host.AB_CAS(tno, None, s.occ, s.xone, indn, s.nbasis, s.ninact, s.nact, igem, rdm, 0.5)
    nd = fx["ndimx"]
    rp = np.array(fx["abplus"]).reshape(nd, nd, order="F")
    rm = np.array(fx["abmin"]).reshape(nd, nd, order="F")
    scale = max(1.0, np.max(np.abs(rp)), np.max(np.abs(rm)))
    assert np.max(np.abs(ap - rp)) < 1e-11 * scale
    assert np.max(np.abs(am - rm)) < 1e-11 * scale
    assert abs(etot - fx["ecasscf"]) < 1e-10
    c = np.sqrt(s.occ) * np.where(s.occ < 0.5, -1.0, 1.0)      # CICoef (interpa.f:3777-3781)
    assert np.allclose(ci, c, rtol=0, atol=1e-14)


@pytest.mark.parametrize("name,alpha,occ_patch", [
    ("n24_cas66", 0.3, None),
    ("n48_cas44", 0.8, None),
    ("n20_cas44_noinact", 0.5, None),
    # an inactive orbital just below 1: its T1..T6 sums take the element loop
    ("n12_cas44", 0.5, (0, 1.0 - 1e-12)),
])
def test_ab_cas_host_matches_oracle(cuda, name, alpha, occ_patch):
    """The full A+B and A-B of the host AB_CAS against the oracle's C
    restatement (oracle/ab_cas.c) on larger active spaces than the reference
    dump: the active-space sums of T1..T6 come from the contracted tables,
    or from the element loop where an outer index is not active."""
    import accas_oracle as O
    from gammcor_amd import host, synth
    fx = gold(name)
    s = synth.system_from_meta(fx["system"])
    tno, indn, igem, rdm = _inputs(s)
    occ = s.occ.copy()
    if occ_patch is not None:
        occ[occ_patch[0]] = occ_patch[1]
    print(f"drop-in AB_CAS {name} ndimx={len(indn)} alpha={alpha}", flush=True)
    ap, am, etot, _ = host.AB_CAS(tno, None, occ, s.xone, indn, s.nbasis, s.ninact, s.nact, igem, rdm, alpha)
    rp, rm, e_ref, _ = O.ab_cas(s.nbasis, s.ninact, s.nact, occ, s.xone, tno, s.gamma_act, igem, indn, alpha)
    scale = max(1.0, np.max(np.abs(rp)), np.max(np.abs(rm)))
    assert np.max(np.abs(ap - rp)) < 1e-11 * scale, np.max(np.abs(ap - rp))
    assert np.max(np.abs(am - rm)) < 1e-11 * scale, np.max(np.abs(am - rm))
    assert abs(etot - e_ref) < 1e-10 * max(1.0, abs(e_ref))


@pytest.mark.parametrize("alpha", [0.0, 0.37, 1.0])
def test_erpasymm1_and_aceneerpa_host_match_oracle(cuda, alpha):
    import accas_oracle as O
    from gammcor_amd import host
    s = _system(dict(nbasis=16, ninact=3, nact=2, nelact=2, rank=30, sigma=0.1, seed=1))
    tno, indn, igem, rdm = _inputs(s)
    print(f"drop-in ERPASYMM1/ACEneERPA n16 alpha={alpha}", flush=True)
    ap, am, _, ci = O.ab_cas(s.nbasis, s.ninact, s.nact, s.occ, s.xone, tno, s.gamma_act, igem, indn, alpha)
    y_ref, e_ref, nneg_ref = O.erpasymm1(ap, am)
    y, e, sq, ams, nneg = host.ERPASYMM1(ap, am)
    scale = max(1.0, np.max(np.abs(e_ref)))
    assert np.max(np.abs(e - e_ref)) < 1e-10 * scale
    assert nneg == nneg_ref
    aps = 0.5 * (ap + ap.T)
    lam, v = np.linalg.eigh(aps)
    assert np.max(np.abs(sq - (v * np.sqrt(np.abs(lam))) @ v.T)) < 1e-10 * max(1.0, np.max(np.abs(sq)))
    assert np.array_equal(ams, 0.5 * (am + am.T)) or np.max(np.abs(ams - 0.5 * (am + am.T))) < 1e-15 * scale
    # Y is basis-dependent inside degenerate clusters; P = Yf Yf^T is not
    keep = (e > 1e-3) & (e < 1e8)
    p = y[:, keep] @ y[:, keep].T
    p_ref = y_ref[:, keep] @ y_ref[:, keep].T
    assert np.max(np.abs(p - p_ref)) < 1e-9 * max(1.0, np.max(np.abs(p_ref)))
    w_ref, _ = O.aceneerpa(y_ref, e_ref, tno, s.occ, ci, igem, indn)
    w = host.ACEneERPA(y_ref, e_ref, tno, s.occ, igem, ci, indn, s.nbasis)
    assert abs(w - w_ref) < 1e-10 * max(1.0, abs(w_ref))


@pytest.mark.parametrize("name", ["n16_cas22", "n24_cas66"])
def test_aceinteg_host_matches_reference(cuda, name):
    from gammcor_amd import host
    fx = gold(name)
    s = _system(fx["system"])
    tno, indn, igem, rdm = _inputs(s)
    print(f"drop-in ACEInteg {name} ndimx={len(indn)}", flush=True)
    for i, (a, w_ref) in enumerate(zip(fx["alphas"], fx["walpha"])):
        r = host.ACEInteg(tno, None, s.occ, s.xone, indn, s.nbasis, s.ninact, s.nact, igem, rdm, a,
                          side_outputs=(i == 0))
        assert abs(r["ecorr"] - w_ref) < 1e-9
        assert abs(r["egone"] - fx["ecasscf"]) < 1e-10
        if i == 0:
            e = r["eig"]
            keep = (e > 1e-3) & (e < 1e8)
            assert int(keep.sum()) == fx["eig"][0]["npos"]
            assert abs(e[keep].sum() - fx["eig"][0]["sum_kept"]) < 1e-9
            w = host.ACEneERPA(r["eigvecr"], e, tno, s.occ, igem,
                               np.sqrt(s.occ) * np.where(s.occ < 0.5, -1.0, 1.0), indn, s.nbasis)
            assert abs(w - w_ref) < 1e-9


@pytest.mark.parametrize("name", ["n16_cas22", "n20_cas44_noinact", "n24_cas66"])
def test_aceinteg_fofo_host_matches_reference(cuda, name):
    """ACEInteg_FOFO from the (FF|OO) and (FO|FO) record arrays: every W(alpha)
    and ECASSCF of the reference fixture (the reference's own FOFO path gives
    the in-core values to 1e-13 on these systems), and in FP64 mode the same
    ERPASYMM1 side outputs as the in-core drop-in."""
    from gammcor_amd import host, synth
    fx = gold(name)
    s = _system(fx["system"])
    tno, indn, igem, rdm = _inputs(s)
    jrec, krec = synth.fofo_records(tno, s.nbasis, s.ninact + s.nact)
    print(f"drop-in ACEInteg_FOFO {name} ndimx={len(indn)}", flush=True)
    for i, (a, w_ref) in enumerate(zip(fx["alphas"], fx["walpha"])):
        r = host.ACEInteg_FOFO(jrec, krec, None, s.occ, s.xone, indn, s.nbasis, s.ninact, s.nact, igem, rdm, a,
                               side_outputs=(i == 0))
        assert abs(r["ecorr"] - w_ref) < 1e-9
        assert abs(r["egone"] - fx["ecasscf"]) < 1e-10
        if i == 0:
            q = host.ACEInteg(tno, None, s.occ, s.xone, indn, s.nbasis, s.ninact, s.nact, igem, rdm, a,
                              side_outputs=True)
            assert np.max(np.abs(r["eig"] - q["eig"])) < 1e-10 * max(1.0, np.max(np.abs(q["eig"])))
            assert np.max(np.abs(r["abplus"] - q["abplus"])) < 1e-10 * max(1.0, np.max(np.abs(q["abplus"])))
            assert np.max(np.abs(r["abmin"] - q["abmin"])) < 1e-10 * max(1.0, np.max(np.abs(q["abmin"])))


def test_acecorr_host_matches_reference_and_tracks_new_integrals(cuda):
    from gammcor_amd import host
    fx = gold("n48_cas44")
    s = _system(fx["system"])
    tno, indn, igem, rdm = _inputs(s)
    print(f"drop-in ACECORR n48 ndimx={len(indn)}", flush=True)
    args = (tno, None, s.occ, s.xone, indn, s.nbasis, s.ninact, s.nact, igem, rdm, fx["alphas"], fx["weights"])
    r = host.ACECORR_host(*args)
    assert abs(r["ecorr"] - fx["ecorr"]) < 1e-9
    assert np.max(np.abs(r["walpha"] - np.array(fx["walpha"]))) < 1e-9
    m = host.ACECORR_host(*args, prec=host.GMC_MIXED)
    assert abs(m["ecorr"] - fx["ecorr"]) < 1e-6
    # same host buffer, different contents: the resident copy must follow
    tno2 = tno.copy()
    tno[:] = 0.5 * tno2
    half = host.ACECORR_host(*args)
    tno[:] = tno2
    again = host.ACECORR_host(*args)
    assert abs(again["ecorr"] - fx["ecorr"]) < 1e-9
    assert abs(half["ecorr"] - fx["ecorr"]) > 1e-3

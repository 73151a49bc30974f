"""Mixed-precision AC-CAS (FP32 eigen solves + FP64 Ogita-Aishima steps)
against the FP64 GPU path and the oracle; alpha-batch invariance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,ni,na,ne,rank,seed,nal", [(12, 2, 4, 4, 36, 0, 4), (48, 6, 4, 4, 144, 0, 8)])
def test_mixed_vs_fp64(cuda, n, ni, na, ne, rank, seed, nal):
    from gammcor_amd import host, synth
    import accas_oracle as O
    s = synth.make_system(n, ni, na, ne, rank, 0.1, seed)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    a, w = host.GauLeg(0.0, 1.0, nal)
    d = host.DeviceSystem(tw, s.occ, s.xone, s.gamma_act, n, ni, na, s.nele)
    ref = host.ACECORR(d, a, w, prec=host.GMC_FP64, return_eig=True)
    one = host.ACECORR(d, a, w, prec=host.GMC_FP64, alpha_batch=1)
    assert abs(one["ecorr"] - ref["ecorr"]) < 1e-10
    m1 = host.ACECORR(d, a, w, prec=host.GMC_MIXED, return_eig=True)
    m2 = host.ACECORR(d, a, w, prec=host.GMC_MIXED2, return_eig=True)
    f32 = host.ACECORR(d, a, w, prec=host.GMC_FP32)
    e1, e2, e32 = (abs(x["ecorr"] - ref["ecorr"]) for x in (m1, m2, f32))
    print(f"n={n}: |dE| fp32 {e32:.2e}  mixed(3 OA) {e1:.2e}  mixed(2 OA) {e2:.2e}", flush=True)
    assert e1 < 1e-6 and e2 < 1e-6
    de = np.max(np.abs(m1["eig"].cpu().numpy() - ref["eig"].cpu().numpy()))
    assert de < 1e-9


def test_unconverged_problems_are_solved_again_in_fp64(cuda, monkeypatch, capfd):
    """With the convergence threshold at zero no refinement converges: after
    six steps every problem is solved again in FP64, and the mixed result is
    the FP64 path's.  The default (Cholesky) route refines only A+'s
    eigenpairs below zero, so only M's problems are solved again; the eigen
    route (GMC_MIXED_ROUTE=eigen) refines A+ in full and solves both again.
    At the default threshold the same system converges in three steps and
    nothing is solved again."""
    from gammcor_amd import host, synth
    import accas_oracle as O
    s = synth.make_system(48, 6, 4, 4, 144, 0.1, 0)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    a, w = host.GauLeg(0.0, 1.0, 8)
    d = host.DeviceSystem(tw, s.occ, s.xone, s.gamma_act, 48, 6, 4, s.nele)
    ref = host.ACECORR(d, a, w, prec=host.GMC_FP64, return_eig=True)
    monkeypatch.setenv("GMC_OA_STATS", "1")
    capfd.readouterr()
    m = host.ACECORR(d, a, w, prec=host.GMC_MIXED, return_eig=True)
    err = capfd.readouterr().err
    assert "solved again in FP64" not in err
    assert abs(m["ecorr"] - ref["ecorr"]) < 1e-6
    monkeypatch.setenv("GMC_OA_EXTEND", "0")
    f = host.ACECORR(d, a, w, prec=host.GMC_MIXED, return_eig=True)
    err = capfd.readouterr().err
    print(err, flush=True)
    assert "OA A+" not in err and "Cholesky route: restricted step" in err
    assert "[gmc] OA M: 8 of 8 problems solved again in FP64" in err
    assert err.count("[gmc] OA M step") == 6
    assert abs(f["ecorr"] - ref["ecorr"]) < 1e-10
    assert np.max(np.abs(f["eig"].cpu().numpy() - ref["eig"].cpu().numpy())) < 1e-10
    monkeypatch.setenv("GMC_MIXED_ROUTE", "eigen")
    f = host.ACECORR(d, a, w, prec=host.GMC_MIXED, return_eig=True)
    err = capfd.readouterr().err
    print(err, flush=True)
    assert "[gmc] OA A+: 8 of 8 problems solved again in FP64" in err
    assert "[gmc] OA M: 8 of 8 problems solved again in FP64" in err
    assert err.count("[gmc] OA M step") == 6
    assert abs(f["ecorr"] - ref["ecorr"]) < 1e-10
    assert np.max(np.abs(f["eig"].cpu().numpy() - ref["eig"].cpu().numpy())) < 1e-10


@pytest.mark.parametrize("system", ["n48", "n160"])
def test_cholesky_route_matches_eigen_route(cuda, monkeypatch, system):
    """The Cholesky route (the default: |A+| = L L^T from the eigenpairs of
    A+ below zero) and the reference's eigen route (S = |A+|^(1/2),
    GMC_MIXED_ROUTE=eigen) give the same W(alpha), within 1 uEh of the FP64
    path, and the same count of negative eigenvalues of A+, which is not
    zero on these systems."""
    from gammcor_amd import host, _lib
    if system == "n48":
        d, a, w = _n48_system()
    else:
        from gammcor_amd import synth
        s = synth.config_system("n160_cas88", with_integrals=False)
        import torch
        ao = synth.ao_integrals_device(s.chol, cuda)
        ure = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(cuda)
        _lib.call("gmc_twono1_dev", ao.data_ptr(), ure.data_ptr(), s.nbasis, ao.numel(), None)
        d = host.DeviceSystem(ao, s.occ, s.xone, s.gamma_act, s.nbasis, s.ninact, s.nact, s.nele, device=cuda)
        a, w = host.GauLeg(0.0, 1.0, 8)
    ref = host.ACECORR(d, a, w, prec=host.GMC_FP64)
    _lib.call("gmc_prof_reset")
    _lib.call("gmc_prof_enable", 1)
    try:
        c = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
        routes_c = (_lib.prof_get("mixed_chol_route")[1], _lib.prof_get("mixed_eigen_route")[1])
        monkeypatch.setenv("GMC_MIXED_ROUTE", "eigen")
        _lib.call("gmc_prof_reset")
        e = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
        routes_e = (_lib.prof_get("mixed_chol_route")[1], _lib.prof_get("mixed_eigen_route")[1])
    finally:
        _lib.call("gmc_prof_enable", 0)
    print(f"{system}: |dE| Cholesky route {abs(c['ecorr'] - ref['ecorr']):.2e}, eigen route "
          f"{abs(e['ecorr'] - ref['ecorr']):.2e}; nneg {list(c['nneg'])} / {list(ref['nneg'])}", flush=True)
    assert routes_c == (1, 0) and routes_e == (0, 1)
    assert np.all(np.asarray(ref["nneg"]) > 0)
    assert np.array_equal(c["nneg"], ref["nneg"]) and np.array_equal(e["nneg"], ref["nneg"])
    assert abs(c["ecorr"] - ref["ecorr"]) < 1e-6 and abs(e["ecorr"] - ref["ecorr"]) < 1e-6
    assert np.max(np.abs(np.asarray(c["walpha"]) - np.asarray(ref["walpha"]))) < 1e-6


def _n48_system():
    from gammcor_amd import host, synth
    import accas_oracle as O
    s = synth.make_system(48, 6, 4, 4, 144, 0.1, 0)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    a, w = host.GauLeg(0.0, 1.0, 8)
    return host.DeviceSystem(tw, s.occ, s.xone, s.gamma_act, 48, 6, 4, s.nele), a, w


def test_clusters_through_the_batched_eigensolver(cuda, monkeypatch, capfd):
    """The refinement's clusters above the workgroup Jacobi's size go to the
    batched FP64 eigensolver (N=400's first step has one of about 1 900
    columns per problem).  GMC_OA_CMAX=2 sends every cluster of three or more
    columns there: the result stays within 1 uEh of the FP64 path, and
    within 1e-9 Eh of the default run, whose clusters the Jacobi solves."""
    import re
    from gammcor_amd import host, synth, _lib
    s = synth.config_system("n160_cas88", with_integrals=False)
    import torch
    ao = synth.ao_integrals_device(s.chol, cuda)
    ure = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(cuda)
    _lib.call("gmc_twono1_dev", ao.data_ptr(), ure.data_ptr(), s.nbasis, ao.numel(), None)
    d = host.DeviceSystem(ao, s.occ, s.xone, s.gamma_act, s.nbasis, s.ninact, s.nact, s.nele, device=cuda)
    a, w = host.GauLeg(0.0, 1.0, 8)
    ref = host.ACECORR(d, a, w, prec=host.GMC_FP64)
    monkeypatch.setenv("GMC_OA_STATS", "1")
    capfd.readouterr()
    m0 = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
    err0 = capfd.readouterr().err
    monkeypatch.setenv("GMC_OA_CMAX", "2")
    m1 = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
    err1 = capfd.readouterr().err
    print(err0, err1, flush=True)
    big0 = [int(x) for x in re.findall(r"OA M step \d+: \d+ pairs in \d+ clusters \(largest \d+, (\d+) to", err0)]
    big1 = [int(x) for x in re.findall(r"OA M step \d+: \d+ pairs in \d+ clusters \(largest \d+, (\d+) to", err1)]
    largest = [int(x) for x in re.findall(r"OA M step \d+: \d+ pairs in \d+ clusters \(largest (\d+)", err1)]
    assert big0 and max(largest) >= 3 and sum(big1) > sum(big0), (largest, big0, big1)
    assert "solved again in FP64" not in err1
    assert abs(m1["ecorr"] - ref["ecorr"]) < 1e-6
    assert abs(m1["ecorr"] - m0["ecorr"]) < 1e-9


def test_half_precision_updates_on_every_step(cuda, monkeypatch, capfd):
    """GMC_OA_HALF at 1e-12 forms every refinement update from plain half
    inputs (the branch larger systems take for their first corrections):
    the refinement still meets 1 uEh against the FP64 path."""
    from gammcor_amd import host
    d, a, w = _n48_system()
    ref = host.ACECORR(d, a, w, prec=host.GMC_FP64)
    monkeypatch.setenv("GMC_OA_HALF", "1e-12")
    monkeypatch.setenv("GMC_OA_STATS", "1")
    capfd.readouterr()
    m = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
    err = capfd.readouterr().err
    print(err, flush=True)
    assert abs(m["ecorr"] - ref["ecorr"]) < 1e-6


def test_close_pairs_over_the_cap_go_to_fp64(cuda, monkeypatch, capfd):
    """With no close pair recorded (GMC_OA_MAXPAIRS=0) a step that finds one
    cannot count as converged: the refinement goes on, and a refinement whose
    last step still found pairs solves its problems again in FP64.  The mixed
    result stays within 1 uEh of the FP64 path."""
    from gammcor_amd import host
    d, a, w = _n48_system()
    ref = host.ACECORR(d, a, w, prec=host.GMC_FP64)
    monkeypatch.setenv("GMC_OA_MAXPAIRS", "0")
    monkeypatch.setenv("GMC_OA_STATS", "1")
    capfd.readouterr()
    m = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
    err = capfd.readouterr().err
    print(err, flush=True)
    assert abs(m["ecorr"] - ref["ecorr"]) < 1e-6
    # a step that found close pairs is not the last one unless its problems
    # are solved again
    import re
    steps = [(what, int(npair)) for what, npair in re.findall(r"OA (A\+|M) step \d+: (\d+) pairs in", err)]
    for what in ("A+", "M"):
        seq = [npair for w_, npair in steps if w_ == what]
        if seq and seq[-1] > 0:
            assert f"[gmc] OA {what}:" in err and "solved again in FP64" in err


def test_fixed_step_count_fails_loudly_over_the_cap(cuda, monkeypatch):
    """A fixed step count (GMC_MIXED2) has no FP64 re-solve: a problem left
    with unrecorded close pairs fails the call instead of returning a value
    its refinement did not converge to."""
    from gammcor_amd import host, _lib
    d, a, w = _n48_system()
    monkeypatch.setenv("GMC_OA_MAXPAIRS", "0")
    monkeypatch.setenv("GMC_OA_STATS", "1")
    try:
        host.ACECORR(d, a, w, prec=host.GMC_MIXED2)
    except _lib.GmcError as e:
        assert "close pairs" in str(e)

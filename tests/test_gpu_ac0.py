"""The linearised AC correction (AC0CAS) on the GPU against the reference's
own AC0CAS (golden fixtures written by the compiled reference on the same
seeded systems) and against the oracle's block-by-block restatement."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["n12_cas44_ac0", "n16_cas22_ac0", "n20_cas44_noinact_ac0", "n24_cas66_ac0", "n48_cas44_ac0"]


def gold(name):
    with open(os.path.join(GOLD, f"ref_{name}.json")) as f:
        return json.load(f)


def _device_system(fx, cuda):
    """Seeded inputs; AO integrals transformed on the GPU by TwoNO1."""
    import torch
    from gammcor_amd import host, synth
    m = fx["system"]
    s = synth.make_system(m["nbasis"], m["ninact"], m["nact"], m["nelact"], m["rank"], m["sigma"], m["seed"])
    tno = torch.from_numpy(s.twoel_ao).to(cuda)
    ure = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(cuda)
    host.TwoNO1(tno, ure, s.nbasis)
    return s, host.DeviceSystem(tno, s.occ, s.xone, s.gamma_act, s.nbasis, s.ninact, s.nact, s.nele)


@pytest.mark.parametrize("name", CASES)
def test_ac0_matches_reference(cuda, name):
    """ECorr of AC0CAS to 1e-10 relative (1e-10 Eh below 1 Eh), ECASSCF to
    1e-10 Eh; the N=20 case has no inactive orbitals (no active-inactive or
    virtual-inactive blocks)."""
    from gammcor_amd import host
    fx = gold(name)
    s, d = _device_system(fx, cuda)
    print(f"AC0 {name}: ndimx={d.ndimx}", flush=True)
    assert d.ndimx == fx["ndimx"]
    r = host.AC0CAS(d)
    assert abs(r["ecasscf"] - fx["ecasscf"]) < 1e-10
    assert abs(r["ecorr"] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"])), (r["ecorr"], fx["ac0_ecorr"])


def test_ac0_matches_oracle_and_repeats(cuda):
    """A system outside the fixtures against the oracle, and the same call
    twice (the workspaces of the first serve the second; the eigensolver sums
    with LDS atomics, so a repeat agrees to rounding, not bit for bit)."""
    from gammcor_amd import host, synth
    import accas_oracle as O
    s = synth.make_system(32, 4, 6, 6, 96, 0.1, 11)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    ref = O.ac0cas(O.system_dict(s), tw)
    d = host.DeviceSystem(tw, s.occ, s.xone, s.gamma_act, 32, 4, 6, s.nele)
    print(f"AC0 N=32 CAS(6,6): ndimx={d.ndimx}", flush=True)
    a = host.AC0CAS(d)
    b = host.AC0CAS(d)
    assert abs(a["ecorr"] - ref["ecorr"]) < 1e-10 * max(1.0, abs(ref["ecorr"]))
    assert abs(a["ecorr"] - b["ecorr"]) < 1e-12 * max(1.0, abs(a["ecorr"]))


FOFO_CASES = ["n12_cas44_ac0fofo", "n16_cas22_ac0fofo", "n20_cas44_noinact_ac0fofo", "n24_cas66_ac0fofo",
              "n48_cas44_ac0fofo"]


@pytest.mark.parametrize("name", FOFO_CASES)
def test_ac0fofo_matches_reference(cuda, name):
    """AC0CAS_FOFO on the device (the (FF|OO) and (FO|FO) classes built from
    the AO integrals) against the reference's own AC0CAS_FOFO on the same
    system, ECorr to 1e-10 relative; it differs from AC0CAS by 1e-3 Eh or
    more on these systems."""
    import torch
    from gammcor_amd import host, synth
    fx = gold(name)
    m = fx["system"]
    s = synth.make_system(m["nbasis"], m["ninact"], m["nact"], m["nelact"], m["rank"], m["sigma"], m["seed"])
    ao = torch.from_numpy(s.twoel_ao).to(cuda)
    u = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(cuda)
    dummy = torch.zeros(1, dtype=torch.float64, device=cuda)
    d = host.DeviceSystem(dummy, s.occ, s.xone, s.gamma_act, s.nbasis, s.ninact, s.nact, s.nele)
    print(f"AC0CAS_FOFO {name}: ndimx={d.ndimx}", flush=True)
    r = host.AC0CAS_FOFO(ao, u, d)
    assert abs(r["ecasscf"] - fx["ecasscf"]) < 1e-10
    assert abs(r["ecorr"] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"])), (r["ecorr"], fx["ac0_ecorr"])
    other = gold(name.replace("_ac0fofo", "_ac0"))["ac0_ecorr"]
    assert abs(r["ecorr"] - other) > 1e-3


def test_ac0fofo_host_drop_in_from_records(cuda):
    """The host-pointer entry with the two classes in the reference's FFOO /
    FOFO record order (gmc_ac0cas_fofo, the Fortran AC0CAS_FOFO's target)."""
    import ctypes
    from gammcor_amd import host, synth, _lib
    import accas_oracle as O
    fx = gold("n24_cas66_ac0fofo")
    m = fx["system"]
    s = synth.make_system(m["nbasis"], m["ninact"], m["nact"], m["nelact"], m["rank"], m["sigma"], m["seed"])
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    jrec, krec = synth.fofo_records(tw, s.nbasis, s.ninact + s.nact)
    indn = host.ACCAS_pairs(s.occ, s.nele)
    ure = np.asfortranarray(np.eye(s.nbasis))
    rdm2 = host.pack_rdm2act(s.gamma_act)
    igem = np.ascontiguousarray(s.igem, dtype=np.int32)
    ec, et = np.zeros(1), np.zeros(1)
    ind = np.ascontiguousarray(indn.reshape(-1), dtype=np.int32)
    _lib.call("gmc_ac0cas_fofo", ec.ctypes.data, et.ctypes.data, jrec.ctypes.data, krec.ctypes.data, ure.ctypes.data,
              np.ascontiguousarray(s.occ).ctypes.data, np.ascontiguousarray(s.xone).ctypes.data, igem.ctypes.data,
              rdm2.ctypes.data, ind.ctypes.data, len(indn), s.nbasis, s.ninact, s.nact)
    assert abs(et[0] - fx["ecasscf"]) < 1e-10
    assert abs(ec[0] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"]))

"""An AC-CAS solve from the reference's input files: the seeded NBasis=12
system written as Molpro files (AOTWOINT.mol, MOLPRO.MOPUN with CASORB,
AOONEINT.mol with ONEHAMIL, the 1RDM/2RDM records), read back by the
library's readers, transformed and solved on the GPU in FP64: ECorr and
W(alpha) within 1e-9 Eh of the compiled reference's fixture for the same
system."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_solve_from_molpro_files(cuda, tmp_path):
    import torch
    from gammcor_amd import formats, host, synth
    with open(os.path.join(GOLD, "ref_n12_cas44.json")) as f:
        fx = json.load(f)
    m = fx["system"]
    s = synth.make_system(m["nbasis"], m["ninact"], m["nact"], m["nelact"], m["rank"], m["sigma"], m["seed"])
    n, ni, na = s.nbasis, s.ninact, s.nact
    cmo = s.uaomo.T.copy()
    h_ao = cmo @ synth.unpack_sym(s.xone, n) @ cmo.T
    formats.write_aotwoint_molpro(tmp_path / "AOTWOINT.mol", s.twoel_ao, [n])
    formats.write_mo_molpro(tmp_path / "MOLPRO.MOPUN", cmo, [n])
    formats.write_oneint_molpro(tmp_path / "AOONEINT.mol", {"ONEHAMIL": 0.5 * (h_ao + h_ao.T)}, [n])
    formats.write_rdm_molpro(tmp_path / "RDM.mol", s.gamma_act, s.occ[ni:ni + na])

    # the inputs of the solve, from the files only
    ao = host.read_aotwoint_molpro(tmp_path / "AOTWOINT.mol", n)
    c = host.read_mo_molpro(tmp_path / "MOLPRO.MOPUN", n)
    h = synth.unpack_sym(host.read_oneint_molpro(tmp_path / "AOONEINT.mol", n), n)
    xone = synth.pack_sym(c.T @ h @ c)
    rdm2 = host.read_rdm2_molpro(tmp_path / "RDM.mol", na)
    d1 = host.read_rdm1_molpro(tmp_path / "RDM.mol", na)
    gam = np.zeros((na,) * 4)
    for i in range(na):
        for j in range(na):
            for k in range(na):
                for l_ in range(na):
                    ij, kl = i * na + j + 1, k * na + l_ + 1
                    a, b = max(ij, kl), min(ij, kl)
                    gam[i, j, k, l_] = rdm2[a * (a - 1) // 2 + b - 1]
    occ = np.zeros(n)
    occ[:ni] = 1.0
    occ[ni:ni + na] = d1[[j * (j + 1) // 2 + j for j in range(na)]]
    assert np.array_equal(occ, s.occ)

    tno = torch.from_numpy(ao).to(cuda)
    ure = torch.from_numpy(np.asfortranarray(c.T).ravel(order="F")).to(cuda)
    host.TwoNO1(tno, ure, n)
    d = host.DeviceSystem(tno, occ, xone, gam, n, ni, na, s.nele)
    r = host.ACECORR(d, fx["alphas"], fx["weights"], prec=host.GMC_FP64)
    print(f"from files: ECorr {r['ecorr']:.12f} reference {fx['ecorr']:.12f}", flush=True)
    assert np.max(np.abs(r["walpha"] - np.array(fx["walpha"]))) < 1e-9
    assert abs(r["ecorr"] - fx["ecorr"]) < 1e-9
    assert abs(r["ecasscf"] - fx["ecasscf"]) < 1e-9

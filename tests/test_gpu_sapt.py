"""SAPT's second-order dispersion on the GPU: each monomer's response
(AB_CAS at alpha = 1 and the ERPA solve, gmc_sapt_response_dev) and the
coupled E2disp contraction (gmc_e2disp_dev) against the compiled
reference's own SAPT chain and e2disp (sapt_pol.f90:1684) on the same
seeded dimers, to 1e-10 relative."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["d12_cas44_e2disp", "d24_cas66_e2disp"])
def test_e2disp_matches_reference(cuda, name):
    import torch
    from gammcor_amd import host, synth
    with open(os.path.join(GOLD, f"ref_{name}.json")) as f:
        fx = json.load(f)
    a, b, vab = synth.dimer(**fx["dimer"])
    resp = []
    for s in (a, b):
        tno = torch.from_numpy(s.twoel_ao).to(cuda)
        ure = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(cuda)
        host.TwoNO1(tno, ure, s.nbasis)
        d = host.DeviceSystem(tno, s.occ, s.xone, s.gamma_act, s.nbasis, s.ninact, s.nact, s.nele)
        resp.append((d, host.SAPT_RESPONSE(d)))
    (da, ra), (db, rb) = resp
    print(f"SAPT {name}: ndimx {da.ndimx} x {db.ndimx}", flush=True)
    assert da.ndimx == fx["ndimx_A"] and db.ndimx == fx["ndimx_B"]
    pa, qa = da.indn_host[:, 0] - 1, da.indn_host[:, 1] - 1
    pb, qb = db.indn_host[:, 0] - 1, db.indn_host[:, 1] - 1
    v = vab[pa[:, None], qa[:, None], pb[None, :], qb[None, :]]
    vdev = torch.from_numpy(np.asfortranarray(v).ravel(order="F")).to(cuda)
    e = host.E2DISP(ra, rb, vdev)
    print(f"SAPT {name}: E2disp {e:.12f} reference {fx['e2disp']:.12f}", flush=True)
    assert abs(e - fx["e2disp"]) < 1e-10 * max(1.0, abs(fx["e2disp"]))
    # the response columns of modes with omega^2 <= 1e-6 are zero, as ERPASYMM0 leaves them
    for r in (ra, rb):
        n = r["ndimx"]
        ev = r["evec"].cpu().numpy().reshape(n, n, order="F")
        small = r["eig"].cpu().numpy() <= 1e-6
        assert np.all(ev[:, small] == 0.0)

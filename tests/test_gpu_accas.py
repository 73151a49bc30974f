"""AC-CAS end to end on the GPU against the oracle (same seeded inputs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, ni, na, ne, rank, seed):
    from gammcor_amd import synth
    import accas_oracle as O
    s = synth.make_system(n, ni, na, ne, rank, 0.1, seed)
    tw = O.two_no1(s.twoel_ao, s.uaomo)
    return s, tw


@pytest.mark.parametrize("n,ni,na,ne,rank,seed,nal", [(12, 2, 4, 4, 36, 0, 4), (16, 3, 2, 2, 30, 1, 5),
                                                      (20, 0, 4, 4, 50, 2, 3), (48, 6, 4, 4, 144, 0, 8)])
def test_acecorr_fp64_vs_oracle(cuda, n, ni, na, ne, rank, seed, nal):
    import accas_oracle as O
    from gammcor_amd import host
    s, tw = _case(n, ni, na, ne, rank, seed)
    a, w = host.GauLeg(0.0, 1.0, nal)
    ref = O.acecorr(O.system_dict(s), tw, a, w)
    d = host.DeviceSystem(tw, s.occ, s.xone, s.gamma_act, n, ni, na, s.nele)
    print(f"acecorr n={n} ndimx={d.ndimx} nalpha={nal}", flush=True)
    got = host.ACECORR(d, a, w, return_eig=True)
    assert d.ndimx == ref["ndimx"]
    assert abs(got["ecasscf"] - ref["ecasscf"]) < 1e-10
    assert np.max(np.abs(got["walpha"] - np.array(ref["walpha"]))) < 1e-9, (got["walpha"], ref["walpha"])
    assert abs(got["ecorr"] - ref["ecorr"]) < 1e-9
    eig = got["eig"].cpu().numpy()
    for i in range(nal):
        assert np.max(np.abs(eig[i] - ref["eig"][i])) < 1e-9
    assert list(got["nneg"]) == ref["nneg_aplus"]

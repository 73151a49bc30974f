"""The batching workload: 64 independent AC-CAS systems of uneven size
(NBasis 40-140, CAS(4e,4o)/CAS(6e,6o) alternating, 16 alpha points each).
Each system is transformed and solved on the GPU with its alpha points
batched, systems one after another; the mixed path must stay within 1 uEh of
the FP64 path on every system, and the FP64 path matches the oracle on the
smallest systems."""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))


def _device_system(cfg, cuda):
    import torch
    from gammcor_amd import host, synth, _lib
    s = synth.make_system(with_integrals=False, **cfg)
    ao = synth.ao_integrals_device(s.chol, cuda)
    ure = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(cuda)
    _lib.call("gmc_twono1_dev", ao.data_ptr(), ure.data_ptr(), s.nbasis, ao.numel(), None)
    d = host.DeviceSystem(ao, s.occ, s.xone, s.gamma_act, s.nbasis, s.ninact, s.nact, s.nele, device=cuda)
    return s, d


def test_batch_workload_mixed_within_1ueh_of_fp64(cuda):
    import torch
    from gammcor_amd import host, synth
    cfgs = synth.batch_configs()
    a, w = host.GauLeg(0.0, 1.0, synth.BATCH["nalpha"])
    worst = 0.0
    t0 = time.perf_counter()
    for k, cfg in enumerate(cfgs):
        print(f"batch system {k}: NBasis={cfg['nbasis']} CAS({cfg['nelact']},{cfg['nact']})", flush=True)
        s, d = _device_system(cfg, cuda)
        ref = host.ACECORR(d, a, w, prec=host.GMC_FP64)
        mix = host.ACECORR(d, a, w, prec=host.GMC_MIXED)
        de = abs(mix["ecorr"] - ref["ecorr"])
        worst = max(worst, de)
        assert np.isfinite(ref["ecorr"]) and ref["ecorr"] < 0.0
        assert de < 1e-6, (k, cfg, de)
        del d
        torch.cuda.empty_cache()
    print(f"64 systems in {time.perf_counter() - t0:.1f} s, max |dEcorr| mixed vs FP64 {worst:.2e}", flush=True)


@pytest.mark.parametrize("k", [0, 1])
def test_batch_smallest_systems_match_oracle(cuda, k):
    """FP64 GPU against the oracle (numpy/LAPACK restatement of ACECORR) on the
    two smallest systems of the workload."""
    import accas_oracle as O
    from gammcor_amd import host, synth
    cfgs = synth.batch_configs()
    order = np.argsort([c["nbasis"] for c in cfgs], kind="stable")
    cfg = cfgs[int(order[k])]
    print(f"batch system NBasis={cfg['nbasis']} vs oracle", flush=True)
    s, d = _device_system(cfg, cuda)
    a, w = host.GauLeg(0.0, 1.0, synth.BATCH["nalpha"])
    gpu = host.ACECORR(d, a, w, prec=host.GMC_FP64)
    so = synth.make_system(**cfg)
    tw = O.two_no1(so.twoel_ao, so.uaomo)
    ref = O.acecorr(O.system_dict(so), tw, a, w)
    assert abs(gpu["ecorr"] - ref["ecorr"]) < 1e-9
    assert np.max(np.abs(np.asarray(gpu["walpha"]) - np.asarray(ref["walpha"]))) < 1e-9

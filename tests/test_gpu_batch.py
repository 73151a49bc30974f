"""The batching workload: 64 independent AC-CAS systems of uneven size
(NBasis 40-140, CAS(4e,4o)/CAS(6e,6o) alternating, 16 alpha points each).

- Per system (each system's alpha points batched): the mixed path stays
  within 1 uEh of the FP64 path on every system, and the FP64 path matches
  the oracle on the smallest systems.
- The batching form (gmc_acecorr_batch_dev: all systems x alpha points in
  padded groups of up to 16 systems per launch sequence): FP64 within 1e-9 Eh
  of the per-system FP64 results on every system, mixed within 1 uEh.
- Its HIP-graph replay: captured on the first call and kept, then replayed
  alone by the next calls, with eager batch solves and synchronisations in
  between, every replay within 1e-9 Eh of the eager result; after a mixed
  batch solve (whose workspaces may grow, which forces a new capture) the
  same again.
"""
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


@pytest.fixture(scope="module")
def workload(cuda):
    """The 64 systems on the device (NO integrals transformed) and their
    per-system FP64 and mixed results."""
    from gammcor_amd import host, synth
    cfgs = synth.batch_configs()
    a, w = host.GauLeg(0.0, 1.0, synth.BATCH["nalpha"])
    systems, fp64, mixed = [], [], []
    t0 = time.perf_counter()
    for k, cfg in enumerate(cfgs):
        print(f"batch system {k}: NBasis={cfg['nbasis']} CAS({cfg['nelact']},{cfg['nact']})", flush=True)
        _, d = _device_system(cfg, cuda)
        systems.append(d)
        fp64.append(host.ACECORR(d, a, w, prec=host.GMC_FP64)["ecorr"])
        mixed.append(host.ACECORR(d, a, w, prec=host.GMC_MIXED)["ecorr"])
    print(f"64 systems one by one in {time.perf_counter() - t0:.1f} s", flush=True)
    return dict(systems=systems, fp64=np.array(fp64), mixed=np.array(mixed), alphas=a, weights=w)


def test_batch_workload_mixed_within_1ueh_of_fp64(workload):
    fp64, mixed = workload["fp64"], workload["mixed"]
    assert np.all(np.isfinite(fp64)) and np.all(fp64 < 0.0)
    de = np.abs(mixed - fp64)
    print(f"max |dEcorr| mixed vs FP64 {de.max():.2e}", flush=True)
    assert np.all(de < 1e-6), (np.argmax(de), de.max())


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


def test_batch_form_fp64_matches_per_system(workload):
    from gammcor_amd import host
    nd = [d.ndimx for d in workload["systems"]]
    grp, npad = host.batch_plan(nd, len(workload["alphas"]))
    print(f"batch form: {grp.max() + 1} groups, padded orders {sorted(set(npad.tolist()), reverse=True)}", flush=True)
    t0 = time.perf_counter()
    r = host.ACECORR_BATCH(workload["systems"], workload["alphas"], workload["weights"], prec=host.GMC_FP64)
    print(f"batch form FP64: {time.perf_counter() - t0:.2f} s", flush=True)
    de = np.abs(r["ecorr"] - workload["fp64"])
    print(f"batch form FP64 vs per system: max |dEcorr| {de.max():.2e}", flush=True)
    assert np.all(de < 1e-9), (np.argmax(de), de.max())


def test_batch_form_mixed_within_1ueh(workload, monkeypatch):
    """Every group on the mixed path (GMC_BATCH_MIXED_MIN=0: by default
    groups below order 3000, all of this workload's, run in FP64)."""
    from gammcor_amd import host
    monkeypatch.setenv("GMC_BATCH_MIXED_MIN", "0")
    t0 = time.perf_counter()
    r = host.ACECORR_BATCH(workload["systems"], workload["alphas"], workload["weights"], prec=host.GMC_MIXED)
    print(f"batch form mixed: {time.perf_counter() - t0:.2f} s", flush=True)
    de = np.abs(r["ecorr"] - workload["fp64"])
    k = int(np.argmax(de))
    print(f"batch form mixed vs per-system FP64: max |dEcorr| {de.max():.2e} (system {k}; its per-system mixed "
          f"{abs(workload['mixed'][k] - workload['fp64'][k]):.2e})", flush=True)
    assert np.all(de < 1e-6), (k, de.max())


def test_batch_form_graph_replays(workload):
    from gammcor_amd import host
    args = (workload["systems"], workload["alphas"], workload["weights"])
    eager = host.ACECORR_BATCH(*args, prec=host.GMC_FP64)["ecorr"]
    for rep in range(4):
        if rep == 2:
            os.environ["GMC_BATCH_MIXED_MIN"] = "0"
            try:
                host.ACECORR_BATCH(*args, prec=host.GMC_MIXED)
            finally:
                del os.environ["GMC_BATCH_MIXED_MIN"]
        t0 = time.perf_counter()
        g = host.ACECORR_BATCH(*args, prec=host.GMC_FP64, graph=True)["ecorr"]
        print(f"graph call {rep}: {time.perf_counter() - t0:.2f} s (capture and replay, or replay alone), "
              f"max |d| vs eager {np.abs(g - eager).max():.2e}", flush=True)
        assert np.all(np.abs(g - eager) < 1e-9)
        e2 = host.ACECORR_BATCH(*args, prec=host.GMC_FP64)["ecorr"]
        assert np.all(np.abs(e2 - eager) < 1e-9)

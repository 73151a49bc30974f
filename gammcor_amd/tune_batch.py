"""Autotuner of the batching form (gmc_acecorr_batch_dev).

The batching form cuts the systems, sorted by ERPA order, into groups whose
systems x alpha points run as one batch padded to the group's largest order;
the cut minimises a cost model of one batch,

    t(npad, B) = c0 + c1 npad + c2 B npad^2 + c3 B npad^3   (ms),

c0 the fixed cost of a batch, c1 the per-column latency of its
tridiagonalisations (one column after another), c2 the streaming work per
problem (the panel SYMV, formation, copies), c3 the MFMA work.  This script measures t on the GPU for batches of B = g x 16 problems
(g copies of one seeded system, 16 alpha points, forced into one batch) over
a grid of orders and group sizes, fits c by least squares and writes
gammcor_amd/tuning/batch_gfx950.json (the measurements and the fit), which
host.ACECORR_BATCH passes to the library.

    python -m gammcor_amd.tune_batch [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "tuning", "batch_gfx950.json")


def measure(nbasis_list, groups, nalpha=16, reps=2):
    import torch
    from . import host, synth, _lib
    dev = torch.device("cuda:0")
    a, w = host.GauLeg(0.0, 1.0, nalpha)
    rows = []
    one_batch = np.array([1e9, 0.0, 0.0, 0.0])     # a single group whenever it fits
    for nb in nbasis_list:
        cfg = dict(nbasis=nb, ninact=nb // 12, nact=4, nelact=4, rank=3 * nb, sigma=0.1, seed=2000 + nb)
        s = synth.make_system(with_integrals=False, **cfg)
        ao = synth.ao_integrals_device(s.chol, dev)
        ure = torch.from_numpy(np.asfortranarray(s.uaomo).ravel(order="F")).to(dev)
        _lib.call("gmc_twono1_dev", ao.data_ptr(), ure.data_ptr(), nb, ao.numel(), None)
        d = host.DeviceSystem(ao, s.occ, s.xone, s.gamma_act, nb, s.ninact, s.nact, s.nele, device=dev)
        for g in groups:
            systems = [d] * g
            host.ACECORR_BATCH(systems, a, w, prec=host.GMC_FP64, cost=one_batch)   # sizes every workspace
            times = []
            for _ in range(reps):
                _lib.call("gmc_prof_reset")
                _lib.call("gmc_prof_enable", 1)
                host.ACECORR_BATCH(systems, a, w, prec=host.GMC_FP64, cost=one_batch)
                torch.cuda.synchronize()
                _lib.call("gmc_prof_enable", 0)
                ms, cnt, _ = _lib.prof_get("batch_group_f64")
                assert cnt == 1, cnt
                times.append(ms)
            npad = int(host.batch_plan([d.ndimx] * g, nalpha, one_batch)[1][0])
            rows.append(dict(nbasis=nb, ndimx=d.ndimx, npad=npad, systems=g, problems=g * nalpha,
                             ms=float(min(times))))
            print(f"npad {npad:5d}  B {g * nalpha:4d}  {min(times):9.2f} ms", flush=True)
        del d, ao
        _lib.call("gmc_finalize")
        torch.cuda.empty_cache()
    return rows


def fit(rows):
    A = np.array([[1.0, r["npad"], r["problems"] * float(r["npad"]) ** 2, r["problems"] * float(r["npad"]) ** 3]
                  for r in rows])
    t = np.array([r["ms"] for r in rows])
    # relative least squares: every measurement weighs alike
    c, *_ = np.linalg.lstsq(A / t[:, None], np.ones_like(t), rcond=None)
    c = np.maximum(c, 0.0)
    pred = A @ c
    return c, float(np.max(np.abs(pred - t) / t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="a smaller grid")
    ap.add_argument("--out", default=OUT, help="where the measurements and the fit go")
    args = ap.parse_args()
    nbs = [40, 80, 140] if args.quick else [40, 60, 80, 100, 120, 140]
    groups = [1, 4, 16] if args.quick else [1, 2, 4, 8, 16]
    t0 = time.perf_counter()
    rows = measure(nbs, groups)
    c, err = fit(rows)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    res = dict(model="t_ms = c0 + c1 * npad + c2 * B * npad^2 + c3 * B * npad^3 (one FP64 batch of B problems padded to npad)",
               cost=c.tolist(), max_rel_fit_error=err, measurements=rows, gpu="MI355X (gfx950)",
               seconds=time.perf_counter() - t0)
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(dict(cost=res["cost"], max_rel_fit_error=err)), flush=True)


if __name__ == "__main__":
    main()

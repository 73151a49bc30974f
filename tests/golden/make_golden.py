#!/usr/bin/env python3
"""Write the golden fixtures of this directory from the compiled reference.

Needs /root/reference (the reference sources) and the reference build of
oracle/ (``make -C oracle ref``); runs only in the build container.  For each
case a seeded synthetic system (gammcor_amd.synth) is written in the driver's
input format, the reference's own TwoNO1 + ACEInteg (mode grid), ACCAS (mode
accas, its 5-point GauLeg and pair selection), AB_CAS (mode dumpab) or AC0CAS
(mode ac0, the linearised AC correction) runs,
and the printed full-precision values are stored as JSON.  Inputs are not
stored: they are regenerated from the seeds recorded in each fixture.

    python tests/golden/make_golden.py small      # N <= 48 cases (seconds)
    python tests/golden/make_golden.py n160       # NBasis=160 config (minutes)
    python tests/golden/make_golden.py n240       # NBasis=240 config (about an hour, 25 GB)
    python tests/golden/make_golden.py dimers     # SAPT E2disp of the seeded dimers
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gammcor_amd import synth  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

# name -> (system args, mode, number of alpha points)
CASES = {
    "n12_cas44": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0), "grid", 8),
    "n16_cas22": (dict(nbasis=16, ninact=3, nact=2, nelact=2, rank=30, sigma=0.1, seed=1), "grid", 5),
    "n20_cas44_noinact": (dict(nbasis=20, ninact=0, nact=4, nelact=4, rank=50, sigma=0.1, seed=2), "grid", 3),
    "n24_cas66": (dict(nbasis=24, ninact=3, nact=6, nelact=6, rank=60, sigma=0.1, seed=5), "grid", 4),
    "n48_cas44": (dict(synth.CONFIGS["n48_cas44"]), "grid", 8),
    "n48_cas44_accas": (dict(synth.CONFIGS["n48_cas44"]), "accas", 5),
    # the NBasis=48 system with a plain normalised Gaussian CI vector (no
    # weight put on the leading determinant)
    "n48_cas44_plainci": (dict(synth.CONFIGS["n48_cas44"], ci_ref_weight=0.0), "grid", 8),
    "n12_cas44_ab": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0), "dumpab", 1),
    "n160_cas88": (dict(synth.CONFIGS["n160_cas88"]), "grid", 8),
    # the NBasis=240 workload with a 4-point rule: the reference needs about
    # 25 GB and an hour here, the GPU test compares every W(alpha) and ECorr
    "n240_cas88": (dict(synth.CONFIGS["n240_cas88"]), "grid", 4),
    "n12_cas44_ac0": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0), "ac0", 1),
    "n16_cas22_ac0": (dict(nbasis=16, ninact=3, nact=2, nelact=2, rank=30, sigma=0.1, seed=1), "ac0", 1),
    "n20_cas44_noinact_ac0": (dict(nbasis=20, ninact=0, nact=4, nelact=4, rank=50, sigma=0.1, seed=2), "ac0", 1),
    "n24_cas66_ac0": (dict(nbasis=24, ninact=3, nact=6, nelact=6, rank=60, sigma=0.1, seed=5), "ac0", 1),
    "n48_cas44_ac0": (dict(synth.CONFIGS["n48_cas44"]), "ac0", 1),
    # edge cases of the reference's branches on the NBasis=12 system:
    # a nearly degenerate active pair discarded (accas.f:67-71); an active
    # occupation of exactly 1.0 (the Occ.Eq.One branches, interpa.f:3926,
    # 3962, 3999, 4028) against 1 - 1e-12 (the active branches); the
    # Hamiltonian scaled so the smallest omega lies just above 1e-3 at three
    # nodes and below it (omega^2 <= 1e-6: left unnormalised and filtered
    # out, interpa.f:1748-1776, acfd.f:4976, 5013) at the fourth
    "n12_cas44_degen": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0,
                             degenerate_active=(1, 2)), "grid", 4),
    "n12_cas44_occ1": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0,
                            occ_act_set=[(0, 1.0)]), "grid", 4),
    "n12_cas44_occ1eps": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0,
                               occ_act_set=[(0, 1.0 - 1e-12)]), "grid", 4),
    "n12_cas44_omega_edge": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0,
                                  energy_scale=1.2823e-3), "grid", 4),
    # AC0CAS_FOFO (abfofo.f90:1485) on the FFOO/FOFO records of the same systems
    "n12_cas44_ac0fofo": (dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed=0), "ac0fofo", 1),
    "n16_cas22_ac0fofo": (dict(nbasis=16, ninact=3, nact=2, nelact=2, rank=30, sigma=0.1, seed=1), "ac0fofo", 1),
    "n20_cas44_noinact_ac0fofo": (dict(nbasis=20, ninact=0, nact=4, nelact=4, rank=50, sigma=0.1, seed=2), "ac0fofo",
                                  1),
    "n24_cas66_ac0fofo": (dict(nbasis=24, ninact=3, nact=6, nelact=6, rank=60, sigma=0.1, seed=5), "ac0fofo", 1),
    "n48_cas44_ac0fofo": (dict(synth.CONFIGS["n48_cas44"]), "ac0fofo", 1),
}
SMALL = [k for k in CASES if not k.startswith(("n160", "n240"))]

# SAPT dimers (mode e2disp): name -> (dimer args of synth.dimer)
DIMERS = {
    "d12_cas44_e2disp": dict(nbasis=12, ninact=2, nact=4, nelact=4, rank=36, sigma=0.1, seed_a=0, seed_b=7),
    "d24_cas66_e2disp": dict(nbasis=24, ninact=3, nact=6, nelact=6, rank=60, sigma=0.1, seed_a=5, seed_b=8),
}


def run_dimer(name):
    """Two monomers through the reference's SAPT response chain and its
    e2disp (driver mode e2disp)."""
    args = DIMERS[name]
    a, b, vab = synth.dimer(**args)
    alphas, weights = synth.gauss_legendre_01(1)
    with tempfile.TemporaryDirectory() as d:
        synth.write_inputs(a, d, alphas, weights, tag="_A")
        synth.write_inputs(b, d, alphas, weights, tag="_B")
        a.twoel_ao.tofile(os.path.join(d, "twoel_ao.bin"))
        synth.twomoab_records(vab, a, b).astype("<f8").tofile(os.path.join(d, "TWOMOAB"))
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {DRIVER} e2disp"], cwd=d, capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference driver failed\n{r.stdout[-3000:]}{r.stderr[-3000:]}")
    res = dict(case=name, mode="e2disp", dimer=args, generator="gammcor_amd.synth.dimer")
    for line in r.stdout.splitlines():
        t = line.split()
        if t and t[0] == "MONOMER":
            res[f"ndimx_{t[1]}"] = int(t[2])
            res[f"ecasscf_{t[1]}"] = float(t[3])
        elif t and t[0] == "E2DISP":
            res["e2disp"] = float(t[1])
    return res


def parse(out):
    res = dict(walpha=[], alphas=[], weights=[], eig=[], twono=[])
    for line in out.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "TWONO_SUMS":
            res["twono_sum"], res["twono_sumsq"] = float(t[1]), float(t[2])
        elif t[0] == "TWONO":
            res["twono"].append([int(t[1]), int(t[2]), int(t[3]), int(t[4]), float(t[5])])
        elif t[0] == "NDIMX":
            res["ndimx"] = int(t[1])
        elif t[0] == "WALPHA":
            res["alphas"].append(float(t[2]))
            res["weights"].append(float(t[3]))
            res["walpha"].append(float(t[4]))
        elif t[0] == "EIG":
            res["eig"].append(dict(npos=int(t[2]), sum_kept=float(t[3]), first=float(t[4]), mid=float(t[5]),
                                   last=float(t[6])))
        elif t[0] == "ECASSCF":
            res["ecasscf"] = float(t[1])
        elif t[0] == "ECORR":
            res["ecorr"] = float(t[1])
        elif t[0] == "AC0ECORR":
            res["ac0_ecorr"] = float(t[1])
        elif line.strip().startswith("ACAlpha"):
            # ACECORR list-directed line: ACAlpha <a> W_ALPHA <w>  (acfd.f:197)
            res.setdefault("accas_alphas", []).append(float(t[1]))
            res.setdefault("accas_walpha", []).append(float(t[3]))
        elif "AC-ERPA-CASSCF" in line:
            v = t[-3:]
            res["accas_line"] = [float(x) for x in v]
    return res


def run_case(name):
    args, mode, nal = CASES[name]
    s = synth.make_system(**args)
    if "occ_act_set" in args:
        # the occupations reach the driver as binary doubles: 1 - 1e-12 stays
        # apart from 1.0
        assert all(s.occ[s.ninact + k] == v for k, v in args["occ_act_set"])
    alphas, weights = synth.gauss_legendre_01(nal)
    with tempfile.TemporaryDirectory() as d:
        synth.write_inputs(s, d, alphas, weights)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {DRIVER} {mode}"], cwd=d, capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference driver failed\n{r.stderr[-3000:]}")
        res = parse(r.stdout)
        if mode == "dumpab":
            n = res["ndimx"]
            ap = np.fromfile(os.path.join(d, "abplus.bin")).reshape(n, n, order="F")
            am = np.fromfile(os.path.join(d, "abmin.bin")).reshape(n, n, order="F")
            res["abplus"] = ap.ravel(order="F").tolist()
            res["abmin"] = am.ravel(order="F").tolist()
    res.update(case=name, mode=mode, nalpha=nal, system=s.meta,
               generator="gammcor_amd.synth.make_system", quadrature="numpy leggauss mapped to [0,1]")
    return res


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "small"
    if which == "dimers" or which in DIMERS:
        for name in (DIMERS if which == "dimers" else [which]):
            res = run_dimer(name)
            with open(os.path.join(HERE, f"ref_{name}.json"), "w") as f:
                json.dump(res, f, indent=1)
            print(name, {k: res.get(k) for k in ("ndimx_A", "ndimx_B", "e2disp")}, flush=True)
        return
    names = SMALL if which == "small" else [n for n in CASES if n.startswith(which)] or [which]
    for name in names:
        res = run_case(name)
        path = os.path.join(HERE, f"ref_{name}.json")
        with open(path, "w") as f:
            json.dump(res, f, indent=1)
        print(name, "->", path, {k: res.get(k) for k in ("ndimx", "ecorr", "ecasscf")}, flush=True)


if __name__ == "__main__":
    main()

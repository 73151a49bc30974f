"""The reference program with the Fortran drop-in layer linked in place of its
own TwoNO1, AB_CAS, ERPASYMM1, ACEneERPA, ACEInteg, ACEInteg_FOFO and AC0CAS
(oracle/_ref/ref_driver_gpu, built by `make -C oracle ref_gpu` from the
reference sources plus gammcor_amd/fortran): the reference's own ACECORR and
driver run unchanged and call the GPU through iso_c_binding.  Its printed
results must match the CPU reference's fixtures on the same seeded inputs."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver_gpu")
sys.path.insert(0, os.path.join(HERE, "golden"))

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER),
                                  reason="oracle/_ref/ref_driver_gpu not built (make -C oracle ref_gpu)")


def _run(name, mode, prec=None):
    import make_golden as G
    from gammcor_amd import synth
    fx = G.json.load(open(os.path.join(HERE, "golden", f"ref_{name}.json")))
    m = fx["system"]
    s = synth.make_system(m["nbasis"], m["ninact"], m["nact"], m["nelact"], m["rank"], m["sigma"], m["seed"])
    alphas, weights = synth.gauss_legendre_01(fx["nalpha"])
    env = dict(os.environ)
    if prec:
        env["GAMMCOR_HIP_PREC"] = prec
    with tempfile.TemporaryDirectory() as d:
        synth.write_inputs(s, d, alphas, weights)
        print(f"fortran drop-in {name} mode={mode} prec={prec or 'fp64'}", flush=True)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; {DRIVER} {mode}"], cwd=d, env=env,
                           capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return fx, G.parse(r.stdout), r.stdout


@needs_driver
def test_reference_accas_driver_through_the_fortran_drop_in(cuda):
    fx, res, out = _run("n48_cas44_accas", "accas")
    assert "gammcor_hip: ACEInteg on the GPU" in out
    assert np.max(np.abs(np.array(res["accas_walpha"]) - np.array(fx["accas_walpha"]))) < 1e-9
    assert np.max(np.abs(np.array(res["accas_line"]) - np.array(fx["accas_line"]))) < 1e-8


@needs_driver
@pytest.mark.parametrize("name", ["n16_cas22", "n24_cas66"])
def test_reference_grid_driver_through_the_fortran_drop_in(cuda, name):
    fx, res, out = _run(name, "grid")
    assert "gammcor_hip: TwoNO1 on the GPU" in out
    assert abs(res["twono_sum"] - fx["twono_sum"]) < 1e-9 * max(1.0, abs(fx["twono_sum"]))
    assert np.max(np.abs(np.array(res["walpha"]) - np.array(fx["walpha"]))) < 1e-9
    assert abs(res["ecorr"] - fx["ecorr"]) < 1e-9
    assert abs(res["ecasscf"] - fx["ecasscf"]) < 1e-10


@needs_driver
def test_reference_grid_driver_mixed_precision(cuda):
    fx, res, out = _run("n24_cas66", "grid", prec="mixed")
    assert "precision mode 2" in out
    assert abs(res["ecorr"] - fx["ecorr"]) < 1e-6


@needs_driver
@pytest.mark.parametrize("name", ["n16_cas22", "n20_cas44_noinact"])
def test_reference_fofo_driver_through_the_fortran_drop_in(cuda, name):
    """The driver writes the FFOO and FOFO record files from the transformed
    integrals and calls ACEInteg_FOFO at every node; the drop-in reads the
    files and runs on the GPU.  The reference's own FOFO path reproduces the
    in-core fixture values to 1e-13 on these systems."""
    fx, res, out = _run(name, "fofo")
    assert "gammcor_hip: ACEInteg_FOFO on the GPU" in out
    assert np.max(np.abs(np.array(res["walpha"]) - np.array(fx["walpha"]))) < 1e-9
    assert abs(res["ecorr"] - fx["ecorr"]) < 1e-9
    assert abs(res["ecasscf"] - fx["ecasscf"]) < 1e-10


@needs_driver
@pytest.mark.parametrize("name", ["n16_cas22_ac0", "n20_cas44_noinact_ac0", "n48_cas44_ac0"])
def test_reference_ac0_driver_through_the_fortran_drop_in(cuda, name):
    """The driver's AC0CAS call (with its TwoNO1) lands on the drop-in, which
    reads COMMON and rdm2.dat as the reference does and runs on the GPU."""
    fx, res, out = _run(name, "ac0")
    assert "gammcor_hip: AC0CAS on the GPU" in out
    assert abs(res["ac0_ecorr"] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"]))
    assert abs(res["ecasscf"] - fx["ecasscf"]) < 1e-10


@needs_driver
@pytest.mark.parametrize("name", ["n16_cas22_ac0fofo", "n24_cas66_ac0fofo"])
def test_reference_ac0fofo_driver_through_the_fortran_drop_in(cuda, name):
    """The driver writes the FFOO and FOFO record files and calls
    AC0CAS_FOFO (a procedure of MODULE abfofo, rebound to the external
    drop-in at link time), which reads them and runs on the GPU."""
    fx, res, out = _run(name, "ac0fofo")
    assert "gammcor_hip: AC0CAS_FOFO on the GPU" in out
    assert abs(res["ac0_ecorr"] - fx["ac0_ecorr"]) < 1e-10 * max(1.0, abs(fx["ac0_ecorr"]))

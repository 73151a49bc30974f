"""Build the Fortran module of the C ABI (gammcor_amd/fortran/gammcor_hip.f90)
with amdflang: object and .mod into gammcor_amd/fortran/_build/.

The drop-in subroutines (gammcor_dropin.f) include the reference's
commons.inc, so they are compiled as part of a reference build, against that
tree's include directory (oracle/Makefile, target ref_gpu, shows the recipe).
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fortran")
OUT = os.path.join(SRC, "_build")
FC = os.environ.get("FC", "amdflang")


def available():
    return shutil.which(FC) is not None


def build():
    os.makedirs(OUT, exist_ok=True)
    obj = os.path.join(OUT, "gammcor_hip.o")
    cmd = [FC, "-O2", "-fPIC", "-module-dir", OUT, "-c", os.path.join(SRC, "gammcor_hip.f90"), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{FC} failed:\n{r.stderr[-4000:]}")
    return obj


if __name__ == "__main__":
    print(build())

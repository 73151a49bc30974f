"""MI355X (gfx950) AC-CAS correlation-energy path of GAMMCOR, rebuilt on HIP.

The compute runs in libgammcor_hip.so (see include/gammcor_hip.h); this
package holds the host-side mirror of the reference interface, the build
recipe and the seeded synthetic-input generator.
"""
from .host import ACECORR, ACCAS_pairs, DeviceSystem, GauLeg, TwoNO1, geminals, pack_rdm2act  # noqa: F401

__all__ = ["ACECORR", "ACCAS_pairs", "DeviceSystem", "GauLeg", "TwoNO1", "geminals", "pack_rdm2act"]

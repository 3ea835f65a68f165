"""MI355X (gfx950) implementation of HAWASSI's Hamiltonian pseudo-spectral
right-hand side (AB2/HS2, 2-D flat and two-reference-depth) and its embedded
RKF45 stepper, behind the C ABI in include/hawassi_hip.h."""
from ._lib import HW_PREC_FP32, HW_PREC_FP64, HipPathMissing, HwError, load  # noqa: F401

__version__ = "0.1.0"

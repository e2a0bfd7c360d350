"""MI355X-native (gfx950) spectral-viscosity incompressible Euler RHS + SSP-RK3.

The compute path is libazeban_hip.so (hand-written HIP kernels behind the C ABI
in include/azeban_hip.h); this package is its Python binding and a mirror of
TobiasRohner/azeban's Equation / TimeIntegrator / Simulation interface.
"""
from .solver import (  # noqa: F401
    Grid,
    IncompressibleEuler,
    Simulation,
    SmoothCutoff1D,
    SSP_RK3,
    axpby_,
    copy_from_padded,
    copy_to_padded,
    energy_spectrum,
    fft_c2r,
    fft_r2c,
    initialize,
    leray_,
    make_smooth_cutoff_1d,
    max_norm,
    scale_,
    to_physical,
)

"""MI355X-native excess-functional (c-proj) hot path of Mini-MDFT.

The compute path is libmdft_hip.so (HIP kernels for gfx950 behind the C ABI of
include/mdft_hip.h); this package is its Python face.
"""
from ._lib import DOUBLE, FLOAT, FP32, FP64, MIXED, MIXED_X6, MIXED_X6F, MDFTError  # noqa: F401
from .mdft import (  # noqa: F401
    AngularGrid,
    Convolution,
    Functional,
    OrientationProjectionMap,
    OrientationProjectionTransform,
    SOLVENTS,
    Server,
    Solver,
    SpatialGrid,
    fft3d,
    fill_uniform,
    get_delta_f,
    get_delta_rho,
    kT,
    prefactor_of,
    rho0_of,
    rotation_matrices,
    vext,
)

"""gacelle_amd -- MI355X (gfx950) implementation of GACELLE's voxelwise
estimation hot path: askadam, mcmc (Metropolis-Hastings, Goodman-Weare) and
the forward models they drive, as hand-written HIP kernels behind a C ABI
(include/gacelle.h, libgacelle_hip.so)."""
from . import _abi, utils
from ._abi import GacelleError, load_library
from .askadam import askadam
from .mcmc import mcmc
from .models import gpuJointR1R2starMapping, gpuMCRMWI, gpuNEXI, gpuNEXIrice, gpuR2starMapping

__all__ = [
    "askadam",
    "mcmc",
    "gpuR2starMapping",
    "gpuJointR1R2starMapping",
    "gpuNEXI",
    "gpuNEXIrice",
    "gpuMCRMWI",
    "utils",
    "GacelleError",
    "load_library",
]

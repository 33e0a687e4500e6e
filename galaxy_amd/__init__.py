"""galaxy_amd — MI355X (gfx950) HIP implementation of GALAXY's resident Barnes-Hut step.

The compute path is libgalaxy_bh.so (HIP kernels + C ABI, include/galaxy_bh.h).
Python modules here are thin host-side mirrors of the reference interfaces.
"""
from .bh import (  # noqa: F401
    BhError,
    Config,
    EvolvingState,
    NbodyGpu,
    ParallelGpuTree,
    ParallelTreeRuntime,
    TreeEvidence,
)

__all__ = [
    "BhError",
    "Config",
    "EvolvingState",
    "NbodyGpu",
    "ParallelGpuTree",
    "ParallelTreeRuntime",
    "TreeEvidence",
]

"""MI355X-native particle tracking for MaCaw-style Monte Carlo transport.

The product is the HIP library behind the C ABI in include/macaw_amd.h
(macaw_amd/lib/libmacaw_amd.so, built by `python -m macaw_amd.build`).
`macaw_amd.lib` binds it; `macaw_amd.study` mirrors the MOOSE objects of
the reference (OpenMCStudy, CollisionKernel, RayBCs, OpenMCTally,
OpenMCTallyAux); `macaw_amd.problems` generates the synthetic inputs.
"""
from . import abi  # noqa: F401

"""LyceanEM raycast + frequency-domain propagation path on MI355X (gfx950).

A drop-in for the reference package's hot path: the module paths
``lyceanem.models.frequency_domain``, ``lyceanem.raycasting.rayfunctions``,
``lyceanem.electromagnetics.empropagation`` and
``lyceanem.models.acceleration_structures`` keep the reference's function names,
arguments and return types; the compute runs in hand-written HIP kernels
(``_lib/liblem.so``, C-ABI in ``include/lem.h``).
"""
__version__ = "0.1.0"

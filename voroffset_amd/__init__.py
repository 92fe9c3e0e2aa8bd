"""voroffset_amd -- MI355X-native dexel morphology (dilation, erosion,
opening, closing of CompressedVolume dexel grids by a Euclidean ball).

The compute path is libvoroffset_hip.so (hand-written HIP kernels for gfx950,
C ABI in include/voroffset_hip.h); this package holds the volume container,
the ctypes binding and the method object with the reference's interface.
"""

from .volume import CompressedVolume  # noqa: F401

__all__ = ["CompressedVolume", "VoronoiMorphoHip"]


def __getattr__(name):
    if name == "VoronoiMorphoHip":
        from .morpho import VoronoiMorphoHip
        return VoronoiMorphoHip
    raise AttributeError(name)

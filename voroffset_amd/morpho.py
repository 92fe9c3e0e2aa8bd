"""Morphology methods with the reference's plugin interface.

Mirrors ``vor3d::VoronoiMorpho`` (src/vor3d/Voronoi.h:13-44): a method object
with ``dilation(input, radius) -> (result, time_1, time_2)`` and
``erosion(...)``, plus the CLI's ``closing`` / ``opening`` composition
(app/cli3d/offset3d.cpp:124-133).  ``VoronoiMorphoHip`` is the ``-m hip``
method: everything runs on the GPU through libvoroffset_hip.so; errors raise
(the reference's vor_assert throws std::runtime_error, Common.cpp:8-18).
"""

from __future__ import annotations

from . import _native
from .volume import CompressedVolume


class VoronoiMorphoHip:
    """The `-m hip` method.  ``precision`` is the device storage precision of
    the endpoints ("fp32" or "fp64"); predicates are FP64 either way."""

    def __init__(self, precision="fp32", device=0):
        if precision not in _native.PRECISIONS:
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self.ctx = _native.Context(device)

    def apply(self, op, volume: CompressedVolume, radius: float):
        """Returns (result volume, time_1 ms, time_2 ms)."""
        zlo, zhi = volume.z_range()
        off, z, (t1, t2) = self.ctx.morph_host(op, self.precision, radius, volume.nx, volume.ny,
                                               volume.off, volume.z, zlo, zhi)
        return volume.like(off, z), t1, t2

    def dilation(self, volume, radius):
        return self.apply("dilation", volume, radius)

    def erosion(self, volume, radius):
        return self.apply("erosion", volume, radius)

    def closing(self, volume, radius):
        return self.apply("closing", volume, radius)

    def opening(self, volume, radius):
        return self.apply("opening", volume, radius)

    def apply_batch(self, items):
        """Batched form: items = [(op, volume, radius), ...] all run together on
        the device (one launch per pass and stage for the whole batch).
        Returns ([result volumes], time_1, time_2)."""
        import numpy as np

        dtype = np.float32 if self.precision == "fp32" else np.float64
        keep, jobs, shifts = [], [], []
        try:
            for op, vol, r in items:
                zlo, zhi = vol.z_range()
                shift = _native.storage_shift(zlo, zhi) if self.precision == "fp32" else 0.0
                d_off = _native.DeviceArray(vol.off)
                d_z = _native.DeviceArray((vol.z - shift).astype(dtype))
                keep += [d_off, d_z]
                shifts.append(shift)
                jobs.append((op, r, vol.nx, vol.ny, d_off.address, d_z.address, zlo - shift, zhi - shift))
            outs, (t1, t2) = self.ctx.morph_batch_device(self.precision, jobs)
            results = []
            for (v, cnt), (op, vol, r), shift in zip(outs, items, shifts):
                off = _native.to_host(v.offsets, np.int64, v.nx * v.ny + 1)
                z = _native.to_host(v.z, dtype, cnt).astype(np.float64) + shift
                results.append(vol.like(off, z))
            return results, t1, t2
        finally:
            for d in keep:
                d.free()

    def _upload(self, vol, shift=0.0):
        import numpy as np
        dtype = np.float32 if self.precision == "fp32" else np.float64
        return _native.DeviceArray(vol.off), _native.DeviceArray((vol.z - shift).astype(dtype))

    def get_volume(self, volume):
        """(measure, number of segments): CompressedVolume::get_volume and
        numSegments (CompressedVolume.cpp:61-83), reduced on the device."""
        import numpy as np
        zlo, zhi = volume.z_range()
        shift = _native.storage_shift(zlo, zhi) if self.precision == "fp32" else 0.0  # lengths do not see it
        d_off, d_z = self._upload(volume, shift)
        try:
            return self.ctx.measure_device(self.precision, volume.nx, volume.ny, d_off.address, d_z.address,
                                           volume.spacing)
        finally:
            d_off.free()
            d_z.free()

    def calculate_xor(self, a, b):
        """VoronoiMorpho::calculateXor (Voronoi.cpp:91-111): returns (measure
        of the symmetric difference, the xor volume), with a's geometry."""
        import numpy as np
        dtype = np.float32 if self.precision == "fp32" else np.float64
        zlo, zhi = a.z_range()
        shift = _native.storage_shift(zlo, zhi) if self.precision == "fp32" else 0.0
        arrays = [*self._upload(a, shift), *self._upload(b, shift)]
        try:
            va = _native.VoVolume(a.nx, a.ny, arrays[0].address, arrays[1].address, zlo - shift, zhi - shift)
            vb = _native.VoVolume(b.nx, b.ny, arrays[2].address, arrays[3].address, zlo - shift, zhi - shift)
            out, cnt, vol = self.ctx.xor_device(self.precision, va, vb, a.spacing)
            off = _native.to_host(out.offsets, np.int64, a.nx * a.ny + 1)
            z = _native.to_host(out.z, dtype, cnt).astype(np.float64) + shift
            return vol, a.like(off, z)
        finally:
            for d in arrays:
                d.free()

    def close(self):
        self.ctx.close()


METHODS = {"hip": VoronoiMorphoHip}

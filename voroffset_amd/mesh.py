"""Triangle meshes: loading, synthetic generators, and mesh -> dexels.

Replaces the parts of the reference's dexelization that sit around the ray
casting (src/vor3d/Dexelize.cpp:231-274): geogram's ``mesh_load`` (own OBJ,
OFF, STL and PLY readers here), the bounding box, the voxel size from
``num_voxels`` and the volume geometry (CompressedVolume.cpp:11-23).  The ray
casting itself runs on the GPU (``vo_rasterize``); there is no CPU fallback.
"""

from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from .volume import CompressedVolume


@dataclass
class TriMesh:
    vertices: np.ndarray   # (nv, 3) float64
    triangles: np.ndarray  # (nf, 3) int32

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, np.float64).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(self.triangles, np.int32).reshape(-1, 3)
        if self.triangles.size and (self.triangles.min() < 0 or self.triangles.max() >= len(self.vertices)):
            raise ValueError("triangle index out of range")

    @property
    def bbox(self):
        """(min corner, max corner), as geogram's get_bbox (Dexelize.cpp:253)."""
        if not len(self.vertices):
            raise ValueError("empty mesh")
        return self.vertices.min(axis=0), self.vertices.max(axis=0)

    @staticmethod
    def concat(meshes):
        vs, fs, base = [], [], 0
        for m in meshes:
            vs.append(m.vertices)
            fs.append(m.triangles + base)
            base += len(m.vertices)
        return TriMesh(np.concatenate(vs) if vs else np.zeros((0, 3)),
                       np.concatenate(fs) if fs else np.zeros((0, 3), np.int32))


# -- readers ----------------------------------------------------------------

def _fan(poly):
    return [(poly[0], poly[i], poly[i + 1]) for i in range(1, len(poly) - 1)]


def load_obj(path):
    V, F = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                V.append([float(t[1]), float(t[2]), float(t[3])])
            elif t[0] == "f":
                idx = []
                for w in t[1:]:
                    i = int(w.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(V) + i)
                F.extend(_fan(idx))
    return TriMesh(np.array(V, np.float64).reshape(-1, 3), np.array(F, np.int32).reshape(-1, 3))


def load_off(path):
    with open(path) as f:
        toks = []
        for line in f:
            line = line.split("#")[0]
            toks.extend(line.split())
    if not toks or not toks[0].endswith("OFF"):
        raise ValueError("not an OFF file")
    k = 1
    nv, nf = int(toks[k]), int(toks[k + 1])
    k += 3
    V = np.array(toks[k:k + 3 * nv], np.float64).reshape(nv, 3)
    k += 3 * nv
    F = []
    for _ in range(nf):
        n = int(toks[k])
        F.extend(_fan([int(v) for v in toks[k + 1:k + 1 + n]]))
        k += 1 + n
    return TriMesh(V, np.array(F, np.int32).reshape(-1, 3))


def load_stl(path):
    data = open(path, "rb").read()
    # binary: 80-byte header, uint32 count, 50 bytes per triangle
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            V = rec["v"].reshape(-1, 3).astype(np.float64)
            return TriMesh(V, np.arange(3 * n, dtype=np.int32).reshape(-1, 3))
    V = []
    for line in data.decode("ascii", "replace").splitlines():
        t = line.split()
        if len(t) == 4 and t[0] == "vertex":
            V.append([float(t[1]), float(t[2]), float(t[3])])
    V = np.array(V, np.float64).reshape(-1, 3)
    return TriMesh(V, np.arange(len(V), dtype=np.int32).reshape(-1, 3))


_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
              "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
              "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def load_ply(path):
    data = open(path, "rb").read()
    end = data.index(b"end_header")
    end = data.index(b"\n", end) + 1
    header = data[:end].decode("ascii").splitlines()
    fmt, elements = None, []
    for line in header:
        t = line.split()
        if not t:
            continue
        if t[0] == "format":
            fmt = t[1]
        elif t[0] == "element":
            elements.append([t[1], int(t[2]), []])
        elif t[0] == "property":
            elements[-1][2].append(t[1:])
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"PLY format {fmt} not supported")
    V = F = None
    if fmt == "ascii":
        toks = data[end:].decode("ascii").split()
        k = 0
        for name, count, props in elements:
            rows = []
            for _ in range(count):
                row = []
                for p in props:
                    if p[0] == "list":
                        n = int(toks[k])
                        row.append([int(v) for v in toks[k + 1:k + 1 + n]])
                        k += 1 + n
                    else:
                        row.append(float(toks[k]))
                        k += 1
                rows.append(row)
            if name == "vertex":
                names = [p[-1] for p in props]
                V = np.array([[r[names.index(c)] for c in "xyz"] for r in rows], np.float64).reshape(-1, 3)
            elif name == "face":
                F = [tri for r in rows for tri in _fan(r[0])]
    else:
        e = "<" if fmt == "binary_little_endian" else ">"
        k = end
        for name, count, props in elements:
            if all(p[0] != "list" for p in props):
                dt = np.dtype([(p[1], e + _PLY_TYPES[p[0]]) for p in props])
                arr = np.frombuffer(data, dtype=dt, count=count, offset=k)
                k += dt.itemsize * count
                if name == "vertex":
                    V = np.stack([arr["x"], arr["y"], arr["z"]], axis=1).astype(np.float64)
                continue
            rows = []
            for _ in range(count):
                row = []
                for p in props:
                    if p[0] == "list":
                        ct, it = np.dtype(e + _PLY_TYPES[p[1]]), np.dtype(e + _PLY_TYPES[p[2]])
                        n = int(np.frombuffer(data, ct, 1, k)[0])
                        k += ct.itemsize
                        row.append(np.frombuffer(data, it, n, k).astype(np.int64).tolist())
                        k += it.itemsize * n
                    else:
                        t = np.dtype(e + _PLY_TYPES[p[0]])
                        row.append(float(np.frombuffer(data, t, 1, k)[0]))
                        k += t.itemsize
                rows.append(row)
            if name == "face":
                F = [tri for r in rows for tri in _fan(r[0])]
    if V is None:
        raise ValueError("PLY without vertices")
    return TriMesh(V, np.array(F or [], np.int32).reshape(-1, 3))


def load(path):
    """Mesh from .obj, .off, .stl or .ply (the formats geogram's mesh_load reads
    for the reference's inputs, Dexelize.cpp:247)."""
    ext = os.path.splitext(path)[1].lower()
    readers = {".obj": load_obj, ".off": load_off, ".stl": load_stl, ".ply": load_ply}
    if ext not in readers:
        raise ValueError(f"unsupported mesh format {ext!r}")
    return readers[ext](path)


def save_obj(mesh, path):
    with open(path, "w") as f:
        for v in mesh.vertices:
            f.write("v %r %r %r\n" % tuple(map(float, v)))
        for t in mesh.triangles:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


# -- generators -------------------------------------------------------------

def icosphere(level=3, center=(0.0, 0.0, 0.0), radius=1.0):
    """Subdivided icosahedron: 10 * 4^level + 2 vertices, 20 * 4^level faces
    (642 and 1280 at level 3), outward-oriented."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nf
    Vn = np.array(V) * radius + np.asarray(center, np.float64)
    return TriMesh(Vn, np.array(F, np.int32))


def icosphere_union(n=32, level=3, seed=0):
    """The 32-sphere plumbing mesh: centres uniform in the unit cube, radii
    U[0.05, 0.2], one seed; overlapping spheres stay separate shells, so
    their overlap fills even-odd."""
    rng = np.random.default_rng(seed)
    base = icosphere(level)
    parts = []
    for _ in range(n):
        c = rng.uniform(0.0, 1.0, 3)
        r = rng.uniform(0.05, 0.2)
        parts.append(TriMesh(base.vertices * r + c, base.triangles))
    return TriMesh.concat(parts)


def prism(p0, p1, radius, sides=8):
    """Closed prism of `sides` sides around the segment p0 -> p1: 2*sides
    vertices, 2*sides side triangles and 2*(sides-2) cap triangles (28 for 8
    sides), outward-oriented."""
    p0, p1 = np.asarray(p0, np.float64), np.asarray(p1, np.float64)
    axis = p1 - p0
    axis /= np.linalg.norm(axis)
    helper = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(axis, helper)
    u /= np.linalg.norm(u)
    v = np.cross(axis, u)
    ang = 2 * np.pi * np.arange(sides) / sides
    ring = radius * (np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * v)
    V = np.concatenate([p0 + ring, p1 + ring])
    F = []
    for i in range(sides):
        j = (i + 1) % sides
        F += [(i, j, sides + j), (i, sides + j, sides + i)]
    for i in range(1, sides - 1):
        F.append((0, i + 1, i))                          # bottom cap, facing -axis
        F.append((sides, sides + i, sides + i + 1))      # top cap, facing +axis
    return TriMesh(V, np.array(F, np.int32))


def prism_tubes(n=2000, box=512.0, seed=0, rmin=0.5, rmax=2.0, lmin=5.0, lmax=60.0):
    """The filigree mesh: n random 8-sided prisms (radius U[rmin, rmax],
    length U[lmin, lmax], direction uniform on the sphere, centre uniform in
    the box shrunk by half a length), in dexel units of a `box`-sized grid."""
    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(n):
        r = rng.uniform(rmin, rmax)
        length = rng.uniform(lmin, lmax)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        m = 0.5 * length + r
        c = rng.uniform(m, box - m, 3)
        parts.append(prism(c - 0.5 * length * d, c + 0.5 * length * d, r))
    return TriMesh.concat(parts)


# -- mesh -> dexels ---------------------------------------------------------

def dexel_grid(mesh, voxel_size, padding, num_voxels=-1):
    """The volume geometry create_dexels derives from the mesh
    (Dexelize.cpp:252-270): bounding box, voxel size from num_voxels along
    the longest axis when num_voxels > 0, then CompressedVolume's constructor
    (padding added on every side).  Returns (empty volume, voxel size)."""
    lo, hi = mesh.bbox
    extent = hi - lo
    if num_voxels > 0:
        voxel_size = float(max(extent)) / num_voxels
    if not (voxel_size > 0):
        raise ValueError("voxel size must be > 0")
    return CompressedVolume.from_bbox(lo, extent, voxel_size, padding), voxel_size


def create_dexels(mesh, voxel_size=1.0, padding=0, num_voxels=-1, ctx=None):
    """Mesh -> CompressedVolume on the GPU (vor3d::create_dexels,
    Dexelize.cpp:231-274).  Column values are z / spacing of the hits, sorted;
    consecutive pairs are segments."""
    from . import _native

    vol, _ = dexel_grid(mesh, voxel_size, padding, num_voxels)
    own = ctx is None
    ctx = ctx or _native.Context(0)
    try:
        off, z = ctx.rasterize(mesh.vertices, mesh.triangles, vol.nx, vol.ny, vol.origin[0], vol.origin[1],
                               vol.spacing)
    finally:
        if own:
            ctx.close()
    return vol.like(off, z)


__all__ = ["TriMesh", "load", "load_obj", "load_off", "load_stl", "load_ply", "save_obj", "icosphere",
           "icosphere_union", "prism", "prism_tubes", "dexel_grid", "create_dexels"]

"""offset3d_hip, the CLI of the GPU path (app/cli3d/offset3d.cpp's options):
argument handling on the CPU; on the GPU, mesh in -> morphology -> volume out
against the oracle, and the C++ mesh readers on every supported format."""

import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from voroffset_amd import mesh as M
from voroffset_amd.volume import CompressedVolume

from .helpers import compare_csr
from .test_raster import _write_off, _write_ply, _write_stl

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "voroffset_amd", "offset3d_hip")


def run(*args):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        pytest.fail("offset3d_hip is not built (python -m voroffset_amd.build)")


def test_help_and_bad_arguments(tmp_path):
    assert run("--help").returncode == 0
    assert run().returncode == 1
    assert run(tmp_path / "missing.obj").returncode == 1
    p = tmp_path / "a.obj"
    M.save_obj(M.icosphere(1), p)
    r = run(p, "--bogus")
    assert r.returncode == 1 and "unknown option" in r.stderr
    assert run(p, "-m", "ours").returncode == 1
    assert run(p, "-x", "melt").returncode == 1


@pytest.mark.gpu
def test_mesh_dilation_matches_oracle(tmp_path):
    mesh = M.icosphere_union(32, level=3, seed=1)
    src = tmp_path / "spheres.obj"
    M.save_obj(mesh, src)
    out, js = tmp_path / "out.dexels", tmp_path / "out.json"
    r = run(src, "-o", out, "-j", js, "-n", 128, "-p", 8, "-r", 4, "-m", "hip64", "-x", "dilation", "-f")
    assert r.returncode == 0, r.stderr
    got = CompressedVolume.load(str(out))
    vol, _ = M.dexel_grid(M.load(str(src)), 1.0, 8, 128)
    off, z = oracle.rasterize(mesh.vertices, mesh.triangles, vol.nx, vol.ny, vol.origin[0], vol.origin[1],
                              vol.spacing)
    ref = oracle.morph(vol.like(off, z), "dilation", 4.0, threads=8)
    assert (got.nx, got.ny) == (vol.nx, vol.ny)
    bad, err = compare_csr(got.off, got.z, *ref)
    assert len(bad) == 0 and err <= 1e-9
    meta = json.load(open(js))
    assert meta["grid_size"] == [vol.nx, vol.ny] and meta["operation"] == "dilation"
    assert meta["num_segments"] == int(np.sum(np.diff(off)) // 2)


@pytest.mark.gpu
def test_mesh_formats_give_the_same_dexels(tmp_path):
    m = M.TriMesh.concat([M.icosphere(2, (4.0, 5.0, 6.0), 3.0), M.prism((1, 1, 1), (9, 8, 7), 1.2)])
    M.save_obj(m, tmp_path / "m.obj")
    _write_off(m, tmp_path / "m.off")
    _write_ply(m, tmp_path / "m.ply", True)
    outs = {}
    for name in ("m.obj", "m.off", "m.ply"):
        o = tmp_path / (name + ".dexels")
        r = run(tmp_path / name, "-o", o, "-d", 0.25, "-n", -1, "-p", 2, "-x", "noop", "-f")
        assert r.returncode == 0, r.stderr
        outs[name] = CompressedVolume.load(str(o))
    a = outs["m.obj"]
    for name in ("m.off", "m.ply"):
        assert np.array_equal(outs[name].off, a.off) and np.array_equal(outs[name].z, a.z), name
    # binary STL stores float32 vertices: same dexels as the float32-rounded mesh
    _write_stl(m, tmp_path / "m.stl", True)
    o = tmp_path / "stl.dexels"
    assert run(tmp_path / "m.stl", "-o", o, "-d", 0.25, "-n", -1, "-p", 2, "-x", "noop", "-f").returncode == 0
    s = CompressedVolume.load(str(o))
    assert s.num_segments() == a.num_segments()


@pytest.mark.gpu
def test_obj_writer_emits_a_hexahedron_per_segment(tmp_path):
    src = tmp_path / "s.obj"
    M.save_obj(M.icosphere(2, (0.0, 0.0, 0.0), 1.0), src)
    out = tmp_path / "hex.obj"
    assert run(src, "-o", out, "-n", 16, "-p", 2, "-r", 1, "-x", "dilation", "-f").returncode == 0
    dex = tmp_path / "d.dexels"
    assert run(src, "-o", dex, "-n", 16, "-p", 2, "-r", 1, "-x", "dilation", "-f").returncode == 0
    nseg = CompressedVolume.load(str(dex)).num_segments()
    r = M.load(str(out))
    assert len(r.vertices) == 8 * nseg and len(r.triangles) == 12 * nseg

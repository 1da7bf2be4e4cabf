"""Flat binary scene files (bgpu_scene_write / bgpu_scene_read,
bling_amd/csrc/scene_file.cpp): a descriptor written and read back holds
the same arrays, records, light maps, camera and film byte for byte; the
oracle renders the read-back scene to the same image; malformed files are
refused with BGPU_E_INVALID.  Host-only calls: no GPU is needed."""
import ctypes as C
import os

import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd.runtime import JobSpec
from bling_amd.scene import SceneFile
from oracle import pyoracle
from tests import cases


def _records(ptr, n, typ):
    return [bytes(ptr[i]) for i in range(n)]


def _light_bytes(desc):
    out = []
    for i in range(desc.num_lights):
        l = abi.Light()
        C.memmove(C.byref(l), C.byref(desc.lights[i]), C.sizeof(abi.Light))
        if l.radiance:
            m = np.ctypeslib.as_array(C.cast(l.radiance, C.POINTER(C.c_float)), (l.width * l.height * 16,)).copy()
        else:  # a sun/sky map: parameters only
            assert l.map == abi.LIGHTMAP_SUNSKY
            m = np.zeros(0, np.float32)
        l.radiance = None
        out.append((bytes(l), m.tobytes()))
    return out


SCENES = {
    "cornell": lambda: scenes.cornell(32),
    "materials_env": lambda: cases.materials_scene(res=24, env=True),
    "textured_uvs": lambda: cases.textured_scene(res=24),
    "node_textures": lambda: cases.node_texture_scene(res=16),
    "sun_sky": lambda: cases.sun_sky_scene(res=16),
    "divergent_small": lambda: scenes.divergent_scene(res=32, hm=16, n_tori=2, torus_tris=200, n_spheres=10,
                                                      sky=(16, 8)),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_round_trip_is_byte_exact(tmp_path, name):
    flat = SCENES[name]()
    path = tmp_path / (name + ".bgs")
    flat.save(path)
    with SceneFile(path) as f:
        d, e = flat.desc, f.desc
        for k in ("num_vertices", "num_triangles", "num_meshes", "num_shapes", "num_materials", "num_lights"):
            assert getattr(d, k) == getattr(e, k), k
        assert bytes(d.camera) == bytes(e.camera) and bytes(d.film) == bytes(e.film)
        assert np.array_equal(f.array("vertices", np.float32, 3).view(np.int32), flat.vertices.view(np.int32))
        assert np.array_equal(f.array("indices", np.int32, 3), flat.indices)
        for arr, cols in (("normals", 3), ("uvs", 2)):
            want = getattr(flat, arr)
            got = f.array(arr, np.float32, cols)
            assert (want is None) == (got is None), arr
            if want is not None:
                assert np.array_equal(got.view(np.int32), want.view(np.int32)), arr
        if len(flat.tri_mesh):
            assert np.array_equal(f.array("tri_mesh", np.int32, 1), flat.tri_mesh)
        assert _records(d.meshes, d.num_meshes, abi.Mesh) == _records(e.meshes, e.num_meshes, abi.Mesh)
        assert _records(d.shapes, d.num_shapes, abi.Shape) == _records(e.shapes, e.num_shapes, abi.Shape)
        assert _records(d.materials, d.num_materials, abi.Material) == \
            _records(e.materials, e.num_materials, abi.Material)
        assert _light_bytes(d) == _light_bytes(e)
        assert f.num_prims == flat.num_prims
        assert (d.num_tex_nodes, d.num_images) == (e.num_tex_nodes, e.num_images)
        assert _records(d.tex_nodes, d.num_tex_nodes, abi.TexNode) == \
            _records(e.tex_nodes, e.num_tex_nodes, abi.TexNode)
        for i in range(d.num_images):
            a, b = d.images[i], e.images[i]
            assert (a.width, a.height, a.channels) == (b.width, b.height, b.channels)
            n = a.width * a.height * a.channels
            assert np.ctypeslib.as_array(a.texels, (n,)).tobytes() == np.ctypeslib.as_array(b.texels, (n,)).tobytes()


@pytest.mark.parametrize("name", ["materials_env", "node_textures", "sun_sky"])
def test_oracle_renders_the_read_back_scene_identically(tmp_path, name):
    flat = {"materials_env": lambda: cases.materials_scene(res=16, env=True),
            "node_textures": lambda: cases.node_texture_scene(res=16),
            "sun_sky": lambda: cases.sun_sky_scene(res=16)}[name]()
    path = tmp_path / "m.bgs"
    flat.save(path)
    spec = JobSpec(spp_nu=2, spp_nv=2, max_depth=4, sample_depth=3, rr_start_depth=3, passes=1,
                   seed=9, variance=True)
    a, asq, ar = pyoracle.render(flat, spec, threads=1)
    with SceneFile(path) as f:
        b, bsq, br = pyoracle.render(f, spec, threads=1)
    assert ar == br
    assert np.array_equal(a, b) and np.array_equal(asq, bsq)


def test_malformed_files_are_refused(tmp_path):
    flat = cases.materials_scene(res=16, env=True)
    good = tmp_path / "good.bgs"
    flat.save(good)
    raw = good.read_bytes()
    assert len(raw) % 64 == 0

    def refused(data, what):
        p = tmp_path / "bad.bgs"
        p.write_bytes(data)
        with pytest.raises(abi.BgpuError) as e:
            SceneFile(p)
        assert e.value.code == abi.E_INVALID, what
        assert "bgpu_scene_read" in str(e.value), what

    refused(raw[:100], "shorter than the header")
    refused(raw[:-64], "truncated")
    refused(raw + b"\0" * 64, "trailing bytes")
    refused(b"XGPUSCN1" + raw[8:], "magic")
    bad = bytearray(raw)
    bad[8] = 0xEE  # version
    refused(bytes(bad), "version")
    bad = bytearray(raw)
    bad[40] ^= 1  # flags: a normals section the file does not hold
    refused(bytes(bad), "flags")
    bad = bytearray(raw)
    bad[48] ^= 1  # size_mesh
    refused(bytes(bad), "record size")
    with pytest.raises(abi.BgpuError):
        SceneFile(tmp_path / "missing.bgs")
    with pytest.raises(abi.BgpuError):
        flat.save(tmp_path / "no-such-dir" / "x.bgs")

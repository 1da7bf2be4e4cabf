"""The Haskell drop-in (bling_amd/haskell) against the C ABI.

GHC is not in this image, so the Haskell files are not compiled here.  What
can be checked is that they bind what include/bgpu.h declares: every entry
point has a `foreign import ccall` with the C function's number of
arguments, and every record's Storable instance reads and writes every
field of its C struct, in declaration order (hsc2hs takes the offsets from
the header itself)."""
import os
import re

import pytest

from bling_amd import abi

HS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bling_amd", "haskell",
                  "Graphics", "Bling")
FFI = os.path.join(HS, "GPU", "FFI.hsc")

STRUCTS = {
    "bgpu_shape": abi.Shape, "bgpu_texture": abi.Texture, "bgpu_material": abi.Material,
    "bgpu_mesh": abi.Mesh, "bgpu_light": abi.Light, "bgpu_camera": abi.Camera, "bgpu_film": abi.Film,
    "bgpu_scene_desc": abi.SceneDesc, "bgpu_rays": abi.Rays, "bgpu_hits": abi.Hits,
    "bgpu_job_desc": abi.JobDesc, "bgpu_stats": abi.Stats, "bgpu_bvh_info": abi.BvhInfo,
    "bgpu_profile": abi.Profile, "bgpu_tex_node": abi.TexNode, "bgpu_image": abi.Image,
}


def _ffi():
    with open(FFI) as f:
        return f.read()


def _imports(text):
    """symbol -> argument count of each foreign import."""
    out = {}
    for m in re.finditer(r'foreign import ccall (?:safe|unsafe) "(\w+)" \w+\s*::(.*?)(?=\nforeign import|\Z)',
                         text, re.S):
        sig = " ".join(m.group(2).split())
        out[m.group(1)] = sig.count("->")
    return out


def test_every_entry_point_is_imported_with_its_arity():
    lib = abi.load()
    imports = _imports(_ffi())
    assert sorted(imports) == sorted(abi.EXPORTS)
    for name, nargs in imports.items():
        argtypes = getattr(lib, name).argtypes
        assert nargs == len(argtypes or []), (name, nargs, argtypes)


@pytest.mark.parametrize("struct", sorted(STRUCTS))
def test_storable_instances_cover_every_field_in_order(struct):
    text = _ffi()
    fields = [f for f, _ in STRUCTS[struct]._fields_]
    start = text.index("  sizeOf _ = #{size %s}" % struct)
    inst = text[start:text.index("\n\n", start)]
    peek_part, poke_part = inst.split("\n  poke p")
    for part in (peek_part, poke_part):
        # arrays are reached through #{ptr ..}, scalars through #{peek ..} / #{poke ..}
        got = re.findall(r"#\{(?:peek|poke|ptr) %s, (\w+)\}" % struct, part)
        assert got == fields, (struct, got, fields)


def test_renderer_module_binds_the_reference_interfaces():
    """The renderer is a Renderer instance reporting PassDone per pass
    (Rendering.hs:60-78) and the parser hook reads the sampler renderer's
    syntax (IO/RendererParser.hs:56-73)."""
    with open(os.path.join(HS, "GPU.hs")) as f:
        gpu = f.read()
    with open(os.path.join(HS, "IO", "GpuParser.hs")) as f:
        parser = f.read()
    assert "instance Renderer GpuRenderer where" in gpu
    assert "PassDone" in gpu and "c_jobCancel" in gpu and "c_jobPoll" in gpu
    for kw in ('namedBlock) "sampled"', 'namedBlock) "sampler"', 'namedBlock) "integrator"',
               '"stratified"', '"path"', '"directLighting"', 'namedInt "maxDepth"', 'namedInt "sampleDepth"'):
        assert kw in parser, kw

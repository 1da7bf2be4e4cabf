"""The C ABI: libbgpu.so loads, exports every entry point include/bgpu.h
declares, the ctypes mirror has the C layout, and argument checks fail
loudly without touching the GPU."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from bling_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bgpu.h")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(bgpu_\w+)\s*\(", src, re.M)))


def test_header_declares_what_abi_mirrors():
    assert sorted(abi.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (bgpu_\w+)", out))
    assert set(declared_functions()) <= exported


STRUCTS = {
    "bgpu_shape": abi.Shape, "bgpu_material": abi.Material, "bgpu_texture": abi.Texture, "bgpu_mesh": abi.Mesh,
    "bgpu_light": abi.Light, "bgpu_camera": abi.Camera, "bgpu_film": abi.Film,
    "bgpu_scene_desc": abi.SceneDesc, "bgpu_rays": abi.Rays, "bgpu_hits": abi.Hits,
    "bgpu_job_desc": abi.JobDesc, "bgpu_stats": abi.Stats, "bgpu_bvh_info": abi.BvhInfo,
    "bgpu_profile": abi.Profile, "bgpu_tex_node": abi.TexNode, "bgpu_image": abi.Image,
}


def test_ctypes_layout_matches_c():
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bgpu.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-I" + os.path.dirname(HEADER), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(l.split() for l in out.strip().splitlines())
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(py, fname).offset, (cname, fname)


def test_null_arguments_fail_with_invalid():
    lib = abi.load()
    assert lib.bgpu_version() == 4
    h = C.c_void_p()
    assert lib.bgpu_scene_create(None, None, C.byref(h)) == abi.E_INVALID
    assert b"NULL" in lib.bgpu_last_error()
    assert lib.bgpu_intersect_closest(None, None, 0, None, 0) == abi.E_INVALID
    assert lib.bgpu_intersect_any(None, None, 0, None, 0) == abi.E_INVALID
    assert lib.bgpu_render(None, None, None, None) == abi.E_INVALID
    assert lib.bgpu_job_submit(None, None, C.byref(h)) == abi.E_INVALID
    assert lib.bgpu_job_wait(None, 0.0) == abi.E_INVALID
    assert lib.bgpu_stats_get(None, None) == abi.E_INVALID
    with pytest.raises(abi.BgpuError):
        abi.check(lib.bgpu_shutdown(None))

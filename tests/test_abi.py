"""The C ABI library loads and exports every function the headers declare.
No compute call is made here (this runs without a GPU)."""
import os
import re
import shutil
import subprocess

import pytest

import flgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fl|rtc)[A-Za-z0-9_]+)\s*\(", text)))


def test_all_declared_symbols_exported():
    L = flgpu.lib()
    names = declared("fl_gpu.h") + declared("fl_scenegen.h") + declared("fl_scenefile.h")
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(flgpu.EXPORTS) <= set(names)


def test_embree_subset_exported():
    """include/fl_rtc.h (the Embree 4 subset the reference's BVH calls) is
    implemented in the same library."""
    L = flgpu.lib()
    names = declared("fl_rtc.h")
    assert len(names) == 17
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version_and_no_device_behaviour():
    L = flgpu.lib()
    assert L.fl_abi_version() == 7 == flgpu.ABI_VERSION
    if L.fl_device_count() == 0:
        h = flgpu.C.c_void_p()
        assert L.fl_ctx_create(0, flgpu.C.byref(h)) == flgpu.FL_E_UNSUPPORTED
        assert not h.value


def test_struct_sizes_match_header():
    C = flgpu.C
    assert C.sizeof(flgpu.Material) == 64
    assert C.sizeof(flgpu.AreaLight) == 16
    assert C.sizeof(flgpu.DeltaLight) == 48
    assert C.sizeof(flgpu.BsdfQuery) == 112
    assert C.sizeof(flgpu.BsdfResult) == 56
    assert C.sizeof(flgpu.Camera) == 48
    assert C.sizeof(flgpu.Stats) == 176
    assert C.sizeof(flgpu.RenderParams) == 96


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof and field offsets of the ctypes mirrors equal the C header's,
    as gcc lays the structs out."""
    C = flgpu.C
    structs = {"fl_material": flgpu.Material, "fl_standard_surface": flgpu.StandardSurface,
               "fl_texture": flgpu.Texture, "fl_area_light": flgpu.AreaLight, "fl_delta_light": flgpu.DeltaLight,
               "fl_scene_desc": flgpu.SceneDesc, "fl_standard_query": flgpu.StandardQuery,
               "fl_bsdf_query": flgpu.BsdfQuery, "fl_bsdf_result": flgpu.BsdfResult,
               "fl_scene_info": flgpu.SceneInfo, "fl_camera": flgpu.Camera, "fl_render_params": flgpu.RenderParams,
               "fl_film": flgpu.Film, "fl_stats": flgpu.Stats, "fl_batch_item": flgpu.BatchItem,
               "fl_tune_result": flgpu.TuneResult, "fl_guide_config": flgpu.GuideConfig,
               "fl_develop_params": flgpu.DevelopParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fl_gpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%%s %%zu\\n", "%s", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%%s.%%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            assert int(got[cname + "." + f[0]]) == getattr(py, f[0]).offset, (cname, f[0])

"""The C ABI library loads and exports every function the headers declare.
No compute call is made here (this runs without a GPU)."""
import os
import re

import flgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fl|rtc)[A-Za-z0-9_]+)\s*\(", text)))


def test_all_declared_symbols_exported():
    L = flgpu.lib()
    names = declared("fl_gpu.h") + declared("fl_scenegen.h")
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
    assert L.fl_abi_version() == 2
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
    assert C.sizeof(flgpu.RenderParams) == 88

"""The C++ host layer (filianore-core_amd/host/filianore_gpu.h): the
Scene / PathIntegrator-shaped interface over the C ABI, exercised by the
example program built beside it."""
import json
import os
import subprocess

import numpy as np
import pytest

import flgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "filianore-core_amd", "host", "fl_render_example")
RTC_EXE = os.path.join(ROOT, "filianore-core_amd", "host", "fl_rtc_example")


def run_example(*args, exe=EXE):
    p = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    line = p.stdout.strip().splitlines()[-1] if p.stdout.strip() else "{}"
    return p.returncode, json.loads(line)


def test_example_built_and_reports_missing_gpu():
    assert os.access(EXE, os.X_OK), "build() did not produce the C++ example"
    if flgpu.lib().fl_device_count() > 0:
        pytest.skip("a GPU is present")
    rc, out = run_example()
    assert rc == 2 and out["ok"] is False and "no usable GPU" in out["error"]


@pytest.mark.gpu
def test_cpp_layer_matches_c_abi_render(tmp_path):
    """The C++ layer rebuilds the Cornell box as shared triangles in meshes;
    with the same seed and parameters its film equals the film rendered from
    the generator's indexed buffers through the C ABI, bit for bit."""
    film = str(tmp_path / "film.bin")
    rc, out = run_example(film, "64", "48", "16")
    assert rc == 0 and out["ok"], out
    assert out["mismatches"] == 0 and out["ray_hits"] == out["ray_occluded"] > 0
    got = np.fromfile(film, np.float32).reshape(48, 64, 3)
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=64, height=48)
    with flgpu.Context(0) as ctx, ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=16, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=7,
                                 flags=flgpu.FL_RENDER_GRAPH | flgpu.FL_RENDER_VNDF_PDF)
        ref, _, _ = sc.render(p)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_cpp_layer_standard_surface_matches_c_abi(tmp_path):
    """Material::standard_surface in the C++ layer: the Cornell box with
    StandardSurfaceMaterials built through it renders the same film, bit for
    bit, as the same materials given to the C ABI as fl_standard_surface."""
    from test_gpu_standard import standard_box

    film = str(tmp_path / "film.bin")
    rc, out = run_example(film, "48", "40", "16", "standard")
    assert rc == 0 and out["ok"], out
    got = np.fromfile(film, np.float32).reshape(40, 48, 3)
    gen = standard_box(48, 40)
    with flgpu.Context(0) as ctx, ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=16, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=7,
                                 flags=flgpu.FL_RENDER_GRAPH | flgpu.FL_RENDER_VNDF_PDF)
        ref, _, _ = sc.render(p)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    plain = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=48, height=40)
    with flgpu.Context(0) as ctx, ctx.scene(plain) as sc:
        other, _, _ = sc.render(p)
    assert not np.array_equal(other, ref)


@pytest.mark.gpu
def test_cpp_layer_textures_match_c_abi(tmp_path):
    """ImageMapTexture, Material::with_texture and BasicTriangle uvs in the
    C++ layer: the textured standard box renders the same film, bit for bit,
    as the same textures and uvs given to the C ABI."""
    from test_gpu_textures import textured_box

    gen = textured_box(48, 40)
    blob = b"".join(np.array([t.shape[1], t.shape[0]], np.uint32).tobytes() + np.ascontiguousarray(t).tobytes()
                    for t in gen.textures)
    texf = tmp_path / "textures.bin"
    texf.write_bytes(blob)
    film = str(tmp_path / "film.bin")
    rc, out = run_example(film, "48", "40", "16", "standard", str(texf))
    assert rc == 0 and out["ok"], out
    got = np.fromfile(film, np.float32).reshape(40, 48, 3)
    with flgpu.Context(0) as ctx, ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=16, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=7,
                                 flags=flgpu.FL_RENDER_GRAPH | flgpu.FL_RENDER_VNDF_PDF)
        ref, _, _ = sc.render(p)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_rtc_example_built_and_reports_missing_gpu():
    assert os.access(RTC_EXE, os.X_OK), "build() did not produce the Embree-API example"
    if flgpu.lib().fl_device_count() > 0:
        pytest.skip("a GPU is present")
    rc, out = run_example(exe=RTC_EXE)
    assert rc == 2 and out["ok"] is False and "no usable GPU" in out["error"]


@pytest.mark.gpu
def test_embree_pattern_matches_c_abi():
    """Code in the reference's Embree call pattern (one geometry per
    triangle, rtcIntersect1 per ray), built against include/fl_rtc.h: hits
    equal fl_intersect bit for bit with geomID as the triangle, Ng is the
    unnormalized face normal, rtcOccluded1 agrees, multi-triangle meshes map
    (geomID, primID) to the same triangles, and misuse reports errors."""
    rc, out = run_example("2048", exe=RTC_EXE)
    print(out)
    assert rc == 0 and out["ok"], out
    assert out["hits"] > 500 and out["mismatches"] == 0 and out["mesh_mismatches"] == 0
    assert out["occluded_mismatches"] == 0 and out["ng_mismatches"] == 0

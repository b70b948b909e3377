"""The scene wire format (include/fl_scenefile.h): a scene written and read
back is the same scene, byte for byte in every array and struct, and renders
the same film (the oracle here, the GPU under -m gpu); damaged files are
rejected with a reason."""
import ctypes as C
import os

import numpy as np
import pytest

import flgpu
import oracle
from test_gpu_lights import lit_box
from test_gpu_textures import textured_box


def rich_scene():
    """The textured standard-surface box (standard surfaces, five textures,
    uvs) with the lit box's delta lights and shadow colours."""
    gen = textured_box(24, 20)
    lb = lit_box(24, 20)
    gen.emission = lb.emission
    gen.n_emission = lb.n_emission
    gen.lights = lb.lights
    gen.n_lights = lb.n_lights
    gen.delta_lights = lb.delta_lights
    gen.n_delta_lights = lb.n_delta_lights
    gen.textures.append(np.zeros((0, 0, 3), np.uint8))   # an image that did not load
    return gen


def struct_bytes(arr, n):
    return b"".join(bytes(arr[i]) for i in range(n))


def test_round_trip_is_exact(tmp_path):
    gen = rich_scene()
    cam2 = flgpu.Camera()
    C.pointer(cam2)[0] = gen.camera
    cam2.width = 7
    path = tmp_path / "box.flscene"
    flgpu.save_scene_file(path, gen, [gen.camera, cam2])
    got = flgpu.load_scene_file(path, spp=gen.spp, depth=gen.depth)
    for name in ("vertices", "indices", "tri_material", "emission", "uvs"):
        a, b = np.asarray(getattr(gen, name)), np.asarray(getattr(got, name))
        assert a.dtype == b.dtype and np.array_equal(a.ravel(), b.ravel()), name
    assert struct_bytes(gen.materials, len(gen.materials)) == struct_bytes(got.materials, len(got.materials))
    for name, n in (("lights", gen.n_lights), ("delta_lights", gen.n_delta_lights),
                    ("standard_surfaces", gen.n_standard_surfaces)):
        assert n == getattr(got, "n_" + name) and n > 0, name
        assert struct_bytes(getattr(gen, name), n) == struct_bytes(getattr(got, name), n), name
    assert len(got.textures) == len(gen.textures) == 6
    for a, b in zip(gen.textures, got.textures):
        assert a.shape == b.shape and np.array_equal(a, b)
    assert [bytes(c) for c in got.cameras] == [bytes(gen.camera), bytes(cam2)]


def test_loaded_scene_renders_the_same_film(tmp_path):
    gen = rich_scene()
    path = tmp_path / "box.flscene"
    flgpu.save_scene_file(path, gen)
    got = flgpu.load_scene_file(path, depth=gen.depth)
    a, _, _ = oracle.Scene(gen).render(8, gen.depth, seed=5, threads=4)
    b, _, _ = oracle.Scene(got).render(8, gen.depth, seed=5, threads=4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_generated_scene_round_trip(tmp_path):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=2, width=16, height=16)
    path = tmp_path / "spheres.flscene"
    flgpu.save_scene_file(path, gen)
    got = flgpu.load_scene_file(path)
    assert np.array_equal(gen.vertices, got.vertices) and np.array_equal(gen.indices, got.indices)
    assert got.uvs is None and got.textures == []
    # header 16 B + sections of 16-byte headers and 8-byte padded payloads
    assert os.path.getsize(path) >= 16 + gen.vertices.nbytes + gen.indices.nbytes


def split_sections(raw):
    """(header, [(tag, count, payload)]) of a scene file: 16-byte header,
    then sections of a 16-byte header (tag, count, bytes) and a payload
    padded to 8 bytes."""
    head, n = bytes(raw[:16]), int(np.frombuffer(raw[12:16], np.uint32)[0])
    off, out = 16, []
    for _ in range(n):
        tag, count = np.frombuffer(raw[off:off + 8], np.uint32)
        nb = int(np.frombuffer(raw[off + 8:off + 16], np.uint64)[0])
        out.append((int(tag), int(count), bytes(raw[off + 16:off + 16 + nb])))
        off += 16 + nb + (8 - nb % 8) % 8
    return head, out


def join_sections(head, secs):
    raw = bytearray(head[:12]) + np.uint32(len(secs)).tobytes()
    for tag, count, payload in secs:
        raw += np.array([tag, count], np.uint32).tobytes() + np.uint64(len(payload)).tobytes() + payload
        raw += bytes((8 - len(payload) % 8) % 8)
    return raw


def tag_of(name):
    return int(np.frombuffer(name.encode(), np.uint32)[0])


@pytest.mark.parametrize("damage", ["magic", "version", "truncate", "size", "index", "texture_ref", "short_tmat",
                                    "short_uvs", "duplicate", "texture_wrap"])
def test_damaged_files_are_rejected(tmp_path, damage):
    gen = rich_scene()
    path = tmp_path / "box.flscene"
    flgpu.save_scene_file(path, gen)
    raw = bytearray(path.read_bytes())
    if damage == "magic":
        raw[0] = ord("X")
    elif damage == "version":
        raw[8] = 9
    elif damage == "truncate":
        raw = raw[: len(raw) // 2]
    elif damage == "size":
        raw[28] ^= 0x04            # the first section's payload size (VERT)
    elif damage == "index":
        # first triangle's first vertex index -> far out of range (INDX follows VERT)
        nv = gen.vertices.shape[0]
        off = 16 + 16 + ((nv * 12 + 7) // 8) * 8 + 16
        raw[off:off + 4] = np.uint32(10 ** 9).tobytes()
    elif damage == "texture_ref":
        gen.materials[0].kind = flgpu.FL_MAT_LAMBERT
        gen.materials[0].texture = 99
        flgpu.save_scene_file(path, gen)
        raw = bytearray(path.read_bytes())
    elif damage in ("short_tmat", "short_uvs"):
        # a per-triangle / per-vertex section one entry short, its own size
        # consistent with its count (the reader would run past it)
        head, secs = split_sections(raw)
        t, elem = (tag_of("TMAT"), 4) if damage == "short_tmat" else (tag_of("UVS_"), 8)
        secs = [(tg, c - 1, p[:-elem]) if tg == t else (tg, c, p) for tg, c, p in secs]
        raw = join_sections(head, secs)
    elif damage == "duplicate":
        head, secs = split_sections(raw)
        secs.append(next(x for x in secs if x[0] == tag_of("MATL")))
        raw = join_sections(head, secs)
    elif damage == "texture_wrap":
        # 3 w h = 2 * 2^64 + 1399: computed in 64 bits it wraps to 1399, so a
        # size check without an overflow test would accept 1399 texel bytes
        head, secs = split_sections(raw)
        w, h = 4294571377, 2863575501
        assert (3 * w * h) % 2 ** 64 == 1399
        secs.append((tag_of("TEXR"), 1, np.array([w, h], np.uint32).tobytes() + bytes(1399)))
        raw = join_sections(head, secs)
    path.write_bytes(bytes(raw))
    with pytest.raises(flgpu.FlError) as e:
        flgpu.load_scene_file(path)
    assert "fl_scene_file_read" in str(e.value)


def test_missing_file_is_reported(tmp_path):
    with pytest.raises(flgpu.FlError, match="cannot open"):
        flgpu.load_scene_file(tmp_path / "nothing.flscene")


@pytest.mark.gpu
def test_loaded_scene_renders_the_same_film_on_the_gpu(ctx, tmp_path):
    gen = rich_scene()
    path = tmp_path / "box.flscene"
    flgpu.save_scene_file(path, gen)
    got = flgpu.load_scene_file(path, depth=gen.depth)
    films = []
    for g in (gen, got):
        with ctx.scene(g) as sc:
            f, _, st = sc.render(flgpu.default_params(spp=16, depth=gen.depth, seed=3))
        assert st.nan_samples == 0
        films.append(f)
    assert np.array_equal(films[0].view(np.uint32), films[1].view(np.uint32))

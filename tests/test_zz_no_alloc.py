"""fl_intersect_device_noalloc against fl_intersect_device: bit-identical
hits on two calls in a row, the second with fewer rays, for closest-hit and
any-hit (emitters skipped) queries, on a BVH of the default width and on a
deeper width-2 BVH whose spill area makes the fixed scratch hold fewer
persistent blocks. The test's device arrays and stream come from the HIP
runtime the library itself links (the no-alloc calls run on a stream the
test creates)."""
import ctypes as C
import sys

import numpy as np
import pytest

import flgpu
from raysets import primary_rays, secondary_rays

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(ctx):
    """The library's HIP runtime (already loaded by libfl_gpu.so)."""
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    h.test_stream = stream
    yield h
    h.hipStreamDestroy(stream)


def dev_alloc(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p


def rays_for(gen, sc, n_primary):
    """Camera rays of the scene plus cosine-diffuse rays from their hits."""
    cam = gen.camera
    W = 512
    H = max(1, n_primary // W)
    prim = primary_rays(cam, W, H)
    t, tri, _ = sc.intersect(prim)
    sec = secondary_rays(gen.triangles(), prim, t, tri, prim.shape[0], seed=9)
    return np.concatenate([prim, sec]).astype(np.float32)


def run(ctx, hip, sc, rays, flags, noalloc):
    n = rays.shape[0]
    d_rays, d_hits = dev_alloc(hip, rays.nbytes), dev_alloc(hip, n * 16)
    try:
        sentinel = np.full((n, 4), 0xFFFFFFFF, np.uint32)        # stale hits would show
        assert hip.hipMemcpy(d_rays, rays.ctypes.data, rays.nbytes, H2D) == 0
        assert hip.hipMemcpy(d_hits, sentinel.ctypes.data, sentinel.nbytes, H2D) == 0
        L = ctx.L
        if noalloc:
            rc = L.fl_intersect_device_noalloc(sc.h, d_rays, d_hits, n, flags, hip.test_stream)
            assert rc == 0, L.fl_last_error(ctx.h)
            assert hip.hipStreamSynchronize(hip.test_stream) == 0
        else:
            rc = L.fl_intersect_device(sc.h, d_rays, d_hits, n, flags)
            assert rc == 0, L.fl_last_error(ctx.h)
        out = np.zeros((n, 4), np.uint32)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, d_hits, out.nbytes, D2H) == 0
        return out
    finally:
        hip.hipFree(d_rays)
        hip.hipFree(d_hits)


@pytest.mark.parametrize("width", [8, 2])
@pytest.mark.parametrize("flags", [flgpu.FL_CLOSEST_HIT, flgpu.FL_ANY_HIT | flgpu.FL_SKIP_LIGHTS])
def test_noalloc_matches_device_entry_point(ctx, hip, width, flags):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=4, width=64, height=64)
    gen.bvh_width = width
    with ctx.scene(gen) as sc:
        info = sc.info()
        rays = rays_for(gen, sc, 1 << 17)
        for n in (rays.shape[0], rays.shape[0] // 3 + 17):   # second call smaller
            print("case width=%d flags=%d n=%d depth=%d" % (width, flags, n, info.depth), flush=True)
            sys.stdout.flush()
            want = run(ctx, hip, sc, rays[:n], flags, noalloc=False)
            got = run(ctx, hip, sc, rays[:n], flags, noalloc=True)
            assert np.array_equal(got, want)
            hit = want[:, 1] != flgpu.FL_NO_HIT
            assert hit.mean() > 0.3


def test_noalloc_rejects_counters(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=8, height=8)
    with ctx.scene(gen) as sc:
        rc = ctx.L.fl_intersect_device_noalloc(sc.h, C.c_void_p(16), C.c_void_p(16), 4, flgpu.FL_COUNTERS, None)
        assert rc == flgpu.FL_E_INVALID
        assert b"FL_COUNTERS" in ctx.L.fl_last_error(ctx.h)

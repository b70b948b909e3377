"""Small analytic scenes shared by CPU and GPU tests."""
import numpy as np

import flgpu


def lambert(rgb):
    m = flgpu.Material()
    m.kind = flgpu.FL_MAT_LAMBERT
    m.rgb[:] = list(rgb)
    m.weight = 1.0
    return m


def none_material():
    m = flgpu.Material()
    m.kind = flgpu.FL_MAT_NONE
    return m


def quad(c, u, v, sx, sy):
    """create_quad's two triangles (creators/shapecreator.cpp:7-38) mapped by
    (x, y) -> c + x*sx*u + y*sy*v; geometric normal = -(u x v)."""
    c, u, v = (np.asarray(a, np.float64) for a in (c, u, v))
    loc = [(1, -1), (-1, -1), (-1, 1), (1, -1), (-1, 1), (1, 1)]
    return np.array([c + x * sx * u + y * sy * v for x, y in loc], np.float32)


def light_over_floor(albedo=(0.5, 0.5, 0.5), half=0.5, height=1.0, intensity=5.0, spd=None):
    """A large Lambert floor (y=0, normal +y) and a square emitter of half-side
    `half` at y=height facing down, plus a 1x1-pixel camera looking straight
    down at the point below the emitter's centre."""
    floor = quad((0, 0, 0), (1, 0, 0), (0, 0, 1), 50, 50)   # u x v = (0,-1,0) -> normal +y
    lamp = quad((0, height, 0), (1, 0, 0), (0, 0, -1), half, half)  # normal -y
    verts = np.concatenate([floor, lamp])
    idx = np.arange(12, dtype=np.uint32).reshape(4, 3)
    mats = [lambert(albedo), none_material()]
    tri_mat = np.array([0, 0, 1, 1], np.uint32)
    if spd is None:
        spd = np.ones(flgpu.FL_SPD_N, np.float32)
    cam = flgpu.Camera()
    cam.position[:] = [0.0, 3.0, 0.0]
    cam.look_at[:] = [0.0, 0.0, 0.0]
    cam.up[:] = [0.0, 0.0, -1.0]
    cam.vfov_deg = 0.01
    cam.width = cam.height = 1
    return flgpu.GenSceneData.from_arrays(verts, idx, tri_mat, mats, [(2, 0, intensity), (3, 0, intensity)],
                                          spd[None, :], cam, spp=1, depth=0)


def rect_form_factor_center(a_half, b_half, h):
    """Irradiance / radiance at a point directly below the centre of a parallel
    a x b rectangle at distance h: pi * sum of four corner form factors
    F = 1/(2 pi) [X/sqrt(1+X^2) atan(Y/sqrt(1+X^2)) + Y/sqrt(1+Y^2) atan(X/sqrt(1+Y^2))]."""
    X, Y = a_half / h, b_half / h
    F = (X / np.sqrt(1 + X * X) * np.arctan(Y / np.sqrt(1 + X * X)) +
         Y / np.sqrt(1 + Y * Y) * np.arctan(X / np.sqrt(1 + Y * Y))) / (2 * np.pi)
    return np.pi * 4 * F

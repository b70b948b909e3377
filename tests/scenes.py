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


def delta_light(kind, position=(0, 0, 0), direction=(0, 0, 0), emission=0, intensity=1.0, cos_cone=0.0,
                cos_penumbra=0.0, shadow_emission=0):
    d = flgpu.DeltaLight()
    d.type = kind
    d.position[:] = list(position)
    d.direction[:] = list(direction)
    d.emission, d.intensity = emission, intensity
    d.cos_cone, d.cos_penumbra = cos_cone, cos_penumbra
    d.shadow_emission = shadow_emission
    return d


FLOOR_CAM_HEIGHT = 3.0
FLOOR_VIEW_HALF = 2.0   # the camera sees the floor square [-2, 2]^2


def delta_floor(delta_lights, spds, albedo=(0.6, 0.5, 0.4), occluder=False, W=32, area_lights=()):
    """A Lambert floor (y = 0, normal +y) lit by delta lights, seen by a W x W
    camera straight above it at y = 3 whose view covers [-2, 2]^2 of the
    floor. With occluder, a 0.5 x 4 Lambert slab at y = 1, x in [1.5, 2]
    (outside the view at that height) casts shadows into the view. spds:
    (n, FL_SPD_N) emission table (light colours and shadow colours)."""
    floor = quad((0, 0, 0), (1, 0, 0), (0, 0, 1), 50, 50)
    parts = [floor]
    mats = [lambert(albedo)]
    tri_mat = [0, 0]
    if occluder:
        parts.append(quad((1.75, 1.0, 0), (1, 0, 0), (0, 0, 1), 0.25, 2.0))   # faces +y
        parts.append(quad((1.75, 1.0, 0), (1, 0, 0), (0, 0, -1), 0.25, 2.0))  # faces -y
        tri_mat += [0, 0, 0, 0]
    lights = []
    if area_lights:
        mats.append(none_material())
    for (c, half, intensity, em) in area_lights:
        parts.append(quad(c, (1, 0, 0), (0, 0, -1), half, half))          # faces -y
        n = sum(len(p) for p in parts) // 3
        tri_mat += [len(mats) - 1, len(mats) - 1]
        lights += [(n - 2, em, intensity), (n - 1, em, intensity)]
    verts = np.concatenate(parts)
    idx = np.arange(len(verts), dtype=np.uint32).reshape(-1, 3)
    cam = flgpu.Camera()
    cam.position[:] = [0.0, FLOOR_CAM_HEIGHT, 0.0]
    cam.look_at[:] = [0.0, 0.0, 0.0]
    cam.up[:] = [0.0, 0.0, -1.0]
    cam.vfov_deg = float(2 * np.degrees(np.arctan(FLOOR_VIEW_HALF / FLOOR_CAM_HEIGHT)))
    cam.width = cam.height = W
    return flgpu.GenSceneData.from_arrays(verts, idx, np.array(tri_mat, np.uint32), mats, lights,
                                          np.asarray(spds, np.float32).reshape(-1, flgpu.FL_SPD_N), cam, spp=1,
                                          depth=0, delta_lights=list(delta_lights))


def floor_points(cam, sub=8):
    """Floor points (y = 0) seen by each pixel of the straight-down camera of
    delta_floor, on a sub x sub grid inside every pixel: (H, W, sub*sub, 3)."""
    from raysets import camera_basis
    pos, llc, hor, ver = camera_basis(cam)
    W, H = cam.width, cam.height
    j = (np.arange(sub) + 0.5) / sub
    jx, jy = np.meshgrid(j, j, indexing="xy")
    xs = (np.arange(W)[None, :, None] + jx.ravel()[None, None, :])
    ys = (np.arange(H)[:, None, None] + jy.ravel()[None, None, :])
    s = 1.0 - xs / W
    t = 1.0 - ys / H
    d = llc + hor * s[..., None] + ver * t[..., None] - pos
    d = np.broadcast_to(d, (H, W, sub * sub, 3))
    tt = pos[1] / -d[..., 1]
    return pos + d * tt[..., None]


def checker_texture(w=5, h=4, seed=0):
    """A w x h RGB8 image of distinct random colours (rows top to bottom)."""
    rng = np.random.default_rng(seed)
    return rng.integers(20, 250, size=(h, w, 3), dtype=np.uint8)


def textured_floor(texture, spds, with_uvs=True, standard=False, W=32, intensity=2.0):
    """delta_floor's floor lit straight from above by a directional light,
    its albedo an image texture: uv = ((x + 2) / 4, (z + 2) / 4), so the
    camera's view [-2, 2]^2 spans the image once (outside it the lookup
    clamps). standard: a StandardSurface with only an Oren-Nayar lobe of
    roughness 0 (Lambert) reads the texture as kd instead of a Lambert
    material."""
    light = delta_light(flgpu.FL_LIGHT_DIRECTIONAL, direction=(0.0, -1.0, 0.0), intensity=intensity)
    gen = delta_floor([light], spds, W=W)
    if standard:
        s = flgpu.StandardSurface()
        s.kd_weight, s.kd_roughness, s.ks_ior = 1.0, 0.0, 1.5
        s.kd[:] = (0.3, 0.3, 0.3)
        s.kd_texture = 1
        gen.materials[0].kind = flgpu.FL_MAT_STANDARD
        gen.materials[0].standard = 0
        gen.standard_surfaces = (flgpu.StandardSurface * 1)(s)
        gen.n_standard_surfaces = 1
    else:
        gen.materials[0].texture = 1
    gen.textures = [texture]
    if with_uvs:
        v = gen.vertices
        gen.uvs = np.stack([(v[:, 0] + 2.0) / 4.0, (v[:, 2] + 2.0) / 4.0], -1).astype(np.float32)
    return gen


def texel_of(texture, u, v):
    """ImageMapTexture::evaluate's texel index (imagemap.cpp:38-55, with j
    clamped) for uv arrays, as float32 arithmetic gives it."""
    h, w = texture.shape[:2]
    i = np.trunc(np.float32(u) * np.float32(w)).astype(np.int64)
    j = np.trunc((np.float32(1.0) - np.float32(v)) * np.float32(h) - np.float32(0.001)).astype(np.int64)
    return np.clip(i, 0, w - 1), np.clip(j, 0, h - 1)

"""ctypes binding of libfl_gpu.so (include/fl_gpu.h, include/fl_scenegen.h).

This is the Python view of the C ABI used by the tests and the benchmark; the
product itself is the shared library.  Loading fails loudly when the library
is missing: there is no CPU fallback for the render path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfl_gpu.so")
# the struct layouts below are those of this version of include/fl_gpu.h
ABI_VERSION = 7

FL_OK = 0
FL_E_INVALID, FL_E_HIP, FL_E_OOM, FL_E_UNSUPPORTED, FL_E_OVERFLOW, FL_E_NUMERIC = -1, -2, -3, -4, -5, -6
FL_MAT_NONE, FL_MAT_LAMBERT, FL_MAT_CONDUCTOR, FL_MAT_DIELECTRIC, FL_MAT_STANDARD = 0, 1, 2, 3, 4
FL_FRESNEL_DIELECTRIC, FL_FRESNEL_METALLIC, FL_FRESNEL_WAVELENGTH_METALLIC = 0, 1, 2
FL_CLOSEST_HIT, FL_ANY_HIT, FL_SKIP_LIGHTS, FL_COUNTERS = 0, 1, 2, 4
FL_SPECTRAL_HERO, FL_SPECTRAL_PARITY = 0, 1
FL_RENDER_GUIDED, FL_RENDER_COUNTERS, FL_RENDER_VNDF_PDF, FL_RENDER_GRAPH = 1, 2, 4, 8
FL_RENDER_TRAIN_GUIDE, FL_RENDER_SORT_SHADING, FL_RENDER_TIMING, FL_RENDER_NORMALS = 16, 32, 64, 128
FL_RENDER_NEE_ONE = 256
FL_LIGHT_POINT, FL_LIGHT_SPOT, FL_LIGHT_DIRECTIONAL = 1, 2, 3
FL_COLOR_SRGB, FL_COLOR_ACES2065_1 = 0, 1
FL_TONEMAP_NONE, FL_TONEMAP_ACES_BASIC, FL_TONEMAP_REINHARD, FL_TONEMAP_FILMIC = 1, 2, 3, 4
FL_GAMMA_NONE, FL_GAMMA_REC_709, FL_GAMMA_1_8, FL_GAMMA_2_2, FL_GAMMA_4_0 = 1, 2, 3, 4, 5
FL_NO_HIT = 0xFFFFFFFF
FL_SPD_N = 471
FL_GEN_CORNELL, FL_GEN_SPHERES, FL_GEN_BATCH, FL_GEN_GUIDED, FL_GEN_TERRAIN, FL_GEN_GUIDED_SKYLIGHT = 0, 1, 2, 3, 4, 5
BSDF_REFLECTION, BSDF_TRANSMISSION, BSDF_DIFFUSE, BSDF_GLOSSY, BSDF_SPECULAR = 1, 2, 4, 8, 16

f3 = C.c_float * 3
f4 = C.c_float * 4


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("rgb", f3), ("rgb_t", f3), ("roughness", C.c_float),
                ("anisotropic", C.c_float), ("weight", C.c_float), ("metal", C.c_int32), ("ior", C.c_float),
                ("cauchy_b", C.c_float), ("standard", C.c_uint32), ("texture", C.c_uint32), ("spd_r", C.c_uint16),
                ("spd_t", C.c_uint16)]


class StandardSurface(C.Structure):
    _fields_ = [("kd_weight", C.c_float), ("kd", f3), ("kd_roughness", C.c_float),
                ("ks_weight", C.c_float), ("ks", f3), ("ks_roughness", C.c_float), ("ks_anisotropic", C.c_float),
                ("fresnel_type", C.c_int32), ("ks_ior", C.c_float), ("metal", C.c_int32),
                ("sheen_weight", C.c_float), ("sheen_color", f3), ("sheen_roughness", C.c_float),
                ("kt_weight", C.c_float), ("kt", f3),
                ("coat_weight", C.c_float), ("coat", f3), ("coat_ior", C.c_float), ("coat_gloss", C.c_float),
                ("thin_film_thickness", C.c_float), ("thin_film_ior", C.c_float),
                ("thin_walled", C.c_int32), ("kd_texture", C.c_uint32), ("ks_texture", C.c_uint32),
                ("sheen_texture", C.c_uint32), ("kt_texture", C.c_uint32), ("coat_texture", C.c_uint32)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_uint8))]


class StandardQuery(C.Structure):
    _fields_ = [("kd", f4), ("ks", f4), ("kt", f4), ("sheen", f4), ("coat", f4), ("eta", f4), ("k", f4), ("lam", f4),
                ("wo", f3), ("wi", f3), ("u", C.c_float * 2), ("flags", C.c_uint32)]


class AreaLight(C.Structure):
    _fields_ = [("tri", C.c_uint32), ("emission", C.c_uint32), ("intensity", C.c_float),
                ("shadow_emission", C.c_uint32)]


class DeltaLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", f3), ("direction", f3), ("emission", C.c_uint32),
                ("intensity", C.c_float), ("cos_cone", C.c_float), ("cos_penumbra", C.c_float),
                ("shadow_emission", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("n_vertices", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32)), ("n_triangles", C.c_uint32),
                ("tri_material", C.POINTER(C.c_uint32)),
                ("materials", C.POINTER(Material)), ("n_materials", C.c_uint32),
                ("lights", C.POINTER(AreaLight)), ("n_lights", C.c_uint32),
                ("emission_spd", C.POINTER(C.c_float)), ("n_emission", C.c_uint32),
                ("delta_lights", C.POINTER(DeltaLight)), ("n_delta_lights", C.c_uint32),
                ("standard_surfaces", C.POINTER(StandardSurface)), ("n_standard_surfaces", C.c_uint32),
                ("textures", C.POINTER(Texture)), ("n_textures", C.c_uint32), ("uvs", C.POINTER(C.c_float)),
                ("bvh_width", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("n_triangles", C.c_uint32), ("n_nodes", C.c_uint32), ("n_leaf_tris", C.c_uint32),
                ("depth", C.c_uint32), ("node_bytes", C.c_uint64), ("tri_bytes", C.c_uint64),
                ("build_ms", C.c_double), ("bounds_min", f3), ("bounds_max", f3)]


class Camera(C.Structure):
    _fields_ = [("position", f3), ("look_at", f3), ("up", f3), ("vfov_deg", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("diffuse_depth", C.c_int32), ("reflection_depth", C.c_int32),
                ("transmission_depth", C.c_int32), ("rr_threshold", C.c_float), ("rr_start", C.c_uint32),
                ("spectral_mode", C.c_uint32), ("flags", C.c_uint32), ("seed", C.c_uint64),
                ("sample_offset", C.c_uint32), ("max_paths", C.c_uint32), ("stack_depth", C.c_uint32),
                ("trav_budget", C.c_uint32), ("shade_waves", C.c_uint32), ("trav_waves", C.c_uint32),
                ("guide", C.c_void_p),
                ("guide_mix", C.c_float),
                ("guide_train_fraction", C.c_float), ("refill_lanes", C.c_uint32), ("trav_prefetch", C.c_uint32),
                ("guide_bounces", C.c_uint32)]


class Film(C.Structure):
    _fields_ = [("xyz_sum", C.POINTER(C.c_float)), ("xyz_sqsum", C.POINTER(C.c_float)),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_generate", C.c_double), ("ms_extend", C.c_double),
                ("ms_shade", C.c_double), ("ms_shadow", C.c_double), ("ms_resolve", C.c_double),
                ("ms_guide", C.c_double),
                ("samples", C.c_uint64), ("rays_extension", C.c_uint64), ("rays_shadow", C.c_uint64),
                ("nodes_visited", C.c_uint64), ("tris_tested", C.c_uint64),
                ("shadow_nodes_visited", C.c_uint64), ("shadow_tris_tested", C.c_uint64),
                ("launches_extend", C.c_uint32), ("launches_shade", C.c_uint32), ("launches_shadow", C.c_uint32),
                ("pad0", C.c_uint32),
                ("nan_samples", C.c_uint64), ("stack_spills", C.c_uint64),
                ("passes", C.c_uint32), ("max_bounce", C.c_uint32),
                ("paths_per_pass", C.c_uint32), ("stack_depth", C.c_uint32),
                ("ms_infer", C.c_double), ("guide_queries", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("film", C.POINTER(Film)), ("spp", C.c_uint32),
                ("pad", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("max_paths", C.c_uint32), ("stack_depth", C.c_uint32), ("trav_waves", C.c_uint32),
                ("trav_budget", C.c_uint32), ("shade_waves", C.c_uint32), ("sort_shading", C.c_uint32),
                ("ms_best", C.c_double), ("ms_default", C.c_double), ("n_tried", C.c_uint32),
                ("refill_lanes", C.c_uint32), ("node_width", C.c_uint32), ("trav_prefetch", C.c_uint32)]

    def apply(self, params):
        """Copies the tuned values into fl_render_params."""
        params.max_paths, params.stack_depth, params.trav_waves = self.max_paths, self.stack_depth, self.trav_waves
        params.trav_budget, params.shade_waves = self.trav_budget, self.shade_waves
        params.refill_lanes = self.refill_lanes
        params.trav_prefetch = self.trav_prefetch
        if self.sort_shading:
            params.flags |= FL_RENDER_SORT_SHADING
        else:
            params.flags &= ~FL_RENDER_SORT_SHADING
        return params


class GuideConfig(C.Structure):
    _fields_ = [("hidden", C.c_uint32), ("layers", C.c_uint32), ("lobes", C.c_uint32), ("bands", C.c_uint32),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("batch", C.c_uint32), ("seed", C.c_uint64), ("scene_min", f3), ("scene_max", f3),
                ("steps_per_pass", C.c_uint32), ("train_rounds", C.c_uint32)]


class DevelopParams(C.Structure):
    _fields_ = [("color_space", C.c_uint32), ("tone_map", C.c_uint32), ("gamma", C.c_uint32), ("pad", C.c_uint32)]


class BsdfQuery(C.Structure):
    _fields_ = [("kind", C.c_int32), ("alpha_x", C.c_float), ("alpha_y", C.c_float), ("eta", f4), ("k", f4),
                ("r", f4), ("t", f4), ("wo", f3), ("wi", f3), ("u", C.c_float * 2), ("flags", C.c_uint32)]


class BsdfResult(C.Structure):
    _fields_ = [("f_eval", f4), ("pdf_eval", C.c_float), ("f_sample", f4), ("pdf_sample", C.c_float),
                ("wi_sample", f3), ("sampled_type", C.c_uint32)]


class GenOptions(C.Structure):
    _fields_ = [("subdiv", C.c_int32), ("n_spheres", C.c_int32), ("grid_x", C.c_int32), ("grid_y", C.c_int32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("scene_index", C.c_uint32),
                ("camera_index", C.c_uint32), ("seed", C.c_uint64)]


class GenScene(C.Structure):
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("n_vertices", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32)), ("n_triangles", C.c_uint32),
                ("tri_material", C.POINTER(C.c_uint32)),
                ("materials", C.POINTER(Material)), ("n_materials", C.c_uint32),
                ("lights", C.POINTER(AreaLight)), ("n_lights", C.c_uint32),
                ("emission_spd", C.POINTER(C.c_float)), ("n_emission", C.c_uint32),
                ("camera", Camera), ("spp", C.c_uint32), ("depth", C.c_int32)]


# exported symbols of include/fl_gpu.h, include/fl_scenegen.h and include/fl_scenefile.h
EXPORTS = [
    "fl_abi_version", "fl_device_count", "fl_probe_copy", "fl_ctx_create", "fl_ctx_destroy", "fl_last_error",
    "fl_scene_create", "fl_scene_destroy", "fl_scene_get_info", "fl_scene_download_bvh",
    "fl_intersect", "fl_intersect_device", "fl_intersect_device_noalloc", "fl_render", "fl_render_batch", "fl_autotune", "fl_autotune_batch",
    "fl_guide_create", "fl_guide_destroy", "fl_guide_train_step", "fl_guide_forward",
    "fl_guide_get_params", "fl_guide_set_params", "fl_guide_gradient", "fl_guide_stats",
    "fl_bsdf_query_batch", "fl_bsdf_query_standard", "fl_sampler_draw", "fl_develop_film", "fl_develop",
    "fl_gen_default_options", "fl_gen_config", "fl_gen_free", "fl_gen_describe", "fl_gen_blackbody",
    "fl_scene_file_write", "fl_scene_file_read", "fl_scene_file_error", "fl_scene_file_desc",
    "fl_scene_file_cameras", "fl_scene_file_free",
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libfl_gpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.fl_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.fl_ctx_destroy.argtypes = [vp]
        L.fl_probe_copy.argtypes = [vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        L.fl_last_error.argtypes = [vp]
        L.fl_last_error.restype = C.c_char_p
        L.fl_scene_create.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(vp)]
        L.fl_scene_destroy.argtypes = [vp]
        L.fl_scene_get_info.argtypes = [vp, C.POINTER(SceneInfo)]
        L.fl_scene_download_bvh.argtypes = [vp, vp, C.c_size_t, vp, C.c_size_t]
        L.fl_intersect.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
        L.fl_intersect_device.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32]
        L.fl_intersect_device_noalloc.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint32, vp]
        L.fl_render.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderParams), C.POINTER(Film), C.POINTER(Stats)]
        L.fl_render_batch.argtypes = [vp, C.POINTER(BatchItem), C.c_uint32, C.POINTER(RenderParams),
                                      C.POINTER(Stats)]
        L.fl_autotune.argtypes = [vp, C.POINTER(Camera), C.POINTER(RenderParams), C.POINTER(TuneResult)]
        L.fl_autotune_batch.argtypes = [vp, C.POINTER(BatchItem), C.c_uint32, C.POINTER(RenderParams),
                                        C.POINTER(TuneResult)]
        L.fl_guide_create.argtypes = [vp, C.POINTER(GuideConfig), C.POINTER(vp)]
        L.fl_guide_destroy.argtypes = [vp]
        L.fl_guide_train_step.argtypes = [vp, vp, C.c_uint32, C.c_int, C.POINTER(C.c_float)]
        L.fl_guide_forward.argtypes = [vp, vp, vp, C.c_uint32]
        L.fl_guide_get_params.argtypes = [vp, vp, C.c_uint32, C.POINTER(C.c_uint32)]
        L.fl_guide_set_params.argtypes = [vp, vp, C.c_uint32]
        L.fl_guide_gradient.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, C.POINTER(C.c_float)]
        L.fl_guide_stats.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.fl_bsdf_query_batch.argtypes = [vp, C.POINTER(BsdfQuery), C.POINTER(BsdfResult), C.c_uint32]
        L.fl_develop_film.argtypes = [vp, C.c_uint32, C.POINTER(DevelopParams), vp, vp]
        L.fl_bsdf_query_standard.argtypes = [vp, C.POINTER(StandardSurface), C.POINTER(StandardQuery),
                                             C.POINTER(BsdfResult), C.c_uint32]
        L.fl_develop.argtypes = [vp, vp, C.c_uint32, C.POINTER(DevelopParams), vp, vp]
        L.fl_sampler_draw.argtypes = [vp, C.c_uint64, vp, C.c_uint32, C.c_uint32, vp]
        L.fl_gen_default_options.argtypes = [C.POINTER(GenOptions)]
        L.fl_gen_config.argtypes = [C.c_int, C.POINTER(GenOptions), C.POINTER(GenScene)]
        L.fl_gen_free.argtypes = [C.POINTER(GenScene)]
        L.fl_gen_describe.argtypes = [C.POINTER(GenScene), C.POINTER(SceneDesc)]
        L.fl_gen_blackbody.argtypes = [C.c_float, C.POINTER(C.c_float)]
        L.fl_scene_file_write.argtypes = [C.c_char_p, C.POINTER(SceneDesc), C.POINTER(Camera), C.c_uint32]
        L.fl_scene_file_read.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.fl_scene_file_error.restype = C.c_char_p
        L.fl_scene_file_desc.argtypes = [vp]
        L.fl_scene_file_desc.restype = C.POINTER(SceneDesc)
        L.fl_scene_file_cameras.argtypes = [vp, C.POINTER(C.POINTER(Camera))]
        L.fl_scene_file_cameras.restype = C.c_uint32
        L.fl_scene_file_free.argtypes = [vp]
        if L.fl_abi_version() != ABI_VERSION:
            raise RuntimeError("libfl_gpu.so has ABI %d, these bindings need %d; rebuild it"
                               % (L.fl_abi_version(), ABI_VERSION))
        _lib = L
    return _lib


_hip_rt = None


def _hip():
    """The HIP runtime libfl_gpu.so links (already loaded with it): device
    buffers for the *_device entry points."""
    global _hip_rt
    if _hip_rt is None:
        lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip_rt = h
    return _hip_rt


class FlError(RuntimeError):
    pass


def _check(rc, ctx=None, what=""):
    if rc != FL_OK:
        msg = lib().fl_last_error(ctx).decode() if ctx else ""
        raise FlError("%s failed (%d): %s" % (what, rc, msg))


# ------------------------------------------------------------------ scenes --
class GenSceneData:
    """A generated scene as numpy arrays (owned copies)."""

    def __init__(self, config, **opts):
        L = lib()
        o = GenOptions()
        L.fl_gen_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        g = GenScene()
        _check(L.fl_gen_config(config, C.byref(o), C.byref(g)), what="fl_gen_config")
        try:
            nv, nt = g.n_vertices, g.n_triangles
            self.vertices = np.ctypeslib.as_array(g.vertices, (nv * 3,)).copy().reshape(nv, 3)
            self.indices = np.ctypeslib.as_array(g.indices, (nt * 3,)).copy().reshape(nt, 3)
            self.tri_material = np.ctypeslib.as_array(g.tri_material, (nt,)).copy()
            self.materials = (Material * g.n_materials)(*[g.materials[i] for i in range(g.n_materials)])
            self.lights = (AreaLight * max(1, g.n_lights))(*[g.lights[i] for i in range(g.n_lights)])
            self.n_lights = g.n_lights
            self.emission = np.ctypeslib.as_array(g.emission_spd, (g.n_emission * FL_SPD_N,)).copy() \
                if g.n_emission else np.zeros(FL_SPD_N, np.float32)
            self.n_emission = g.n_emission
            self.delta_lights = (DeltaLight * 1)()
            self.n_delta_lights = 0
            self.standard_surfaces = (StandardSurface * 1)()
            self.n_standard_surfaces = 0
            self.textures = []
            self.uvs = None
            self.camera = Camera()
            C.pointer(self.camera)[0] = g.camera
            self.spp = g.spp
            self.depth = g.depth
        finally:
            L.fl_gen_free(C.byref(g))

    @classmethod
    def from_arrays(cls, vertices, indices, tri_material, materials, lights, emission, camera, spp=16, depth=8,
                    delta_lights=(), standard_surfaces=(), textures=(), uvs=None):
        """A scene from explicit buffers: materials is a list of Material,
        lights a list of (tri, emission_index, intensity[, shadow_emission]),
        delta_lights a list of DeltaLight, emission an (n, FL_SPD_N) array,
        camera a Camera, textures a list of (height, width, 3) uint8 images
        (rows top to bottom), uvs an (n_vertices, 2) array or None."""
        self = cls.__new__(cls)
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        self.tri_material = np.ascontiguousarray(tri_material, np.uint32)
        self.materials = (Material * len(materials))(*materials)
        ls = [AreaLight(*(tuple(l) + (0,) * (4 - len(l)))) for l in lights]
        self.lights = (AreaLight * max(1, len(ls)))(*ls)
        self.n_lights = len(ls)
        self.delta_lights = (DeltaLight * max(1, len(delta_lights)))(*delta_lights)
        self.n_delta_lights = len(delta_lights)
        self.standard_surfaces = (StandardSurface * max(1, len(standard_surfaces)))(*standard_surfaces)
        self.n_standard_surfaces = len(standard_surfaces)
        self.textures = list(textures)
        self.uvs = uvs
        em = np.ascontiguousarray(emission, np.float32).reshape(-1, FL_SPD_N)
        self.emission = em.ravel().copy()
        self.n_emission = em.shape[0]
        self.camera = camera
        self.spp = spp
        self.depth = depth
        return self

    @property
    def n_triangles(self):
        return self.indices.shape[0]

    def desc(self):
        d = SceneDesc()
        self._keep = [np.ascontiguousarray(self.vertices, np.float32), np.ascontiguousarray(self.indices, np.uint32),
                      np.ascontiguousarray(self.tri_material, np.uint32), np.ascontiguousarray(self.emission, np.float32)]
        v, i, m, e = self._keep
        d.vertices = v.ctypes.data_as(C.POINTER(C.c_float))
        d.n_vertices = v.shape[0]
        d.indices = i.ctypes.data_as(C.POINTER(C.c_uint32))
        d.n_triangles = i.shape[0]
        d.tri_material = m.ctypes.data_as(C.POINTER(C.c_uint32))
        d.materials = C.cast(self.materials, C.POINTER(Material))
        d.n_materials = len(self.materials)
        d.lights = C.cast(self.lights, C.POINTER(AreaLight))
        d.n_lights = self.n_lights
        d.emission_spd = e.ctypes.data_as(C.POINTER(C.c_float))
        d.n_emission = self.n_emission
        d.delta_lights = C.cast(self.delta_lights, C.POINTER(DeltaLight))
        d.n_delta_lights = self.n_delta_lights
        d.standard_surfaces = C.cast(self.standard_surfaces, C.POINTER(StandardSurface))
        d.n_standard_surfaces = self.n_standard_surfaces
        texs = [np.ascontiguousarray(t, np.uint8) for t in getattr(self, "textures", [])]
        for t in texs:
            if t.ndim != 3 or t.shape[2] != 3:
                raise ValueError("a texture is a (height, width, 3) uint8 image")
        self._keep_tex = texs
        self._tex = (Texture * max(1, len(texs)))(
            *[Texture(t.shape[1], t.shape[0], t.ctypes.data_as(C.POINTER(C.c_uint8))) for t in texs])
        d.textures = C.cast(self._tex, C.POINTER(Texture))
        d.n_textures = len(texs)
        d.bvh_width = getattr(self, "bvh_width", 0)
        uvs = getattr(self, "uvs", None)
        if uvs is not None:
            self._keep_uv = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
            if self._keep_uv.shape[0] != v.shape[0]:
                raise ValueError("uvs: one (u, v) per vertex")
            d.uvs = self._keep_uv.ctypes.data_as(C.POINTER(C.c_float))
        return d

    def triangles(self):
        """(n, 3, 3) float32 vertex positions per triangle."""
        return self.vertices[self.indices]


def default_params(spp=16, depth=8, mode=FL_SPECTRAL_HERO, seed=1, flags=0):
    p = RenderParams()
    p.spp = spp
    p.diffuse_depth = p.reflection_depth = p.transmission_depth = depth
    p.rr_threshold = 1.0
    p.rr_start = 3
    p.spectral_mode = mode
    p.flags = flags
    p.seed = seed
    p.guide_mix = 0.5
    p.guide_train_fraction = 0.15
    return p


class Context:
    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.fl_ctx_create(device, C.byref(h))
        if rc != FL_OK:
            raise FlError("fl_ctx_create failed (%d): no usable GPU" % rc)
        self.h = h

    def close(self):
        if self.h:
            self.L.fl_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def err(self):
        return self.L.fl_last_error(self.h).decode()

    def scene(self, gen):
        return Scene(self, gen)

    def autotune(self, scene, params, camera=None):
        r = TuneResult()
        cam = camera if camera is not None else scene.gen.camera
        _check(self.L.fl_autotune(scene.h, C.byref(cam), C.byref(params), C.byref(r)), self.h, "fl_autotune")
        return r

    def render_batch(self, entries, params, want_film=True):
        """entries: [(Scene, Camera or None, spp or 0)]; one batched render.
        Returns ([(sum, sq)] per entry, Stats)."""
        n = len(entries)
        items = (BatchItem * n)()
        keep = []
        for k, (sc, cam, spp) in enumerate(entries):
            cam = cam if cam is not None else sc.gen.camera
            items[k].scene = sc.h
            items[k].camera = cam
            items[k].spp = spp
            if want_film:
                W, H = cam.width, cam.height
                s = np.zeros((H, W, 3), np.float32)
                q = np.zeros((H, W, 3), np.float32)
                f = Film(s.ctypes.data_as(C.POINTER(C.c_float)), q.ctypes.data_as(C.POINTER(C.c_float)), W, H)
                keep.append((s, q, f))
                items[k].film = C.pointer(f)
        st = Stats()
        _check(self.L.fl_render_batch(self.h, items, n, C.byref(params), C.byref(st)), self.h, "fl_render_batch")
        return [(s, q) for s, q, _ in keep], st

    def guide(self, seed=1, lr=0.0, batch=0, scene_min=(0, 0, 0), scene_max=(1, 1, 1), steps_per_pass=0,
              train_rounds=0):
        return Guide(self, seed=seed, lr=lr, batch=batch, scene_min=scene_min, scene_max=scene_max,
                     steps_per_pass=steps_per_pass, train_rounds=train_rounds)

    def probe_copy(self, nbytes=1 << 30, iters=5):
        """Device-to-device streaming copy; returns (read + write) GB/s."""
        g = C.c_double(0.0)
        _check(self.L.fl_probe_copy(self.h, nbytes, iters, C.byref(g)), self.h, "fl_probe_copy")
        return g.value

    def develop(self, xyz, color_space=FL_COLOR_SRGB, tone_map=FL_TONEMAP_NONE, gamma=FL_GAMMA_NONE):
        """fl_develop of n pixel means (n x 3 XYZ): (rgb n x 3 float32, rgba8 n x 4 uint8)."""
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        rgb = np.zeros(x.shape, np.float32)
        u8 = np.zeros((x.shape[0], 4), np.uint8)
        p = DevelopParams(color_space, tone_map, gamma, 0)
        _check(self.L.fl_develop(self.h, x.ctypes.data, x.shape[0], C.byref(p), rgb.ctypes.data, u8.ctypes.data),
               self.h, "fl_develop")
        return rgb, u8

    def develop_film(self, item, width, height, color_space=FL_COLOR_SRGB, tone_map=FL_TONEMAP_NONE,
                     gamma=FL_GAMMA_NONE):
        """fl_develop_film of item `item` of the last render: (rgb H x W x 3, rgba8 H x W x 4)."""
        rgb = np.zeros((height, width, 3), np.float32)
        u8 = np.zeros((height, width, 4), np.uint8)
        p = DevelopParams(color_space, tone_map, gamma, 0)
        _check(self.L.fl_develop_film(self.h, item, C.byref(p), rgb.ctypes.data, u8.ctypes.data), self.h,
               "fl_develop_film")
        return rgb, u8

    def bsdf_queries_standard(self, material, queries):
        """fl_bsdf_query_standard: the full StandardSurface BSDF of one
        material, over a list of StandardQuery (returns a list of BsdfResult)
        or a numpy array of dtype(StandardQuery) (returns an array of
        dtype(BsdfResult))."""
        if isinstance(queries, np.ndarray):
            q = np.ascontiguousarray(queries, np.dtype(StandardQuery))
            r = np.zeros(q.shape[0], np.dtype(BsdfResult))
            _check(self.L.fl_bsdf_query_standard(self.h, C.byref(material), q.ctypes.data_as(C.POINTER(StandardQuery)),
                                                 r.ctypes.data_as(C.POINTER(BsdfResult)), q.shape[0]),
                   self.h, "fl_bsdf_query_standard")
            return r
        n = len(queries)
        q = (StandardQuery * n)(*queries)
        r = (BsdfResult * n)()
        _check(self.L.fl_bsdf_query_standard(self.h, C.byref(material), q, r, n), self.h, "fl_bsdf_query_standard")
        return list(r)

    def sampler_draw(self, seed, keys, n_dims):
        """fl_sampler_draw: keys (n, 3) uint32 (pixel, sample, bounce) -> (n, n_dims) float32."""
        k = np.ascontiguousarray(keys, np.uint32).reshape(-1, 3)
        out = np.zeros((k.shape[0], n_dims), np.float32)
        _check(self.L.fl_sampler_draw(self.h, C.c_uint64(seed), k.ctypes.data, k.shape[0], n_dims, out.ctypes.data),
               self.h, "fl_sampler_draw")
        return out

    def bsdf_queries(self, queries):
        n = len(queries)
        q = (BsdfQuery * n)(*queries)
        r = (BsdfResult * n)()
        _check(self.L.fl_bsdf_query_batch(self.h, q, r, n), self.h, "fl_bsdf_query_batch")
        return r


def save_scene_file(path, gen, cameras=None):
    """fl_scene_file_write: gen's scene description and its cameras (default:
    gen.camera) to path."""
    cams = [gen.camera] if cameras is None else list(cameras)
    arr = (Camera * max(1, len(cams)))(*cams)
    d = gen.desc()
    if lib().fl_scene_file_write(os.fsencode(path), C.byref(d), arr, len(cams)) != FL_OK:
        raise FlError("fl_scene_file_write: " + lib().fl_scene_file_error().decode())


def load_scene_file(path, spp=16, depth=8):
    """fl_scene_file_read: the file's scene as a GenSceneData (owned copies),
    its first camera as .camera and all of them as .cameras."""
    L = lib()
    h = C.c_void_p()
    if L.fl_scene_file_read(os.fsencode(path), C.byref(h)) != FL_OK:
        raise FlError("fl_scene_file_read: " + L.fl_scene_file_error().decode())
    try:
        d = L.fl_scene_file_desc(h).contents
        cp = C.POINTER(Camera)()
        nc = L.fl_scene_file_cameras(h, C.byref(cp))
        cams = []
        for k in range(nc):
            c = Camera()
            C.pointer(c)[0] = cp[k]
            cams.append(c)

        def arr(ptr, n, dtype, shape):
            return np.ctypeslib.as_array(ptr, (n,)).copy().astype(dtype).reshape(shape) if n else \
                np.zeros(shape, dtype)

        nv, nt = d.n_vertices, d.n_triangles
        verts = arr(d.vertices, 3 * nv, np.float32, (nv, 3))
        idx = arr(d.indices, 3 * nt, np.uint32, (nt, 3))
        tmat = arr(d.tri_material, nt, np.uint32, (nt,)) if d.tri_material else np.zeros(nt, np.uint32)
        mats = [d.materials[i] for i in range(d.n_materials)]
        lights = [(d.lights[i].tri, d.lights[i].emission, d.lights[i].intensity, d.lights[i].shadow_emission)
                  for i in range(d.n_lights)]
        em = arr(d.emission_spd, d.n_emission * FL_SPD_N, np.float32, (d.n_emission, FL_SPD_N))
        textures = []
        for j in range(d.n_textures):
            t = d.textures[j]
            if t.rgb and t.width and t.height:
                textures.append(np.ctypeslib.as_array(t.rgb, (t.height * t.width * 3,)).copy()
                                .reshape(t.height, t.width, 3))
            else:
                textures.append(np.zeros((0, 0, 3), np.uint8))
        uvs = arr(d.uvs, 2 * nv, np.float32, (nv, 2)) if d.uvs else None
        gen = GenSceneData.from_arrays(verts, idx, tmat, mats, lights, em if d.n_emission else
                                       np.zeros((1, FL_SPD_N), np.float32),
                                       cams[0] if cams else Camera(), spp=spp, depth=depth,
                                       delta_lights=[d.delta_lights[i] for i in range(d.n_delta_lights)],
                                       standard_surfaces=[d.standard_surfaces[i]
                                                          for i in range(d.n_standard_surfaces)],
                                       textures=textures, uvs=uvs)
        if not d.n_emission:
            gen.n_emission = 0
        if not d.tri_material:
            gen.tri_material = np.zeros(nt, np.uint32)
        gen.cameras = cams
        return gen
    finally:
        L.fl_scene_file_free(h)


class Scene:
    def __init__(self, ctx, gen):
        self.ctx = ctx
        self.gen = gen
        d = gen.desc()
        h = C.c_void_p()
        _check(ctx.L.fl_scene_create(ctx.h, C.byref(d), C.byref(h)), ctx.h, "fl_scene_create")
        self.h = h

    def close(self):
        if self.h:
            self.ctx.L.fl_scene_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        i = SceneInfo()
        _check(self.ctx.L.fl_scene_get_info(self.h, C.byref(i)), self.ctx.h, "fl_scene_get_info")
        return i

    def bvh(self):
        i = self.info()
        nodes = np.zeros(i.node_bytes, np.uint8)
        tris = np.zeros(max(1, i.tri_bytes), np.uint8)
        _check(self.ctx.L.fl_scene_download_bvh(self.h, nodes.ctypes.data, nodes.nbytes, tris.ctypes.data,
                                                tris.nbytes), self.ctx.h, "fl_scene_download_bvh")
        return nodes.reshape(-1, 128), tris[: i.tri_bytes].reshape(-1, 48)

    def intersect(self, rays, flags=FL_CLOSEST_HIT, counters=False):
        """rays: (n, 8) float32 [ox, oy, oz, tmin, dx, dy, dz, tmax] -> hits (n, 4) [t, tri(u32), u, v]."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        c = (C.c_uint64 * 2)()
        f = flags | (FL_COUNTERS if counters else 0)
        _check(self.ctx.L.fl_intersect(self.h, rays.ctypes.data, hits.ctypes.data, n, f, c), self.ctx.h,
               "fl_intersect")
        t = hits[:, 0].copy()
        tri = hits[:, 1].view(np.uint32).copy()
        uv = hits[:, 2:4].copy()
        if counters:
            return t, tri, uv, (int(c[0]), int(c[1]))
        return t, tri, uv

    def intersect_resident(self, rays, flags=FL_CLOSEST_HIT, reps=5):
        """fl_intersect_device on rays copied once to device memory: returns
        (t, tri, uv) and the wall time per call in ms over `reps` calls
        queued back to back (PCIe excluded)."""
        import time
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hip = _hip()
        d_r, d_h = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(d_r), rays.nbytes) or hip.hipMalloc(C.byref(d_h), n * 16):
            raise FlError("hipMalloc failed")
        try:
            if hip.hipMemcpy(d_r, rays.ctypes.data, rays.nbytes, 1) or hip.hipDeviceSynchronize():
                raise FlError("upload failed")
            _check(self.ctx.L.fl_intersect_device(self.h, d_r, d_h, n, flags), self.ctx.h, "fl_intersect_device")
            hip.hipDeviceSynchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                _check(self.ctx.L.fl_intersect_device(self.h, d_r, d_h, n, flags), self.ctx.h, "fl_intersect_device")
            if hip.hipDeviceSynchronize():
                raise FlError("fl_intersect_device failed")
            ms = (time.perf_counter() - t0) / reps * 1e3
            hits = np.zeros((n, 4), np.float32)
            if hip.hipMemcpy(hits.ctypes.data, d_h, hits.nbytes, 2):
                raise FlError("download failed")
        finally:
            hip.hipFree(d_r)
            hip.hipFree(d_h)
        return hits[:, 0].copy(), hits[:, 1].view(np.uint32).copy(), hits[:, 2:4].copy(), ms

    def render(self, params, camera=None, want_film=True):
        cam = camera if camera is not None else self.gen.camera
        W, H = cam.width, cam.height
        st = Stats()
        film = None
        if want_film:
            s = np.zeros((H, W, 3), np.float32)
            q = np.zeros((H, W, 3), np.float32)
            film = Film(s.ctypes.data_as(C.POINTER(C.c_float)), q.ctypes.data_as(C.POINTER(C.c_float)), W, H)
        rc = self.ctx.L.fl_render(self.h, C.byref(cam), C.byref(params), C.byref(film) if film else None,
                                  C.byref(st))
        _check(rc, self.ctx.h, "fl_render")
        if want_film:
            return s, q, st
        return None, None, st


GUIDE_PARAMS = 5 * (64 * 64 + 64)
ITEM_SEED_STRIDE = 0x9E3779B97F4A7C15   # item k of a batch renders with seed + k * stride (mod 2^64)


class Guide:
    """The learned vMF-mixture proposal (fl_guide_*)."""

    def __init__(self, ctx, seed=1, lr=0.0, batch=0, scene_min=(0, 0, 0), scene_max=(1, 1, 1), steps_per_pass=0,
                 train_rounds=0):
        self.ctx = ctx
        cfg = GuideConfig()
        cfg.seed, cfg.lr, cfg.batch, cfg.steps_per_pass = seed, lr, batch, steps_per_pass
        cfg.train_rounds = train_rounds
        cfg.scene_min = f3(*scene_min)
        cfg.scene_max = f3(*scene_max)
        h = C.c_void_p()
        _check(ctx.L.fl_guide_create(ctx.h, C.byref(cfg), C.byref(h)), ctx.h, "fl_guide_create")
        self.h = h

    def close(self):
        if self.h:
            self.ctx.L.fl_guide_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((x.shape[0], 40), np.float32)
        _check(self.ctx.L.fl_guide_forward(self.h, x.ctypes.data, out.ctypes.data, x.shape[0]), self.ctx.h,
               "fl_guide_forward")
        return out

    def gradient(self, records):
        r = np.ascontiguousarray(records, np.float32)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = C.c_float(0)
        _check(self.ctx.L.fl_guide_gradient(self.h, r.ctypes.data, r.shape[0], g.ctypes.data, GUIDE_PARAMS,
                                            C.byref(loss)), self.ctx.h, "fl_guide_gradient")
        return loss.value, g

    def train_step(self, records):
        r = np.ascontiguousarray(records, np.float32)
        loss = C.c_float(0)
        rc = self.ctx.L.fl_guide_train_step(self.h, r.ctypes.data, r.shape[0], 0, C.byref(loss))
        return rc, loss.value

    def params(self):
        P = np.zeros(3 * GUIDE_PARAMS, np.float32)
        n = C.c_uint32(0)
        _check(self.ctx.L.fl_guide_get_params(self.h, P.ctypes.data, P.size, C.byref(n)), self.ctx.h,
               "fl_guide_get_params")
        return P[:GUIDE_PARAMS], P[GUIDE_PARAMS:2 * GUIDE_PARAMS], P[2 * GUIDE_PARAMS:]

    def set_params(self, P):
        P = np.ascontiguousarray(P, np.float32)
        _check(self.ctx.L.fl_guide_set_params(self.h, P.ctypes.data, P.size), self.ctx.h, "fl_guide_set_params")

    def stats(self):
        steps, loss, div = C.c_uint32(0), C.c_float(0), C.c_int(0)
        _check(self.ctx.L.fl_guide_stats(self.h, C.byref(steps), C.byref(loss), C.byref(div)), self.ctx.h,
               "fl_guide_stats")
        return steps.value, loss.value, div.value

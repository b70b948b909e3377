"""StandardSurfaceMaterial test helpers: materials that switch on each lobe of
compute_scattering_functions (shading/materials/standardsurface.cpp:72-159)
alone and in stacks, and vectorised fl_standard_query arrays."""
import numpy as np

import flgpu

QDT = np.dtype(flgpu.StandardQuery)
RDT = np.dtype(flgpu.BsdfResult)

# per-lane spectra of a query (lane k): reflectances, forge-like eta/k, wavelengths
LANES = {
    "kd": (0.55, 0.62, 0.70, 0.81),
    "ks": (0.90, 0.75, 0.60, 0.95),
    "kt": (0.85, 0.90, 0.95, 0.70),
    "sheen": (0.40, 0.65, 0.30, 0.90),
    "coat": (0.95, 0.90, 0.85, 0.99),
    "eta": (0.18, 0.45, 1.10, 2.40),
    "k": (3.4, 2.9, 4.1, 5.5),
    "lam": (412.0, 487.0, 563.0, 671.0),
}


def surface(**kw):
    """A StandardSurface with every lobe off; keyword arguments set fields."""
    s = flgpu.StandardSurface()
    s.ks_ior = 1.5
    s.coat_ior = 1.5
    s.thin_film_ior = 1.33
    s.kd[:] = s.ks[:] = s.kt[:] = s.sheen_color[:] = s.coat[:] = (1.0, 1.0, 1.0)
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(s, k)[:] = v
        else:
            setattr(s, k, v)
    return s


# One material per lobe kind, then stacks (names say which lobes)
MATERIALS = {
    "oren": dict(kd_weight=0.8, kd_roughness=0.5),
    "difftrans": dict(kd_weight=0.7, kd_roughness=0.3, thin_walled=1),
    "conductor": dict(ks_weight=0.9, ks_roughness=0.5, ks_anisotropic=0.4,
                      fresnel_type=flgpu.FL_FRESNEL_WAVELENGTH_METALLIC),
    "fblend_dielectric": dict(kd_weight=0.6, kd_roughness=0.4, ks_weight=0.8, ks_roughness=0.45,
                              fresnel_type=flgpu.FL_FRESNEL_DIELECTRIC, ks_ior=1.6),
    "fblend_metallic": dict(kd_weight=0.5, kd_roughness=0.2, ks_weight=0.9, ks_roughness=0.55, ks_anisotropic=0.6,
                            fresnel_type=flgpu.FL_FRESNEL_METALLIC),
    "thinfilm": dict(ks_roughness=0.5, thin_film_thickness=320.0, thin_film_ior=1.33, ks_ior=1.5),
    "sheen": dict(sheen_weight=0.9, sheen_roughness=0.45),
    "coat": dict(coat_weight=0.8, coat_gloss=0.0, coat_ior=1.5),
    "oren+conductor+thinfilm+coat": dict(kd_weight=0.5, kd_roughness=0.3, ks_weight=0.7, ks_roughness=0.5,
                                          fresnel_type=flgpu.FL_FRESNEL_WAVELENGTH_METALLIC,
                                          thin_film_thickness=250.0, coat_weight=0.4, coat_gloss=0.2),
    "fblend+thinfilm+sheen+coat": dict(kd_weight=0.6, kd_roughness=0.5, ks_weight=0.6, ks_roughness=0.35,
                                       ks_anisotropic=0.3, fresnel_type=flgpu.FL_FRESNEL_DIELECTRIC,
                                       thin_film_thickness=410.0, thin_film_ior=1.45, sheen_weight=0.5,
                                       sheen_roughness=0.6, coat_weight=0.5, coat_gloss=0.5, coat_ior=1.4),
    "difftrans+fspec+coat": dict(kd_weight=0.5, kd_roughness=0.6, ks_weight=0.8, kt_weight=0.9, ks_ior=1.45,
                                 fresnel_type=flgpu.FL_FRESNEL_WAVELENGTH_METALLIC, thin_walled=1,
                                 coat_weight=0.3, coat_gloss=0.4),
    "fspec": dict(ks_weight=1.0, kt_weight=1.0, ks_ior=1.5),
}

# materials whose every lobe samples the density its pdf reports (the sheen
# lobe does not: EstevezSheenDistribution samples cosine-weighted half vectors
# but reports D * cos as their density)
MATCHED = ["oren", "difftrans", "conductor", "fblend_dielectric", "fblend_metallic", "thinfilm", "coat",
           "oren+conductor+thinfilm+coat"]


def material(name):
    return surface(**MATERIALS[name])


def has_specular_lobe(m):
    """FresnelSpecularBXDF is in the stack when ksweight > 0 and ktweight > 0
    (standardsurface.cpp:113-117)."""
    return m.ks_weight > 0 and m.kt_weight > 0


def n_lobes(m):
    """Number of lobes compute_scattering_functions adds (standardsurface.cpp:88-158)."""
    wl = m.fresnel_type == flgpu.FL_FRESNEL_WAVELENGTH_METALLIC
    n = int((m.kd_weight > 0 and m.ks_weight <= 0) or (m.kd_weight > 0 and m.ks_weight > 0 and wl))
    n += int(m.ks_weight > 0)
    n += int(m.thin_film_thickness > 0) + int(m.sheen_weight > 0) + int(m.coat_weight > 0)
    return n


def queries(wo, wi=None, u=None, lanes=None, vndf=False):
    """fl_standard_query array for n directions wo (and wi, u), the per-lane
    spectra of LANES (overridden by lanes)."""
    wo = np.atleast_2d(np.asarray(wo, np.float64))
    n = wo.shape[0]
    q = np.zeros(n, QDT)
    L = dict(LANES)
    if lanes:
        L.update(lanes)
    for k, v in L.items():
        q[k] = np.broadcast_to(np.asarray(v, np.float32), (n, 4))
    q["wo"] = wo
    q["wi"] = np.broadcast_to(np.asarray((0, 0, 1) if wi is None else wi, np.float64), (n, 3))
    q["u"] = np.broadcast_to(np.asarray((0.5, 0.5) if u is None else u, np.float64), (n, 2))
    q["flags"] = flgpu.FL_RENDER_VNDF_PDF if vndf else 0
    return q


def unit_rows(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_dirs(rng, n, zmin=0.02):
    v = unit_rows(rng.normal(size=(n, 3)))
    small = np.abs(v[:, 2]) < zmin
    v[small, 2] = np.sign(v[small, 2] + 1e-9) * zmin
    return unit_rows(v)


def sphere_grid(nt, nph):
    """Midpoint grid over the whole sphere in (theta, phi): directions and
    solid-angle weights."""
    th = (np.arange(nt) + 0.5) * np.pi / nt
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    w = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nph)).ravel()
    return d, w


def strata(n):
    """n x n stratified (midpoint) sample grid on [0,1)^2."""
    g = (np.arange(n) + 0.5) / n
    A, B = np.meshgrid(g, g, indexing="ij")
    return np.stack([A.ravel(), B.ravel()], -1)

"""Byte-exact numpy layouts of the reference's POD scene types.

These mirror ``include/render_kernel.h:9-94`` of the reference (TriangleData
96 B, CameraData 56 B, EmitterData 32 B, SpotlightData 48 B, BsdfData 68 B) so
that flattened scene arrays can be handed to the C ABI (``include/bte_api.h``)
without repacking.
"""
import numpy as np

FLOAT3 = ("<f4", (3,))
FLOAT2 = ("<f4", (2,))
FLOAT4 = ("<f4", (4,))

TRIANGLE = np.dtype([("v0", FLOAT3), ("v1", FLOAT3), ("v2", FLOAT3),
                     ("n0", FLOAT3), ("n1", FLOAT3), ("n2", FLOAT3),
                     ("uv0", FLOAT2), ("uv1", FLOAT2), ("uv2", FLOAT2)])
CAMERA = np.dtype([("origin", FLOAT3), ("forward", FLOAT3), ("right", FLOAT3), ("up", FLOAT3),
                   ("fovYRadians", "<f4"), ("aspect", "<f4")])
EMITTER = np.dtype([("triIndexOffset", "<i4"), ("triCount", "<i4"), ("cdfOffset", "<i4"),
                    ("radiance", FLOAT3), ("areaSum", "<f4"), ("powerWeight", "<f4")])
SPOTLIGHT = np.dtype([("position", FLOAT3), ("direction", FLOAT3), ("radiance", FLOAT3),
                      ("intensity", "<f4"), ("innerConeCosine", "<f4"), ("outerConeCosine", "<f4")])
BSDF = np.dtype([("type", "<i4"), ("p0", FLOAT4), ("p1", FLOAT4), ("p2", FLOAT4), ("p3", FLOAT4)])

assert TRIANGLE.itemsize == 96
assert CAMERA.itemsize == 56
assert EMITTER.itemsize == 32
assert SPOTLIGHT.itemsize == 48
assert BSDF.itemsize == 68

# BsdfType (include/render_kernel.h:79-85)
BSDF_DIFFUSE = 1
BSDF_DIELECTRIC = 2
BSDF_MIRROR = 3
BSDF_MICROFACET = 4
BSDF_DISNEY = 5


def bsdf_diffuse(albedo):
    b = np.zeros((), BSDF)
    b["type"] = BSDF_DIFFUSE
    b["p0"][:3] = albedo
    return b


def bsdf_mirror(tint=(1.0, 1.0, 1.0)):
    b = np.zeros((), BSDF)
    b["type"] = BSDF_MIRROR
    b["p0"][:3] = [t if t > 0 else 1.0 for t in tint]
    return b


def bsdf_dielectric(int_ior=1.5, ext_ior=1.0):
    b = np.zeros((), BSDF)
    b["type"] = BSDF_DIELECTRIC
    b["p1"][:2] = (int_ior, ext_ior)
    return b


def bsdf_microfacet(kd, alpha=0.1, int_ior=1.5, ext_ior=1.0):
    """Packing of src/main.cpp:646-651 (ks = 1 - max(kd))."""
    b = np.zeros((), BSDF)
    b["type"] = BSDF_MICROFACET
    kd = np.asarray(kd, np.float32)
    ks = np.float32(1.0) - np.float32(max(kd[0], kd[1], kd[2]))
    b["p0"][:3] = kd
    b["p1"] = (alpha, int_ior, ext_ior, ks)
    return b


def bsdf_disney(base_color=(0.5, 0.5, 0.5), roughness=0.1, metallic=0.0, specular=0.5,
                spec_trans=0.0, spec_tint=0.0, sheen=0.0, sheen_tint=0.5, subsurface=0.0,
                anisotropic=0.0, clearcoat=0.0, clearcoat_gloss=1.0, eta=1.5):
    """Packing of src/main.cpp:652-657; defaults of include/scene.h:67-82."""
    b = np.zeros((), BSDF)
    b["type"] = BSDF_DISNEY
    b["p0"] = (*base_color, roughness)
    b["p1"] = (metallic, specular, spec_trans, spec_tint)
    b["p2"] = (sheen, sheen_tint, subsurface, anisotropic)
    b["p3"] = (clearcoat, clearcoat_gloss, eta, 0.0)
    return b

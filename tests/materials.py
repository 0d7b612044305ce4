"""Scenes whose materials take the OpenPBR parameters across their ranges:
anisotropic GGX, partial metalness and specular weight, coloured specular,
varied IOR, tangent-space normal maps at several scales, and specular colour,
specular weight and emission textures, on the Cornell
box geometry. Shared by the CPU test of the kernel source and the GPU parity
tests.
"""
import numpy as np

import ptcore


def random_openpbr_world(seed, image_w=128, image_h=128, atlas_size=16, huge_roughness=False):
    """Cornell geometry whose seven surface materials are drawn at random
    (the emitter keeps its emission). Every parameter the shading reads
    varies: trace.slang:369-375 (anisotropic alpha), 1097-1181 (BSDF setup),
    1026-1039 (normal map). huge_roughness: three surfaces get a specular
    roughness of 3e8 (GGX alpha near 9e16, beyond the 2^56 bound under which
    the E_dielectric estimate takes its restricted-domain normalisation)."""
    rng = np.random.default_rng(seed)
    g = ptcore.Generated(ptcore.SCENE_CORNELL, 1, image_w, image_h)
    a = g.arrays()
    geo, att, mat = a["geometries"].copy(), a["attributes"].copy(), a["materials"].copy()
    surf = np.flatnonzero(np.all(mat["emission_color_and_luminance"][:, :3] == 0, axis=1))
    n = len(surf)
    pick = lambda *vals: rng.choice(np.array(vals, dtype=np.float32), n)  # noqa: E731
    mat["base_weight"][surf] = rng.uniform(0.3, 1.0, n)
    mat["base_metalness"][surf] = np.where(rng.uniform(size=n) < 0.5, pick(0.0, 1.0), rng.uniform(0, 1, n))
    mat["base_diffuse_roughness"][surf] = rng.uniform(0.0, 1.0, n)
    mat["specular_color_and_weight"][surf, :3] = rng.uniform(0.3, 1.0, (n, 3))
    mat["specular_color_and_weight"][surf, 3] = np.where(rng.uniform(size=n) < 0.5, pick(0.0, 1.0),
                                                         rng.uniform(0, 1, n))
    mat["specular_roughness"][surf] = rng.uniform(0.01, 1.0, n)
    mat["specular_roughness_anisotropy"][surf] = np.where(rng.uniform(size=n) < 0.3, pick(0.0, 1.0),
                                                          rng.uniform(0, 1, n))
    mat["specular_ior"][surf] = rng.uniform(1.1, 2.5, n)
    if huge_roughness:
        mat["specular_roughness"][surf[:3]] = 3.0e8
    # tangent-space normal maps on about half the surfaces
    mapped = surf[rng.uniform(size=n) < 0.6]
    mat["normal_tex_layer"][mapped] = 0
    mat["normal_scale"][mapped] = rng.uniform(0.25, 1.5, len(mapped))
    mat["normal_uv"][mapped] = [0.0, 0.0, 1.0, 1.0]
    att["uv"] = rng.uniform(0.0, 1.0, att["uv"].shape).astype(np.float32)
    # tangents: random directions in each triangle's plane or off it, both handedness signs
    for k in ("t0", "t1", "t2"):
        att[k][:, :3] = rng.normal(size=(len(att), 3)).astype(np.float32)
    att["t0"][:, 3] = rng.choice(np.array([-1.0, 1.0], dtype=np.float32), len(att))
    # specular colour (sRGB layer 0) and specular weight (linear layer 1, alpha) textures
    full = [0.0, 0.0, 1.0, 1.0]
    coloured = surf[rng.uniform(size=n) < 0.5]
    mat["specular_color_tex_layer"][coloured] = 0
    mat["specular_color_uv"][coloured] = full
    weighted = surf[rng.uniform(size=n) < 0.5]
    mat["specular_weight_tex_layer"][weighted] = 1
    mat["specular_weight_uv"][weighted] = full
    # the emitter's colour comes from sRGB layer 1 (trace.slang:455)
    lights = np.setdiff1d(np.arange(len(mat)), surf)
    mat["emission_tex_layer"][lights] = 1
    mat["emission_uv"][lights] = full
    srgb = rng.integers(40, 256, (2, atlas_size, atlas_size, 4), dtype=np.uint8)
    # layer 0: normals mostly toward +z in tangent space, some strongly tilted; layer 1: random
    lin = rng.integers(0, 256, (2, atlas_size, atlas_size, 4), dtype=np.uint8)
    lin[0, ..., 2] = rng.integers(100, 256, (atlas_size, atlas_size))
    lin[0, ..., 3] = 255
    inp = ptcore.WorldInput(geo, att, mat, a["skybox"], srgb=srgb, linear=lin, atlas_size=atlas_size)
    return inp.build()

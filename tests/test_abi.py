"""The C ABI boundary: layouts and exported symbols. CPU only (no compute calls)."""
import ctypes as C
import os
import re

import numpy as np

import ptcore

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("pt_api.h", "pt_scenes.h", "pt_assets.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+(pt_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ptcore.lib()
    declared = declared_functions()
    assert len(declared) >= 35
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding gives every one a signature
    assert set(declared) <= set(ptcore.exported_symbols())


def test_layout_sizes_match_reference():
    # byte layouts of world.rs:132-537 and path_tracer.rs:45-57
    assert ptcore.GEOMETRY_DTYPE.itemsize == 36
    assert ptcore.ATTRIBUTE_DTYPE.itemsize == 112
    assert ptcore.ATTRIBUTE_DTYPE.fields["t0"][1] == 64
    assert ptcore.MATERIAL_DTYPE.itemsize == 224
    assert ptcore.MATERIAL_DTYPE.fields["alpha_mode"][1] == 92
    assert ptcore.MATERIAL_DTYPE.fields["base_color_uv"][1] == 128
    assert ptcore.BVH_DTYPE.itemsize == 32
    assert C.sizeof(ptcore.PtCamera) == 64
    assert C.sizeof(ptcore.PtSceneUniforms) == 16
    assert C.sizeof(ptcore.PtSkyboxUniforms) == 16
    assert C.sizeof(ptcore.PtTraceParameters) == 32
    assert ptcore.SCENE_ALIAS_DTYPE.itemsize == 16 and ptcore.SKY_ALIAS_DTYPE.itemsize == 12


def test_no_device_fails_loudly_or_works():
    """Without a GPU, creating a context reports a device error (no CPU fallback)."""
    if ptcore.gpu_available():
        return
    try:
        ptcore.Context(0)
    except ptcore.PtError as e:
        assert e.code == ptcore.PT_ERR_DEVICE
    else:
        raise AssertionError("context creation succeeded without a HIP device")


def test_last_error_is_set_on_invalid_arguments():
    lib = ptcore.lib()
    h = C.c_void_p()
    assert lib.pt_tracer_create(None, None, 16, 16, C.byref(h)) == ptcore.PT_ERR_INVALID
    assert b"null" in lib.pt_last_error()
    assert lib.pt_tracer_set_max_bounces(None, 3) == ptcore.PT_ERR_INVALID


def test_material_defaults_follow_reference():
    m = ptcore.default_material()
    assert np.array_equal(m["base_color"], [1, 1, 1, 1])
    assert m["specular_ior"] == np.float32(1.5) and m["specular_roughness"] == np.float32(0.3)
    assert m["base_color_tex_layer"] == -1

"""ctypes binding of liboracle.so - the CPU oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
never by the product. See oracle.h for the functions.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the oracle (a variant with a planted mistake, for checking the tests)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "path-tracer_amd"))
import ptcore  # noqa: E402  (structs only)

_p = C.c_void_p


class OracleStats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("pair_tests", C.c_uint64),
                ("tri_tests", C.c_uint64)]


_SIG = [
    ("oracle_render_f32", C.c_int, [C.POINTER(ptcore.SceneDesc), C.POINTER(ptcore.PtTraceParameters), C.c_uint32,
                                    C.c_uint32, C.c_uint32, _p, _p, C.c_int, C.POINTER(OracleStats)]),
    ("oracle_render_f64", C.c_int, [C.POINTER(ptcore.SceneDesc), C.POINTER(ptcore.PtTraceParameters), C.c_uint32,
                                    C.c_uint32, C.c_uint32, _p, C.c_int, C.POINTER(OracleStats)]),
    ("oracle_render_rows_f32", C.c_int, [C.POINTER(ptcore.SceneDesc), C.POINTER(ptcore.PtTraceParameters),
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _p, C.c_int,
                                         C.POINTER(OracleStats)]),
    ("oracle_primary_hits", C.c_int, [C.POINTER(ptcore.SceneDesc), C.POINTER(ptcore.PtTraceParameters), C.c_uint32,
                                      C.c_uint32, _p, _p]),
    ("oracle_intersect", C.c_int, [C.POINTER(ptcore.SceneDesc), _p, C.c_uint32, _p, _p, _p, _p, _p, _p]),
    ("oracle_intersect_brute", C.c_int, [C.POINTER(ptcore.SceneDesc), _p, C.c_uint32, _p, _p]),
    ("oracle_occluded", C.c_int, [C.POINTER(ptcore.SceneDesc), _p, _p, _p, C.c_uint32, _p]),
    ("oracle_fresnel_exact", C.c_float, [C.c_float, C.c_float]),
    ("oracle_ggx_lambda", C.c_float, [_p, C.c_float, C.c_float]),
    ("oracle_sample_vndf", None, [_p, C.c_float, C.c_float, C.c_float, C.c_float, _p]),
    ("oracle_e_fon_exact", C.c_float, [C.c_float, C.c_float]),
    ("oracle_ggx_d", C.c_float, [_p, C.c_float, C.c_float]),
    ("oracle_estimate_e_dielectric", None, [_p, C.c_float, C.c_float, C.c_float, _p, C.POINTER(C.c_uint32), _p]),
    ("oracle_alpha_anisotropic", None, [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("oracle_evaluate_skybox", None, [C.POINTER(ptcore.SceneDesc), _p, _p]),
    ("oracle_srgb_decode", C.c_float, [C.c_uint32]),
    ("oracle_orthonormal_tangent", None, [_p, _p, _p]),
    ("oracle_sample_atlas", None, [_p, C.c_uint32, C.c_int, C.c_float, C.c_float, C.c_int, _p]),
    ("oracle_primary_dir", None, [C.POINTER(ptcore.SceneDesc), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.POINTER(C.c_uint32), _p]),
    ("oracle_specular_ior_modulated", C.c_float, [C.c_float, C.c_float]),
    ("oracle_power_heuristic", C.c_float, [C.c_float, C.c_float]),
    ("oracle_evaluate_lobes", None, [_p, _p, _p, _p, _p, _p]),
    ("oracle_sample_bsdf", C.c_int, [_p, _p, _p, C.c_float, C.c_float, C.c_float, _p]),
    ("oracle_sample_light", C.c_int, [C.POINTER(ptcore.SceneDesc), _p, C.POINTER(C.c_uint32), _p,
                                      C.POINTER(C.c_float), _p, C.POINTER(C.c_float)]),
    ("oracle_skybox_light_pdf", C.c_float, [C.POINTER(ptcore.SceneDesc), _p, C.c_float]),
    ("oracle_triangle_light_pdf", C.c_float, [C.POINTER(ptcore.SceneDesc), C.c_uint32, _p, C.c_float, C.c_float]),
    ("oracle_rand", C.c_float, [C.POINTER(C.c_uint32)]),
    ("oracle_seed", C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    ("oracle_display_channel", C.c_float, [C.c_float]),
    ("oracle_math", C.c_int, [C.c_int, _p, _p, _p, C.c_uint32]),
]

MATH_FUNCS = {"sin": 0, "cos": 1, "acos": 2, "atan2": 3, "pow": 4, "log2": 5, "exp2": 6, "sincos_sin": 7, "sincos_cos": 8}


def detmath(name, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float32)
    out = np.empty_like(x)
    assert lib().oracle_math(MATH_FUNCS[name], _ptr(x), _ptr(y), _ptr(out), x.size) == 0
    return out

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIG:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params(image_w, image_h, max_bounces, sample_count=1, origin=(0, 0)):
    p = ptcore.PtTraceParameters()
    p.image_size[0], p.image_size[1] = image_w, image_h
    p.tile_origin[0], p.tile_origin[1] = origin
    p.max_bounces = max_bounces
    p.sample_count = sample_count
    return p


def render_f32(desc, image_w, image_h, max_bounces, samples, first_sample=1, tile=None, origin=(0, 0),
               threads=0, accum=None):
    tw, th = tile or (image_w, image_h)
    acc = np.zeros((th, tw, 4), np.float32) if accum is None else accum
    storage = np.zeros((th, tw, 4), np.float32)
    st = OracleStats()
    p = params(image_w, image_h, max_bounces, first_sample, origin)
    rc = lib().oracle_render_f32(C.byref(desc), C.byref(p), tw, th, samples, _ptr(acc), _ptr(storage), threads,
                                 C.byref(st))
    assert rc == 0
    return acc, storage, {k: getattr(st, k) for k, _ in OracleStats._fields_}


def render_rows_f32(desc, image_w, image_h, max_bounces, samples, row_begin, row_end, threads=0):
    acc = np.zeros((image_h, image_w, 4), np.float32)
    st = OracleStats()
    p = params(image_w, image_h, max_bounces, 1)
    rc = lib().oracle_render_rows_f32(C.byref(desc), C.byref(p), image_w, image_h, row_begin, row_end, samples,
                                      _ptr(acc), threads, C.byref(st))
    assert rc == 0
    return acc, {k: getattr(st, k) for k, _ in OracleStats._fields_}


def render_f64(desc, image_w, image_h, max_bounces, samples, first_sample=1, threads=0, tile=None, origin=(0, 0)):
    tw, th = tile or (image_w, image_h)
    acc = np.zeros((th, tw, 4), np.float64)
    st = OracleStats()
    p = params(image_w, image_h, max_bounces, first_sample, origin)
    rc = lib().oracle_render_f64(C.byref(desc), C.byref(p), tw, th, samples, _ptr(acc), threads, C.byref(st))
    assert rc == 0
    return acc


def primary_hits(desc, image_w, image_h, sample_count=1, tile=None, origin=(0, 0)):
    tw, th = tile or (image_w, image_h)
    tri = np.empty((th, tw), np.int32)
    t = np.empty((th, tw), np.float32)
    p = params(image_w, image_h, 0, sample_count, origin)
    assert lib().oracle_primary_hits(C.byref(desc), C.byref(p), tw, th, _ptr(tri), _ptr(t)) == 0
    return tri, t


def intersect(desc, rays):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    hit = np.empty(n, np.int32)
    t, v, w = (np.empty(n, np.float32) for _ in range(3))
    pairs, tris = np.empty(n, np.uint32), np.empty(n, np.uint32)
    assert lib().oracle_intersect(C.byref(desc), _ptr(rays), n, _ptr(hit), _ptr(t), _ptr(v), _ptr(w), _ptr(pairs),
                                  _ptr(tris)) == 0
    return hit, t, v, w, pairs, tris


def intersect_brute(desc, rays):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    hit = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    assert lib().oracle_intersect_brute(C.byref(desc), _ptr(rays), n, _ptr(hit), _ptr(t)) == 0
    return hit, t


def occluded(desc, rays, tmax, skip=None):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    tmax = np.ascontiguousarray(tmax, np.float32)
    skip = np.full(n, -1, np.int32) if skip is None else np.ascontiguousarray(skip, np.int32)
    out = np.empty(n, np.uint8)
    assert lib().oracle_occluded(C.byref(desc), _ptr(rays), _ptr(tmax), _ptr(skip), n, _ptr(out)) == 0
    return out


def f32vec(*xs):
    return np.ascontiguousarray(np.array(xs, np.float32))


# ---- the kernel source compiled for the host (kernel_on_host.hip)
_koh = None


def kernel_on_host(desc, image_w, image_h, max_bounces, sample, tile, origin=(0, 0), pair_leaves=False):
    """One sample (1-based index `sample`) of a tile by the device code of
    pt_device.h run on the CPU; returns the per-pixel path radiance (h, w, 3).
    pair_leaves selects the traversal's paired triangle loads for small leaves."""
    global _koh
    if _koh is None:
        path = os.path.join(_HERE, "libkernel_on_host.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        _koh = C.CDLL(path)
        _koh.kernel_on_host_render.restype = C.c_int
        _koh.kernel_on_host_render.argtypes = [C.POINTER(ptcore.SceneDesc)] + [C.c_uint32] * 9 + [_p]
    tw, th = tile
    out = np.zeros((th, tw, 3), np.float32)
    assert _koh.kernel_on_host_render(C.byref(desc), image_w, image_h, origin[0], origin[1], tw, th, max_bounces,
                                      sample, 1 if pair_leaves else 0, _ptr(out)) == 0
    return out

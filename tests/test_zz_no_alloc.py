"""GPU: the tile render that allocates nothing (pt_render_tile_noalloc).

The caller brings the device buffers (allocated here with the HIP runtime
that libptcore.so itself binds); the call's traversal stacks live in the
library's static scratch array. Its accumulation and storage must equal those
of a megakernel tracer over the same tile bit for bit, on two calls in a row,
the second on a smaller tile of another scene, so the second call reuses
scratch the first one left behind.
"""
import ctypes as C

import numpy as np
import pytest

import ptcore

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A device buffer from libptcore's HIP runtime (the caller's memory)."""

    def __init__(self, nbytes):
        L = ptcore.lib()
        self.hip = L
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipMalloc.restype = C.c_int
        L.hipFree.argtypes = [C.c_void_p]
        L.hipFree.restype = C.c_int
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemcpy.restype = C.c_int
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipMemset.restype = C.c_int
        self.ptr = C.c_void_p()
        assert L.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes
        assert L.hipMemset(self.ptr, 0x7f, nbytes) == 0  # garbage the call must not read

    def read(self, shape):
        out = np.empty(shape, np.float32)
        assert out.nbytes == self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


def noalloc_render(ctx, scene, image, origin, tile, bounces, first_sample, samples):
    tw, th = tile
    acc = DeviceArray(tw * th * 16)
    sto = DeviceArray(tw * th * 16)
    p = ptcore.PtTraceParameters()
    p.image_size[0], p.image_size[1] = image
    p.tile_origin[0], p.tile_origin[1] = origin
    p.max_bounces = bounces
    p.sample_count = first_sample
    try:
        ptcore._check(ptcore.lib().pt_render_tile_noalloc(scene.handle, C.byref(p), tw, th, samples, acc.ptr, sto.ptr))
        ctx.synchronize()
        return acc.read((th, tw, 4)), sto.read((th, tw, 4))
    finally:
        acc.free()
        sto.free()


def tracer_render(ctx, scene, image, origin, tile, bounces, samples):
    t = ptcore.Tracer(ctx, scene, tile[0], tile[1], ptcore.SCHEDULE_MEGAKERNEL)
    t.set_tile(image, origin, tile)
    t.set_max_bounces(bounces)
    t.trace_all(samples)
    out = t.accumulation(), t.storage()
    t.close()
    return out


def test_noalloc_render_equals_the_tracer(ctx, field, cornell):
    field_scene = ptcore.Scene(ctx, field.desc)
    cornell_scene = ptcore.Scene(ctx, cornell.desc)
    try:
        cases = [(field_scene, (1920, 1080), (900, 500), (96, 64), 6, 4),
                 (cornell_scene, (256, 256), (100, 37), (40, 24), 8, 3)]
        for scene, image, origin, tile, bounces, spp in cases:
            print(f"no-alloc render: image {image}, tile {tile} at {origin}, depth {bounces}, {spp} spp", flush=True)
            acc, sto = noalloc_render(ctx, scene, image, origin, tile, bounces, 1, spp)
            ref_acc, ref_sto = tracer_render(ctx, scene, image, origin, tile, bounces, spp)
            assert np.array_equal(acc, ref_acc)
            assert np.array_equal(sto, ref_sto)
            assert np.isfinite(acc).all() and acc[..., :3].sum() > 0
    finally:
        field_scene.close()
        cornell_scene.close()


def test_noalloc_rejects_bad_calls(ctx, cornell):
    scene = ptcore.Scene(ctx, cornell.desc)
    p = ptcore.PtTraceParameters()
    p.image_size[0] = p.image_size[1] = 64
    p.max_bounces = 4
    p.sample_count = 0  # not 1-based
    try:
        assert ptcore.lib().pt_render_tile_noalloc(scene.handle, C.byref(p), 8, 8, 1, C.c_void_p(1), None) == \
            ptcore.PT_ERR_INVALID
        p.sample_count = 1
        assert ptcore.lib().pt_render_tile_noalloc(scene.handle, C.byref(p), 8, 8, 1, None, None) == \
            ptcore.PT_ERR_INVALID
        p.tile_origin[0] = 64  # outside the image
        assert ptcore.lib().pt_render_tile_noalloc(scene.handle, C.byref(p), 8, 8, 1, C.c_void_p(1), None) == \
            ptcore.PT_ERR_INVALID
    finally:
        scene.close()

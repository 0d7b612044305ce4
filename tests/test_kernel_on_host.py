"""The kernel's own source (path-tracer_amd/csrc/pt_device.h), compiled for the
host CPU, against the C++ oracle: bit-identical per-sample radiance. CPU only.

This pins the kernel's expressions and their evaluation order against the
restatement of trace.slang without a GPU (it found the kernel adding
throughput * (emission * weight) where the shader adds
(throughput * emission) * weight, trace.slang:991). The GPU parity tests then
compare the hardware run with the same oracle.
"""
import numpy as np
import pytest

import binding as oracle
import ptcore


def check(desc, image_w, image_h, max_bounces, sample, tile, origin):
    ref = oracle.render_f32(desc, image_w, image_h, max_bounces, 1, first_sample=sample, tile=tile, origin=origin)[0]
    # both leaf forms of the traversal (paired triangle loads, plain loop)
    for pair in (False, True):
        got = oracle.kernel_on_host(desc, image_w, image_h, max_bounces, sample, tile, origin, pair_leaves=pair)
        diff = np.any(got != ref[..., :3], axis=-1)
        assert not diff.any(), f"pair={pair}: {int(diff.sum())} pixels differ, first at {np.argwhere(diff)[:3].tolist()}"
        assert np.isfinite(got).all() and got.max() > 0


@pytest.mark.parametrize("sample,max_bounces", [(1, 8), (5, 3)])
def test_cornell(cornell, sample, max_bounces):
    check(cornell.desc, 256, 256, max_bounces, sample, (96, 96), (80, 80))


@pytest.mark.parametrize("sample,max_bounces,origin", [(1, 6, (900, 420)), (3, 6, (200, 700)), (2, 1, (1300, 300))])
def test_field(field, sample, max_bounces, origin):
    check(field.desc, 1920, 1080, max_bounces, sample, (64, 64), origin)


def test_alpha_materials():
    w = ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 256, 256)
    m = w.arrays()["materials"]
    for k, mode, alpha in ((5, 2, 0.55), (2, 1, 0.3), (6, 1, 0.8)):
        m["alpha_mode"][k] = mode
        m["base_color"][k, 3] = alpha
    check(w.desc, 256, 256, 8, 2, (64, 64), (96, 96))


def test_textured_ingested_scene():
    from test_gpu_assets import scene_files
    glb, sky = scene_files()
    w = ptcore.Assets(glb, sky).world(skybox_luminance=2.0, camera_position=(0.0, 1.2, 3.5), pitch=-12.0, fov=50.0)
    check(w.desc, 128, 128, 6, 1, (64, 64), (32, 32))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_openpbr_materials(seed):
    from materials import random_openpbr_world
    w = random_openpbr_world(seed)
    check(w.desc, 128, 128, 6, seed, (64, 64), (32, 32))

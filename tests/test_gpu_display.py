"""GPU: the display pass (display.wgsl:19-46) against an independent numpy
restatement of its mathematics.

display.wgsl maps each colour channel c of the storage texture through the
ACES filmic fit m = clamp(c(2.51c + 0.03) / (c(2.43c + 0.59) + 0.14), 0, 1)
and the sRGB encode (12.92m below 0.0031308, else 1.055 m^(1/2.4) - 0.055),
and writes (rgb, alpha) to an Rgba8Unorm target, which stores
round(255 * clamp(x, 0, 1)). The numpy side computes this in float64; the
device computes it in float32, so a channel may differ by one level where the
float64 value lies within float32 rounding of a rounding boundary. The tests
hold it to at most one level everywhere and exactly to the FP32 oracle.
"""
import numpy as np
import pytest

import binding as oracle
import ptcore

pytestmark = pytest.mark.gpu


def display_f64(c):
    c = np.asarray(c, np.float64)
    m = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0.0, 1.0)
    s = np.where(m > 0.0031308, 1.055 * np.power(m, 1.0 / 2.4) - 0.055, 12.92 * m)
    return np.rint(np.clip(s, 0.0, 1.0) * 255.0).astype(np.int64)


def knee_inputs():
    """Storage values whose ACES output lies near the sRGB knee 0.0031308."""
    # m(c) is increasing on c >= 0; bisect for m(c) = 0.0031308 in float64
    lo, hi = 0.0, 0.1
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        m = (mid * (2.51 * mid + 0.03)) / (mid * (2.43 * mid + 0.59) + 0.14)
        lo, hi = (mid, hi) if m < 0.0031308 else (lo, mid)
    return (np.float32(lo) * (1 + np.linspace(-0.02, 0.02, 401))).astype(np.float32)


def test_tonemap_matches_display_wgsl(ctx):
    rng = np.random.default_rng(7)
    chans = np.concatenate([
        np.linspace(-1.0, 10.0, 20001, dtype=np.float32),           # negative, the toe, the knee, saturation
        knee_inputs(),
        np.float32(10.0) ** rng.uniform(-6, 1.2, 20000).astype(np.float32),  # log-spread radiance
        # up to 1e18: above about 1.8e19 the fit's c * (2.51 c) overflows float32 and it is inf / inf,
        # whose clamp WGSL leaves to the implementation (storage texels never come near)
        np.array([0.0, -0.0, 1e-30, 1e18, 16.0, 100.0], np.float32),
    ])
    n = (len(chans) + 2) // 3
    rgb = np.resize(chans, n * 3).reshape(n, 3)
    alpha = rng.uniform(-0.5, 1.5, n).astype(np.float32)
    rgba = np.concatenate([rgb, alpha[:, None]], axis=1)
    out = ptcore.tonemap_rgba8(ctx, rgba).astype(np.int64)
    want = display_f64(rgb)
    diff = np.abs(out[:, :3] - want)
    print(f"tone map: {rgb.size} channels, max level difference {int(diff.max())}, "
          f"{int((diff > 0).sum())} channels one level apart", flush=True)
    assert diff.max() <= 1
    assert (diff > 0).mean() < 1e-3
    assert np.array_equal(out[:, 3], np.rint(np.clip(alpha.astype(np.float64), 0, 1) * 255).astype(np.int64))
    # the FP32 oracle's channel function, rounded as the target stores it
    f32 = np.array([oracle.lib().oracle_display_channel(float(c)) for c in rgb.reshape(-1)[:6000]], np.float64)
    assert np.array_equal(out[:, :3].reshape(-1)[:6000], np.rint(np.clip(f32, 0, 1) * 255).astype(np.int64))


def test_tracer_display_is_the_tone_map_of_its_storage(ctx):
    """PathTracer::display on a rendered tile: the RGBA8 frame is the tone map
    of the storage texture (EXPOSURE * sum / samples) read back."""
    world = ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 128, 128)
    world.edit_emission_scale(4.0e4)  # bright enough that the frame spans the whole curve
    world.take_updates()
    scene = ptcore.Scene(ctx, world.desc)
    t = ptcore.Tracer(ctx, scene, 128, 128)
    t.set_max_bounces(4)
    t.trace(8)
    frame = t.display_rgba8()
    storage = t.storage()
    t.close()
    scene.close()
    assert np.array_equal(frame, ptcore.tonemap_rgba8(ctx, storage))
    lv = frame[..., :3]
    print(f"display frame levels: min {int(lv.min())}, max {int(lv.max())}, "
          f"saturated {float((lv == 255).mean()):.2%}", flush=True)
    assert lv.min() < 32 and lv.max() == 255
    assert np.abs(frame[..., :3].astype(np.int64) - display_f64(storage[..., :3])).max() <= 1

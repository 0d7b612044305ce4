"""RNG known answers and the deterministic FP32 libm. CPU only.

* The oracle's RNG and seeding against tests/golden/rng_vectors.json, written
  by an independent numpy restatement of trace.slang:195-202, 736-737.
* pt_math.h (the transcendentals the kernels and the FP32 oracle share)
  against float64 numpy/libm: error bounds in ulp over the argument ranges
  the kernel uses, plus exact special values.
"""
import json
import os

import numpy as np
import pytest

import binding as oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rng_vectors.json")


def test_rng_known_answers():
    import ctypes as C

    with open(GOLDEN) as f:
        cases = json.load(f)["cases"]
    assert len(cases) >= 8
    for c in cases:
        seed = oracle.lib().oracle_seed(c["px"] & 0xFFFFFFFF, c["py"] & 0xFFFFFFFF, c["sample"] & 0xFFFFFFFF)
        assert seed == c["seed"]
        st = C.c_uint32(seed)
        bits = [int(np.float32(oracle.lib().oracle_rand(C.byref(st))).view(np.uint32)) for _ in c["draw_bits"]]
        assert bits == c["draw_bits"]
        assert st.value == c["state_after"]


def test_rng_range_and_odd_seed():
    import ctypes as C

    st = C.c_uint32(oracle.lib().oracle_seed(17, 3, 5))
    assert st.value & 1
    vals = np.array([oracle.lib().oracle_rand(C.byref(st)) for _ in range(20000)])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert abs(vals.mean() - 0.5) < 0.01
    # 23 random mantissa bits: every value is a multiple of 2^-23
    assert np.all(np.float64(vals) * 2 ** 23 == np.round(np.float64(vals) * 2 ** 23))


def ulp_error(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    w32 = want.astype(np.float32)
    ulp = np.abs(np.spacing(np.abs(w32))).astype(np.float64)
    ulp = np.maximum(ulp, np.float64(np.finfo(np.float32).tiny))
    return np.abs(got - want) / ulp


@pytest.mark.parametrize("name,lo,hi,fn,bound", [
    ("sin", -7.0, 7.0, np.sin, 2.0),
    ("cos", -7.0, 7.0, np.cos, 2.0),
    ("acos", -1.0, 1.0, np.arccos, 3.0),
    ("log2", 1e-30, 1e3, np.log2, 3.0),
    ("exp2", -30.0, 30.0, np.exp2, 3.0),
])
def test_detmath_ulp(name, lo, hi, fn, bound):
    rng = np.random.default_rng(1)
    x = rng.uniform(lo, hi, 200_000).astype(np.float32)
    if name == "log2":
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 200_000)).astype(np.float32)
    got = oracle.detmath(name, x)
    want = fn(x.astype(np.float64))
    err = ulp_error(got, want)
    if name in ("sin", "cos"):
        # absolute error near zeros of sin/cos is bounded by ulp(1)
        err = np.minimum(err, np.abs(got.astype(np.float64) - want) / np.spacing(np.float32(1.0)))
    assert err.max() <= bound, f"{name}: max {err.max():.2f} ulp"


def test_detmath_atan2():
    rng = np.random.default_rng(2)
    y = rng.normal(size=200_000).astype(np.float32)
    x = rng.normal(size=200_000).astype(np.float32)
    got = oracle.detmath("atan2", y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert ulp_error(got, want).max() <= 3.0
    # IEEE special cases the kernel can reach (pole directions)
    sp = oracle.detmath("atan2", np.array([0.0, -0.0, 0.0, 1.0, -1.0], np.float32),
                        np.array([1.0, 1.0, -1.0, 0.0, 0.0], np.float32))
    assert sp[0] == 0.0 and np.signbit(sp[1]) and sp[2] == np.float32(np.pi)
    assert sp[3] == np.float32(np.pi / 2) and sp[4] == np.float32(-np.pi / 2)


@pytest.mark.parametrize("y", [0.1, 1.0 / 2.4])
def test_detmath_pow(y):
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 1.0, 200_000).astype(np.float32)
    got = oracle.detmath("pow", x, np.float32(y))
    want = np.power(x.astype(np.float64), np.float64(np.float32(y)))
    assert ulp_error(got, want).max() <= 8.0
    assert oracle.detmath("pow", np.array([0.0], np.float32), np.float32(y))[0] == 0.0
    assert oracle.detmath("pow", np.array([1.0], np.float32), np.float32(y))[0] == 1.0


def test_detmath_exact_points():
    assert oracle.detmath("sin", np.array([0.0], np.float32))[0] == 0.0
    assert oracle.detmath("cos", np.array([0.0], np.float32))[0] == 1.0
    assert oracle.detmath("acos", np.array([1.0], np.float32))[0] == 0.0
    assert abs(oracle.detmath("acos", np.array([-1.0], np.float32))[0] - np.float32(np.pi)) <= np.spacing(np.float32(np.pi))
    assert oracle.detmath("exp2", np.array([10.0], np.float32))[0] == 1024.0


def test_sincos_equals_sin_and_cos():
    """The kernels' one-reduction sincos_ gives sin_ and cos_ bit for bit: every
    float angle of [0, 2pi) sampled densely (the sampling angles 2 pi rand), and
    angles across all quadrants up to the reduction's range."""
    rng = np.random.default_rng(7)
    x = np.concatenate([
        np.float32(2.0 * np.pi) * (np.arange(1 << 20, dtype=np.float32) / np.float32(1 << 20)),
        rng.uniform(-1e4, 1e4, 1 << 18).astype(np.float32),
        np.float32(np.pi / 4) * np.arange(-64, 65, dtype=np.float32),
    ])
    for part, ref in (("sincos_sin", "sin"), ("sincos_cos", "cos")):
        got = oracle.detmath(part, x).view(np.uint32)
        want = oracle.detmath(ref, x).view(np.uint32)
        assert np.array_equal(got, want), part

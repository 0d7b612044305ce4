"""World-preparation oracle (test infrastructure only).

A pure-Python restatement of the host producers of the reference's kernel
inputs, used on small scenes to check the library's C++ world build bit for
bit:

    bvh_build            binned SAH, 8 bins           src/render/world.rs:1339-1345, 1430-1591
    reorder              triangles into BVH order     src/render/world.rs:1348-1355
    alias_table          Vose, f64, LIFO worklists    src/render/world.rs:539-575
    scene_lights         2*pi*area*luminance weights  src/render/world.rs:1691-1735
    sky_prepare          normalisation + alias + flux src/render/world.rs:1872-1987
    light_probabilities  geometry vs sky              src/render/world.rs:484-498
    camera_new/gpu_camera                             src/render/world.rs:2159-2193

Arithmetic follows the reference's types: float32 via numpy scalars (IEEE
single, correctly rounded), f32 mul_add as an exactly rounded fused
multiply-add (computed with fractions), f64 where the reference widens, and
the platform sinf/cosf/tanf through libm as Rust's f32 methods do.
"""
from __future__ import annotations

import ctypes
import math
from fractions import Fraction

import numpy as np

f32 = np.float32
INF = f32(np.inf)
BINS = 8

_libm = ctypes.CDLL("libm.so.6")
for _n in ("sinf", "cosf", "tanf"):
    getattr(_libm, _n).restype = ctypes.c_float
    getattr(_libm, _n).argtypes = [ctypes.c_float]


def round_f32(x: Fraction) -> np.float32:
    """Nearest float32 to an exact rational, ties to even."""
    if x == 0:
        return f32(0.0)
    c = f32(float(x))
    best = None
    for cand in (c, np.nextafter(c, INF), np.nextafter(c, -INF)):
        if not np.isfinite(cand):
            continue
        err = abs(Fraction(float(cand)) - x)
        key = (err, int(np.array(cand).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return f32(best[1])


def fma32(a, b, c) -> np.float32:
    return round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def fma64(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def half_area(mn, mx):
    ex, ey, ez = mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]
    return fma32(ez, ex, fma32(ex, ey, f32(ey * ez)))


def as_usize(v):
    v = float(v)
    if not v > 0.0:
        return 0
    return int(v)


class Tri:
    __slots__ = ("mn", "mx", "c", "orig")


def bvh_build(geometries: np.ndarray):
    """geometries: (N, 9) float32. Returns (nodes list of dicts, order list)."""
    g = np.asarray(geometries, dtype=np.float32).reshape(-1, 9)
    tris = []
    for i, row in enumerate(g):
        v = [row[0:3], row[3:6], row[6:9]]
        t = Tri()
        t.mn = [min(min(v[0][k], v[1][k]), v[2][k]) for k in range(3)]
        t.mx = [max(max(v[0][k], v[1][k]), v[2][k]) for k in range(3)]
        t.c = [f32(f32(v[0][k] + v[1][k]) + v[2][k]) / f32(3.0) for k in range(3)]
        t.orig = i
        tris.append(t)
    nodes = [dict(mn=[INF] * 3, mx=[-INF] * 3, left_first=0, tri_count=0)]

    def set_box(ni, s, e):
        mn, mx = [INF] * 3, [-INF] * 3
        for t in tris[s:e]:
            mn = [min(mn[k], t.mn[k]) for k in range(3)]
            mx = [max(mx[k], t.mx[k]) for k in range(3)]
        nodes[ni]["mn"], nodes[ni]["mx"] = mn, mx

    set_box(0, 0, len(tris))

    def subdivide(ni, s, e):
        n = e - s
        if n <= 2:
            nodes[ni]["left_first"], nodes[ni]["tri_count"] = s, n
            return
        cmn = [min(t.c[k] for t in tris[s:e]) for k in range(3)]
        cmx = [max(t.c[k] for t in tris[s:e]) for k in range(3)]
        best_axis, best_split, best_cost = 0, 0, f32(np.finfo(np.float32).max)
        for axis in range(3):
            amin, amax = cmn[axis], cmx[axis]
            if amin == amax:
                continue
            scale = f32(BINS) / f32(amax - amin)
            bmn = [[INF] * 3 for _ in range(BINS)]
            bmx = [[-INF] * 3 for _ in range(BINS)]
            cnt = [0] * BINS
            for t in tris[s:e]:
                b = min(as_usize(f32(f32(t.c[axis] - amin) * scale)), BINS - 1)
                cnt[b] += 1
                bmn[b] = [min(bmn[b][k], t.mn[k]) for k in range(3)]
                bmx[b] = [max(bmx[b][k], t.mx[k]) for k in range(3)]
            la, lc, ra, rc = [0] * (BINS - 1), [0] * (BINS - 1), [0] * (BINS - 1), [0] * (BINS - 1)
            mn, mx, acc = [INF] * 3, [-INF] * 3, 0
            for i in range(BINS - 1):
                acc += cnt[i]
                mn = [min(mn[k], bmn[i][k]) for k in range(3)]
                mx = [max(mx[k], bmx[i][k]) for k in range(3)]
                lc[i], la[i] = acc, half_area(mn, mx)
            mn, mx, acc = [INF] * 3, [-INF] * 3, 0
            for i in range(BINS - 1, 0, -1):
                acc += cnt[i]
                mn = [min(mn[k], bmn[i][k]) for k in range(3)]
                mx = [max(mx[k], bmx[i][k]) for k in range(3)]
                rc[i - 1], ra[i - 1] = acc, half_area(mn, mx)
            for i in range(BINS - 1):
                cost = fma32(f32(lc[i]), la[i], f32(f32(rc[i]) * ra[i]))
                if cost < best_cost:
                    best_cost, best_axis, best_split = cost, axis, i
        leaf_cost = f32(f32(n) * half_area(nodes[ni]["mn"], nodes[ni]["mx"]))
        if best_cost >= leaf_cost:
            nodes[ni]["left_first"], nodes[ni]["tri_count"] = s, n
            return
        amin, amax = cmn[best_axis], cmx[best_axis]
        scale = f32(BINS) / f32(amax - amin)

        def get_bin(c):
            return min(as_usize(f32(f32(c - amin) * scale)), BINS - 1)

        left, right = s, e - 1
        while True:
            while left <= right and get_bin(tris[left].c[best_axis]) <= best_split:
                left += 1
            while right > left and get_bin(tris[right].c[best_axis]) > best_split:
                right -= 1
            if left >= right:
                break
            tris[left], tris[right] = tris[right], tris[left]
            left += 1
            right -= 1
        split = left
        lci = len(nodes)
        nodes.append(dict(mn=[INF] * 3, mx=[-INF] * 3, left_first=0, tri_count=0))
        nodes.append(dict(mn=[INF] * 3, mx=[-INF] * 3, left_first=0, tri_count=0))
        set_box(lci, s, split)
        set_box(lci + 1, split, e)
        nodes[ni]["left_first"], nodes[ni]["tri_count"] = lci, 0
        subdivide(lci, s, split)
        subdivide(lci + 1, split, e)

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10000))
    try:
        subdivide(0, 0, len(tris))
    finally:
        sys.setrecursionlimit(old)
    return nodes, [t.orig for t in tris]


def nodes_to_array(nodes, dtype):
    out = np.zeros(len(nodes), dtype=dtype)
    for i, n in enumerate(nodes):
        out[i]["aabb_min"] = n["mn"]
        out[i]["aabb_max"] = n["mx"]
        out[i]["left_first"] = n["left_first"]
        out[i]["tri_count"] = n["tri_count"]
    return out


def luminance(rgb) -> float:
    r, g, b = (max(float(x), 0.0) for x in rgb[:3])
    return (r * 0.2126 + g * 0.7152) + b * 0.0722


def alias_table(weights):
    """Vose's method in f64; returns list of (threshold f32, alias, selection_probability f32)."""
    total = 0.0
    for w in weights:
        total += w
    if not total > 0.0:
        return []
    n = len(weights)
    entries = [[f32(1.0), i, f32(w / total)] for i, w in enumerate(weights)]
    scaled = [w / total * float(n) for w in weights]
    small = [i for i in range(n) if scaled[i] < 1.0]
    large = [i for i in range(n) if not scaled[i] < 1.0]
    while small and large:
        low, high = small.pop(), large.pop()
        entries[low][0] = f32(min(max(scaled[low], 0.0), 1.0))
        entries[low][1] = high
        scaled[high] = (scaled[high] + scaled[low]) - 1.0
        (small if scaled[high] < 1.0 else large).append(high)
    return entries


def scene_lights(geometries, attributes, materials):
    """Returns (total_flux f32, count, [(threshold, alias, tri, sel)])."""
    idx, weights = [], []
    for i in range(len(geometries)):
        m = materials[attributes[i]["material"]]
        ec = m["emission_color_and_luminance"]
        emission = luminance(ec[:3]) * float(max(f32(ec[3]), f32(0.0))) * float(max(f32(m["emission_weight"]), f32(0.0)))
        if emission == 0.0:
            continue
        v = [float(x) for x in geometries[i]["v"]]
        a = (v[3] - v[0], v[4] - v[1], v[5] - v[2])
        b = (v[6] - v[0], v[7] - v[1], v[8] - v[2])
        c = (a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1])
        area = math.sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]) * 0.5
        flux = 2.0 * math.pi * area * emission
        if flux > 0.0 and math.isfinite(flux):
            idx.append(i)
            weights.append(flux)
    total = 0.0
    for w in weights:
        total += w
    total_f = f32(min(total, float(np.finfo(np.float32).max)))
    table = alias_table(weights)
    return total_f, len(idx), [(t, a, idx[k], s) for k, (t, a, s) in enumerate(table)]


def row_weight(y, h):
    step = math.pi / float(h)
    return math.cos(float(y) * step) - math.cos(float(y + 1) * step)


def sky_prepare(rgba: np.ndarray, w: int, h: int, luminance_override: float, radius_squared: float):
    """Skybox::new + edit luminance + update_alias_table + update_total_luminous_flux."""
    px = np.array(rgba, dtype=np.float32).reshape(h, w, 4).copy()
    total_luma = 0.0
    for y in range(h):
        wt = row_weight(y, h)
        for x in range(w):
            total_luma = fma64(luminance(px[y, x]), wt, total_luma)
    if total_luma == 0.0:
        return np.zeros(4, np.float32), dict(width=1, height=1, average_luminance=f32(max(luminance_override, 0.0))
                                             if luminance_override >= 0 else f32(0.0), total_luminous_flux=f32(0.0)), []
    avg = total_luma / (2.0 * float(w))
    scale = 1.0 / avg
    for y in range(h):
        for x in range(w):
            for c in range(3):
                px[y, x, c] = f32(float(px[y, x, c]) * scale)
    avg_lum = f32(avg)
    if luminance_override >= 0:
        avg_lum = f32(max(f32(luminance_override), f32(0.0)))
    weights = []
    for y in range(h):
        wt = row_weight(y, h)
        for x in range(w):
            weights.append(luminance(px[y, x]) * wt)
    table = alias_table(weights)
    flux = f32(0.0) if not table else f32(min(4.0 * (math.pi * math.pi) * radius_squared * float(avg_lum),
                                                 float(np.finfo(np.float32).max)))
    return px.reshape(-1), dict(width=w, height=h, average_luminance=avg_lum, total_luminous_flux=flux), table


def light_probabilities(total_flux, emissive_count, sky_flux):
    geometry = total_flux if emissive_count > 0 else f32(0.0)
    scale = max(f32(geometry), f32(sky_flux))
    if scale > f32(0.0):
        g, s = f32(geometry) / scale, f32(sky_flux) / scale
        return [f32(g / f32(g + s)), f32(s / f32(g + s))]
    return [f32(0.0), f32(0.0)]


def gpu_camera(pos, yaw, pitch, fov, aspect):
    yaw, pitch, fov, aspect = f32(yaw), f32(pitch), f32(fov), f32(aspect)
    sy, cy, sp, cp = (f32(_libm.sinf(yaw)), f32(_libm.cosf(yaw)), f32(_libm.sinf(pitch)), f32(_libm.cosf(pitch)))
    ry = [[cy, f32(0), -sy], [f32(0), f32(1), f32(0)], [sy, f32(0), cy]]
    rx = [[f32(1), f32(0), f32(0)], [f32(0), cp, sp], [f32(0), -sp, cp]]

    def mulv(m, v):
        return [f32(f32(f32(m[0][k] * v[0]) + f32(m[1][k] * v[1])) + f32(m[2][k] * v[2])) for k in range(3)]

    rot = [mulv(ry, rx[c]) for c in range(3)]
    height = f32(f32(2.0) * f32(_libm.tanf(f32(fov * f32(0.5)))))
    rxv = mulv(rot, [f32(1), f32(0), f32(0)])
    rny = mulv(rot, [f32(0), f32(-1), f32(0)])
    rnz = mulv(rot, [f32(0), f32(0), f32(-1)])
    hor = [f32(f32(rxv[k] * aspect) * height) for k in range(3)]
    ver = [f32(rny[k] * height) for k in range(3)]
    llc = [f32(rnz[k] - f32(f32(hor[k] + ver[k]) * f32(0.5))) for k in range(3)]
    return dict(position=[f32(p) for p in pos], llc=llc, horizontal=hor, vertical=ver)


def camera_new_position(root_min, root_max):
    mn, mx = root_min, root_max
    return [f32(f32(mn[0] + mx[0]) * f32(0.5)), f32(f32(mn[1] + mx[1]) * f32(0.5)),
            fma32(max(f32(mx[0] - mn[0]), f32(mx[1] - mn[1])), f32(0.6), mx[2])]


def prepare_world(geometries, attributes, materials, sky_rgba, sky_w, sky_h, sky_luminance, cam):
    """The whole World::new + take_updates pipeline. cam: dict(position or None, yaw_deg, pitch_deg, fov_deg, aspect)."""
    nodes, order = bvh_build(np.stack([g["v"] for g in geometries]))
    geo = geometries[order]
    attr = attributes[order]
    total, count, scene_alias = scene_lights(geo, attr, materials)
    root = nodes[0]
    d = [float(root["mx"][k]) - float(root["mn"][k]) for k in range(3)]
    r2 = ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) * 0.25
    sky, sky_u, sky_alias = sky_prepare(sky_rgba, sky_w, sky_h, sky_luminance, r2)
    probs = light_probabilities(total, count, sky_u["total_luminous_flux"])
    rads = f32(f32(math.pi) / f32(180.0))
    pos = cam.get("position")
    if pos is None:
        pos = camera_new_position(root["mn"], root["mx"])
    yaw = f32(0.0) if cam.get("yaw_deg") is None else f32(f32(cam["yaw_deg"]) * rads)
    pitch = f32(0.0) if cam.get("pitch_deg") is None else f32(f32(cam["pitch_deg"]) * rads)
    half_pi = f32(math.pi / 2)
    pitch = min(max(pitch, -half_pi), half_pi)
    fov = half_pi
    if cam.get("fov_deg") is not None:
        fov = f32(f32(cam["fov_deg"]) * rads)
        fov = min(max(fov, f32(0.0)), f32(f32(math.pi) - np.finfo(np.float32).eps))
    camera = gpu_camera(pos, yaw, pitch, fov, cam.get("aspect", 1.0))
    return dict(nodes=nodes, order=order, geometries=geo, attributes=attr, total_flux=total, emissive_count=count,
                scene_alias=scene_alias, sky=sky, sky_uniforms=sky_u, sky_alias=sky_alias,
                light_probabilities=probs, camera=camera)

"""CPU checks pinning the oracle's mirror-reflection pass
(ReflectionTraceFrag.glsl:764-1035, traversal core) against closed forms:
in a corridor of two full-height walls over a floor, every reflected ray's
first hit, its distance, its face and block and the sun shadow at the hit
follow from plane intersections; the quarter-resolution G-buffer sampling
is checked against an independent numpy GL LINEAR / NEAREST sampler."""
from __future__ import annotations

import numpy as np

import oracle

W, H = 96, 64
NX, NY, NZ = 64, 32, 64
X_LEFT, X_RIGHT, Y_FLOOR = 20, 44, 8
FACE_NORMALS = {0: (0, 0, 1), 1: (0, 0, -1), 2: (0, 1, 0), 3: (0, -1, 0), 4: (-1, 0, 0), 5: (1, 0, 0)}


def h2f(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float64)


def _corridor():
    """Floor y < 8 (id 1), left wall x < 20 (id 3), right wall x >= 44 (id 2), full height."""
    g = np.zeros((NZ, NY, NX), np.uint8)
    g[:, :Y_FLOOR, :] = 1
    g[:, :, :X_LEFT] = 3
    g[:, :, X_RIGHT:] = 2
    return g


def _first_hit(o, d):
    """First surface hit by the ray o + s d from inside the corridor:
    (s, axis, face, block) or None when it leaves the world first."""
    cands = []
    if d[0] > 0:
        cands.append(((X_RIGHT - o[0]) / d[0], 0, 4, 2))
    if d[0] < 0:
        cands.append(((X_LEFT - o[0]) / d[0], 0, 5, 3))
    if d[1] < 0:
        cands.append(((Y_FLOOR - o[1]) / d[1], 1, 2, 1))
    if not cands:
        return None
    s, axis, face, block = min(cands)
    p = o + d * s
    if not (p[1] < NY and 0.0 <= p[2] < NZ):
        return None
    return s, axis, face, block, p


def _edge_margin(p, axis):
    """Distance of a point on a face with normal along `axis` to the
    nearest edge of the corridor's surfaces (distances within its plane)."""
    m = [p[2], NZ - p[2]]
    if axis != 0:
        m += [abs(p[0] - X_LEFT), abs(p[0] - X_RIGHT)]
    if axis != 1:
        m += [abs(p[1] - NY), abs(p[1] - Y_FLOOR)]
    return min(m)


def _setup():
    g = _corridor()
    df = oracle.df_build(g)
    cam = oracle.camera((30.5, 17.5, 3.5), 70.0, -28.0, aspect=W / H)
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], W, H, 350)
    return g, df, cam, p


def test_corridor_closed_forms():
    g, df, cam, p = _setup()
    sun = np.array([0.35, 1.0, 0.25])
    sun = (sun / np.linalg.norm(sun)).astype(np.float32)
    r = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun, div=1, max_iter=350,
                             shadow_iter=350)
    dist, hit = h2f(r["dist"]), r["hit"]
    sky = p["t"] < 0
    assert (dist[sky] == -1.0).all() and (hit[sky] == 255).all() and (r["rays"][sky] == 0).all()

    iv = cam["inv_view"].astype(np.float64)
    eye = iv[12:15]
    checked = {"wall_to_wall": 0, "to_floor": 0, "floor_to_wall": 0, "miss": 0, "short": 0, "shadowed": 0, "lit": 0}
    for py in range(H):
        for px in range(W):
            t = float(p["t"][py, px])
            if t < 0:
                continue
            cell = int(p["voxel"][py, px])
            P = eye + _pixel_dir(cam, px, py) * t
            N = np.array(FACE_NORMALS[int(p["face"][py, px])], np.float64)
            I = (P - eye) / np.linalg.norm(P - eye)
            o = P + N * 0.035
            R = I - 2.0 * np.dot(N, I) * N
            if _edge_margin(P, int(np.argmax(np.abs(N)))) < 0.05 or cell == 0xFFFFFFFF:
                continue
            fh = _first_hit(o, R)
            if fh is None:
                if dist[py, px] == -1.0 and hit[py, px] == 255:
                    checked["miss"] += 1
                    continue
                raise AssertionError(f"pixel {px},{py}: expected a miss, got {dist[py, px]} {hit[py, px]:#x}")
            s, axis, face, block, hp = fh
            if _edge_margin(hp, axis) < 0.05:
                continue
            end = R * s
            end[axis] += np.sign(R[axis]) * 1e-4        # the step epsilon past the face
            T = np.linalg.norm(end)
            want = min(0.001 + T, 200.0)
            # binary16 rounding, plus per traversal step at most the 1e-4
            # push past a crossed cell boundary (one-cell steps) and half an
            # float32 ulp at coordinate 64 (3.8e-6), seen through the hit
            # axis' direction component
            ulp = np.spacing(np.float16(want)).astype(np.float64)
            tol = ulp / 2 + (1.04e-4 * r["samples"][py, px] + 2e-5) / abs(R[axis])
            assert abs(dist[py, px] - want) <= tol, (px, py, dist[py, px], want, tol)
            assert hit[py, px] & 0xFF == face and (hit[py, px] >> 8) & 0xFF == block, (px, py, hex(hit[py, px]))
            # the sun shadow ray from the hit + normal * 0.055
            so = o + R * s + np.array(FACE_NORMALS[face], np.float64) * 0.055
            sh = _first_hit(so, sun.astype(np.float64))
            if sh is not None and _edge_margin(sh[4], sh[1]) < 0.05:
                continue
            shadowed = sh is not None
            assert (hit[py, px] >> 16) & 1 == int(shadowed), (px, py, hex(hit[py, px]))
            checked["shadowed" if shadowed else "lit"] += 1
            kind = {2: "to_floor"}.get(face, "floor_to_wall" if p["face"][py, px] == 2 else "wall_to_wall")
            checked[kind] += 1
            if T < 0.5 and tol < 7e-4:   # a dropped 0.001 offset would fail here
                checked["short"] += 1
    # every configuration shows up, including short hits where the 0.001
    # offset exceeds the binary16 rounding
    assert min(checked.values()) >= 5, checked


def _pixel_dir(cam, px, py):
    """normalize(GetRayDirectionAt(texcoord)) in float64."""
    sx, sy = (px + 0.5) / W, (py + 0.5) / H
    ip = cam["inv_proj"].astype(np.float64).reshape(4, 4).T
    iv = cam["inv_view"].astype(np.float64).reshape(4, 4).T
    e = ip @ np.array([sx * 2 - 1, sy * 2 - 1, -1.0, 1.0])
    d = (iv @ np.array([e[0], e[1], -1.0, 0.0]))[:3]
    return d / np.linalg.norm(d)


def _np_linear(plane, s, t):
    """GL LINEAR + REPEAT of a float32 plane with float32 weights."""
    Hh, Ww = plane.shape
    u = np.float32(s) * np.float32(Ww) - np.float32(0.5)
    v = np.float32(t) * np.float32(Hh) - np.float32(0.5)
    fu, fv = np.floor(u), np.floor(v)
    a, b = np.float32(u - fu), np.float32(v - fv)
    x0, x1 = int(fu) % Ww, (int(fu) + 1) % Ww
    y0, y1 = int(fv) % Hh, (int(fv) + 1) % Hh
    one = np.float32(1.0)
    top = plane[y0, x0] * (one - a) + plane[y0, x1] * a
    bot = plane[y1, x0] * (one - a) + plane[y1, x1] * a
    return np.float32(top * (one - b) + bot * b)


def _np_nearest(plane, s, t):
    Hh, Ww = plane.shape
    return plane[int(np.floor(np.float32(t) * np.float32(Hh))) % Hh, int(np.floor(np.float32(s) * np.float32(Ww))) % Ww]


def test_reduced_resolution_sampling():
    """div = 4 equals div = 1 on the G-buffer sampled at the target's texel
    centres by an independent GL LINEAR (t) / NEAREST (face) sampler."""
    g, df, cam, p = _setup()
    sun = np.float32([0.0, 1.0, 0.0])
    for Wf, Hf, div in ((W, H, 4), (W - 2, H - 3, 4), (W, H, 2), (W - 1, H, 3)):
        t, face = p["t"][:Hf, :Wf].copy(), p["face"][:Hf, :Wf].copy()
        cam2 = oracle.camera((30.5, 17.5, 3.5), 70.0, -28.0, aspect=Wf / Hf)
        Wr, Hr = Wf // div, Hf // div
        st = np.empty((Hr, Wr), np.float32)
        sf = np.empty((Hr, Wr), np.uint8)
        for ry in range(Hr):
            for rx in range(Wr):
                s = np.float32(np.float32(rx + 0.5) / np.float32(Wr))
                tt = np.float32(np.float32(ry + 0.5) / np.float32(Hr))
                st[ry, rx] = _np_linear(t, s, tt)
                sf[ry, rx] = _np_nearest(face, s, tt)
        a = oracle.trace_reflect(g, df, cam2["inv_view"], cam2["inv_proj"], t, face, sun, div=div)
        b = oracle.trace_reflect(g, df, cam2["inv_view"], cam2["inv_proj"], st, sf, sun, div=1)
        assert a["dist"].shape == (Hr, Wr)
        np.testing.assert_array_equal(a["dist"], b["dist"])
        np.testing.assert_array_equal(a["hit"], b["hit"])
        # the averaged t mixes sky (-1) into hits at silhouettes: such pixels
        # still trace when the average stays >= 0
        assert (h2f(a["dist"]) > 0).sum() > 0


def test_counts_and_caps():
    g, df, cam, p = _setup()
    sun = np.float32([0.0, 1.0, 0.0])
    r = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun, div=1)
    hit = r["hit"]
    has = hit != 255
    assert ((r["rays"] == 2) == has).all()
    assert ((r["rays"] == 0) == (p["t"] < 0)).all()
    # one iteration each: a reflected ray reads one DF sample and can only
    # hit through a single one-cell step, i.e. within one cell's diagonal
    r1 = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun, div=1, max_iter=1,
                              shadow_iter=1)
    assert (r1["samples"] <= 2).all()
    d1 = h2f(r1["dist"])
    assert (d1[r1["hit"] != 255] <= 0.001 + np.sqrt(3.0) + 1e-3).all()
    assert ((r1["hit"] != 255) <= (hit != 255)).all()   # a capped hit is also an uncapped one


def test_clamp_of_long_hits():
    """Reflections that travel more than 200 cells report 200 (the clamp of
    o_HitDistance); the others stay below it."""
    g = np.zeros((600, 128, 48), np.uint8)
    g[:, :8, :] = 1
    g[560:, :, :] = 4                                  # end wall 560 cells down the corridor
    df = oracle.df_build(g)
    cam = oracle.camera((24.5, 9.5, 2.5), 90.0, -1.5, aspect=W / H)
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], W, H, 350)
    r = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], np.float32([0, 1, 0]), div=1,
                             max_iter=350)
    d = h2f(r["dist"])
    far = (r["hit"] & 0xFF00) == (4 << 8)              # reflected onto the end wall
    assert far.sum() > 20 and (d[far] == 200.0).all()
    assert (d <= 200.0).all()


def test_hit_normal_is_never_zero():
    """The reflection's shadow offset uses the hit's face normal.  The
    traversal's normal vector is zero only after a step along an axis whose
    direction component is 0; such a step moves the other coordinates by
    d * inf, so the ray leaves the volume and no hit carries it.  Horizontal
    rays (d.y = +0 and -0, also from integer heights) under a ceiling: every
    hit's face opposes the ray on its axis."""
    g = np.zeros((64, 32, 64), np.uint8)
    g[:, :8, :] = 1
    g[:, 20:, :] = 2
    g[40:, :, :] = 3
    df = oracle.df_build(g)
    rng = np.random.default_rng(0)
    n = 8000
    o = np.stack([rng.uniform(1, 63, n), rng.uniform(8.05, 19.95, n), rng.uniform(1, 39, n)], 1).astype(np.float32)
    o[n // 2:, 1] = np.floor(o[n // 2:, 1])
    a = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.cos(a), np.zeros(n), np.sin(a)], 1).astype(np.float32)
    d[::2, 1] = -0.0
    d[: n // 8, 0] = 0.0
    d[: n // 8, 2] = np.where(d[: n // 8, 2] < 0, -1.0, 1.0)
    total = 0
    for mi in (2, 5, 350):
        r = oracle.trace_rays(g, df, np.concatenate([o, d], 1), mi)
        h = r["t"] > 0
        total += h.sum()
        comp = {0: (2, -1), 1: (2, 1), 2: (1, -1), 3: (1, 1), 4: (0, 1), 5: (0, -1)}
        for f, dd in zip(r["face"][h], d[h]):
            axis, sgn = comp[int(f)]
            assert dd[axis] * sgn > 0, (f, dd)
    assert total > 1000


def test_mirror_samples_average():
    """Mirror samples repeat one ray: spp = 2 reports (0.001 + 2 T) / 2 and
    two hits, with one reflected and one shadow ray traced."""
    g, df, cam, p = _setup()
    sun = np.float32([0.0, 1.0, 0.0])
    r1 = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun, div=1)
    r2 = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun, div=1, spp=2)
    hit = r1["hit"] != 255
    assert hit.sum() > 500
    np.testing.assert_array_equal(r2["hit"][hit] & 0xFFFFFF, r1["hit"][hit] & 0xFFFFFF)
    assert ((r2["hit"][hit] >> 24) == 2).all() and ((r1["hit"][hit] >> 24) == 1).all()
    np.testing.assert_array_equal(r2["rays"], r1["rays"])
    d1, d2 = h2f(r1["dist"])[hit], h2f(r2["dist"])[hit]
    T = d1 - 0.001                                       # binary16-rounded
    ulp = np.spacing(np.float16(d2)).astype(np.float64)
    assert (np.abs(d2 - (0.0005 + T)) <= ulp / 2 + np.spacing(np.float16(d1)).astype(np.float64) / 2 + 1e-6).all()


def _down_camera(pos, W=64, H=64):
    """View rays within 0.11 degrees of straight down."""
    return oracle.camera(pos, 0.0, -89.99, fov=0.2, aspect=W / H)


def _ggx_cos2(dist_t, oy, ceiling):
    """cos^2 of the half vector's angle from N = +y for a reflected ray of
    length T from height oy to the ceiling, the view ray pointing down:
    R.y = (ceiling - oy) / T = 2 cos^2 - 1."""
    ry = (ceiling - oy) / dist_t
    return (1.0 + ry) / 2.0


def test_ggx_polar_distribution():
    """Rough reflections off a floor under a flat ceiling, view rays within
    0.11 degrees of straight down: each ray's length gives its half vector's
    angle, and xi = (1 - c^2) / (1 + c^2 (alpha^2 - 1)) of the chosen one of
    three samples is the smallest of three uniforms on [0, 0.65): u = xi /
    0.65 has the CDF 1 - (1 - u)^3 (ImportanceSampleGGX with Xi.y scaled by
    0.65, GetReflectionDirection keeping the sample nearest N)."""
    g = np.zeros((256, 48, 256), np.uint8)
    g[:, :8, :] = 1
    g[:, 40:, :] = 2
    df = oracle.df_build(g)
    W = H = 64
    cam = _down_camera((128.5, 30.5, 128.5), W, H)
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], W, H, 350)
    assert (p["face"] == 2).all()
    us = []
    # roughness 0.6-0.8: 1 - R.y stays well above the binary16 rounding of T
    # (relative 2^-11), which would smear the small-angle samples
    for rough_in in (0.6, 0.7, 0.8):
        for frame in range(3):
            r = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"],
                                     np.float32([0, 1, 0]), div=1, max_iter=350, spp=1, rough=1,
                                     roughness=rough_in, frame=frame)
            on = (r["hit"] & 0xFF) == 3                  # the ceiling's underside (-y face)
            assert on.mean() > 0.999
            T = h2f(r["dist"])[on] - 0.001
            c2 = np.clip(_ggx_cos2(T, 8.0 + 0.035, 40.0), 0.0, 1.0)
            a2 = (rough_in ** 2) ** 2
            xi = (1.0 - c2) / (1.0 + c2 * (a2 - 1.0))
            us.append(np.clip(xi / 0.65, 0.0, 1.0))
    u = np.sort(np.concatenate(us))
    cdf = 1.0 - (1.0 - u) ** 3
    emp = np.arange(1, u.size + 1) / u.size
    ks = np.max(np.abs(emp - cdf))
    assert ks < 1.63 / np.sqrt(u.size) + 0.01, ks         # 99% KS bound plus the binary16 rounding of T
    assert u.max() <= 1.0 + 1e-3


def test_ggx_azimuth_sectors():
    """In a tall square room, rough reflections of straight-down view rays
    leave along psi = pi - phi with phi uniform on [0, 0.9 * 2 pi): the
    walls at +x, +z and -z take 90 / 324 of the hits each and the wall at -x
    54 / 324 (the tangent frame of ImportanceSampleGGX around N = +y is
    t = (-1, 0, 0), b = (0, 0, 1))."""
    g = np.zeros((64, 128, 64), np.uint8)
    g[:, :8, :] = 1
    g[:, :, :16] = 3
    g[:, :, 48:] = 3
    g[:16, :, :] = 3
    g[48:, :, :] = 3
    df = oracle.df_build(g)
    W = H = 64
    cam = _down_camera((32.0, 30.5, 32.0), W, H)
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], W, H, 350)
    faces = []
    for frame in range(4):
        r = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], np.float32([0, 1, 0]),
                                 div=1, max_iter=350, spp=1, rough=1, roughness=0.6, frame=frame)
        f = (r["hit"] & 0xFF)[r["hit"] != 255]
        faces.append(f)
    f = np.concatenate(faces)
    assert f.size > 8000
    frac = {k: float((f == k).mean()) for k in (0, 1, 4, 5)}
    assert set(np.unique(f)) <= {0, 1, 4, 5}
    # wall at +x: face 4; -x: face 5; +z: face 1; -z: face 0
    want = {4: 90 / 324, 1: 90 / 324, 0: 90 / 324, 5: 54 / 324}
    for k in want:
        assert abs(frac[k] - want[k]) < 0.03, (frac, want)


def test_rough_sample_counts():
    """spp reflected rays per pixel, and shadow rays for the first
    max(spp / 4, 1) hits; the hit count field equals the hits."""
    g = np.zeros((128, 48, 128), np.uint8)
    g[:, :8, :] = 1
    g[:, 40:, :] = 2
    df = oracle.df_build(g)
    cam = _down_camera((64.5, 30.5, 64.5))
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], 64, 64, 350)
    for spp in (1, 2, 4, 8, 16):
        r = oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], np.float32([0, 1, 0]),
                                 div=1, max_iter=350, spp=spp, rough=1, roughness=0.5, frame=3)
        assert ((r["hit"] >> 24) == spp).all()
        assert (r["rays"] == spp + max(spp // 4, 1)).all()
        assert (((r["hit"] >> 16) & 1) == 1).all()        # the sun straight up: the ceiling shadows every hit


def test_roughness_floor():
    """GetReflectionDirection samples at max(roughness, 0.05) after the
    caller's clamp to [0.01, 1]: roughness 0.01 and 0.04 reproduce 0.05 bit
    for bit, while 0.2 and 0.25 differ."""
    g = np.zeros((128, 48, 128), np.uint8)
    g[:, :8, :] = 1
    g[:, 40:, :] = 2
    df = oracle.df_build(g)
    cam = _down_camera((64.5, 30.5, 64.5))
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], 64, 64, 350)

    def run(rough_in):
        return oracle.trace_reflect(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], np.float32([0, 1, 0]),
                                    div=1, max_iter=350, spp=2, rough=1, roughness=rough_in, frame=5)["dist"]

    base = run(0.05)
    np.testing.assert_array_equal(run(0.01), base)
    np.testing.assert_array_equal(run(0.04), base)
    np.testing.assert_array_equal(run(-3.0), base)
    assert (run(0.2) != run(0.25)).mean() > 0.2

"""CPU pins of the oracle's renderer (VoxelRender.comp:28-94 over the shader method): closed
forms for an open floor under the sun, the G-buffer packing and the depth projection."""
from __future__ import annotations

import numpy as np

from test_oracle import dense_world

SKY = (0.25, 0.5, 0.75)


def f32(x):
    return np.float32(x)


def render_constants(near=0.01, far=1000.0):
    """Sun direction and depth terms in FP32, in the order vrt_render_frame states them."""
    x, y, z = f32(0.3), f32(0.9), f32(-0.28)
    il = f32(1.0) / np.sqrt(x * x + y * y + z * z, dtype=np.float32)
    n, f = f32(near), f32(far)
    pa = -(f + n) / (f - n)
    pb = -(f32(2.0) * f * n) / (f - n)
    return (x * il, y * il, z * il), float(pa), float(pb)


def down_camera_rays(w, h, eye, W):
    """Rays of a camera looking straight down (forward -y), pixel centres over a 90 degree FOV."""
    u = ((np.arange(w) + 0.5) / w * 2 - 1).astype(np.float32)
    v = (1 - (np.arange(h) + 0.5) / h * 2).astype(np.float32)
    uu, vv = np.meshgrid(u, v)
    d = np.stack([uu.ravel(), -np.ones(w * h, np.float32), -vv.ravel()]).astype(np.float64)
    d /= np.linalg.norm(d, axis=0)
    d = d.astype(np.float32)
    d[d == 0] = np.float32(2.0**-24)
    o = np.repeat(np.asarray(eye, np.float32)[:, None], w * h, axis=1)
    return np.ascontiguousarray(o), np.ascontiguousarray(d)


def floor_world(oracle_mod, material=7, size=128):
    vox = np.zeros((size, size, size), np.uint8)
    vox[0:4, :, :] = material
    return vox, dense_world(oracle_mod, vox)


def test_open_floor_closed_form(oracle_mod):
    """A floor under an open sky: the sun ray escapes and every bounce leaves upward, so
    irradiance = albedo * emission + sun + sky for one bounce, albedo for none."""
    vox, ow = floor_world(oracle_mod)
    pal = np.zeros(256, np.uint64)
    pal[7] = (0x3C00 << 16) | 0xF81F  # emission 1.0 (half), RGB565 (31, 0, 31)
    ow.set_palette(pal)
    W = (64, 40, 64)
    o, d = down_camera_rays(24, 16, (0.5, 0.5, 0.5), W)
    sun, pa, pb = render_constants()
    sun_rgb = np.array([f32(1.2) * f32(5.0), f32(1.1) * f32(5.0), f32(1.0) * f32(5.0)], np.float32)
    for bounces in (0, 1, 3):
        g = ow.render_frame(24, 16, o, d, W, (0.0, -1.0, 0.0), sun, pa, pb, bounces, 3, False, SKY)
        irr = g["irradiance"].reshape(-1, 4)
        an = g["albedo_normal"].ravel()
        assert np.all(an & 0xFFFFFF == 0xFF00FF)  # albedo (1, 0, 1) as RGBA8
        assert np.all(an >> 24 == 1 | 2 << 2 | 1 << 4)  # normal (0, +1, 0)
        if bounces == 0:
            want = np.array([1.0, 0.0, 1.0], np.float32)
        else:
            want = (np.array([1.0, 0.0, 1.0], np.float32) * f32(1.0) + sun_rgb) + np.array(SKY, np.float32)
        assert np.array_equal(irr[:, :3], np.broadcast_to(want, irr[:, :3].shape)), bounces
        # depth of the floor top, y = 4 - 40 relative to W (the shader projects pos relative to
        # the world origin, not to the eye), seen along -y: q = 36 / 16
        q = 36.0 / 16
        want_depth = (1000.01 / 999.99) - (2 * 1000 * 0.01 / 999.99) / q
        assert np.allclose(g["depth"], want_depth, rtol=0, atol=1e-5)


def test_shadowed_floor_and_sky(oracle_mod):
    """A roof between floor and sun: the shadow ray hits, throughput halves and, with a black
    roof, one bounce that reaches the roof adds nothing; pixels that see no floor show the sky."""
    vox, _ = floor_world(oracle_mod)
    vox[60:62, :, :] = 9
    ow = dense_world(oracle_mod, vox)
    pal = np.zeros(256, np.uint64)
    pal[7] = 0xFFFF
    ow.set_palette(pal)
    W = (64, 30, 64)
    o, d = down_camera_rays(16, 16, (0.5, 0.5, 0.5), W)
    sun, pa, pb = render_constants()
    g = ow.render_frame(16, 16, o, d, W, (0.0, -1.0, 0.0), sun, pa, pb, 1, 5, False, SKY)
    irr = g["irradiance"].reshape(-1, 4)
    # the sun is blocked (thr = 0.5); a bounce that meets the black, non-emissive roof adds
    # nothing, one that leaves through the open sides adds 0.5 * sky
    half_sky = np.array(SKY, np.float32) * f32(0.5)
    roof = np.all(irr[:, :3] == 0.0, axis=1)
    side = np.all(irr[:, :3] == half_sky, axis=1)
    assert np.all(roof | side) and roof.mean() > 0.5
    # a camera above the roof looking up sees only sky
    o2, d2 = down_camera_rays(8, 8, (0.5, 0.5, 0.5), W)
    d2[1] = -d2[1]
    g2 = ow.render_frame(8, 8, o2, np.ascontiguousarray(d2), (64, 100, 64), (0.0, 1.0, 0.0), sun, pa, pb, 2, 1,
                         True, SKY)
    assert np.all(g2["irradiance"].reshape(-1, 4)[:, :3] == np.array(SKY, np.float32))
    assert np.all(g2["depth"] == -1.0)
    assert np.all(g2["albedo_normal"] == (0xFFFFFF | 21 << 24))

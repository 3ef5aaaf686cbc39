"""Seeded ray batches for traversal parity and measurement.

random_rays   origins uniform in the scene box grown by 20 %, directions
              uniform on the sphere (z = 1-2u, phi = 2 pi v as
              uniformSampleSphere, Sampling.hpp:10-19), normalised in FP32.
camera_rays   primary rays of a pinhole/defocus camera, FP32 restatement of
              calculateCameraRay (kernel.cu:260-337) with uniform jitter.
tie_rays      rays aimed at shared triangle edges and vertices, where several
              triangles report the same t and the lower index must win.
"""
from __future__ import annotations

import numpy as np


def _norm(d):
    d = d.astype(np.float32)
    n = np.sqrt((d * d).sum(-1, dtype=np.float32)).astype(np.float32)
    return (d / n[:, None]).astype(np.float32)


def random_rays(verts, n, seed=1, grow=1.2):
    v = np.asarray(verts, np.float32).reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    c, e = (lo + hi) / 2, (hi - lo) / 2 * grow
    rng = np.random.default_rng(seed)
    o = c + (rng.uniform(-1, 1, size=(n, 3)) * e)
    u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
    z = 1 - 2 * u1
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * u2
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)
    return np.concatenate([o.astype(np.float32), _norm(d)], -1).astype(np.float32)


def camera_rays(cam, spp=1, seed=2, max_rays=None):
    """All W*H*spp primary rays of camera `cam` (an el_camera)."""
    W, H = cam.x_res, cam.y_res
    rng = np.random.default_rng(seed)
    ys, xs = np.meshgrid(np.arange(H, dtype=np.float32), np.arange(W, dtype=np.float32), indexing="ij")
    xs = np.tile(xs.ravel(), spp)
    ys = np.tile(ys.ravel(), spp)
    n = xs.size
    r = rng.uniform(size=(n, 5)).astype(np.float32)
    f = np.float32
    px, py, pz = f(cam.position.x), f(cam.position.y), f(cam.position.z)
    sw, sh = f(cam.sensor_width), f(cam.sensor_height)
    dx = px + xs / f(W) * sw
    dy = py + ys / f(H) * sh
    odx = (-sw / 2 + dx).astype(f)
    ody = (-sh / 2 + dy).astype(f)
    rx = ((1.0 / W) * (r[:, 0] - 0.5) * sw).astype(f)
    ry = ((1.0 / H) * (r[:, 1] - 0.5) * sh).astype(f)
    d = np.stack([odx + rx - px, ody + ry - py, np.full(n, pz + f(cam.focal_length) - pz)], -1).astype(f)
    rot = np.array([cam.rotation.x, cam.rotation.y, cam.rotation.z], f) * f(np.pi / 180)
    cx, sx, cy, sy, cz, sz = np.cos(rot[0]), np.sin(rot[0]), np.cos(rot[1]), np.sin(rot[1]), np.cos(rot[2]), np.sin(rot[2])
    d = np.stack([d[:, 0], d[:, 1] * cx - d[:, 2] * sx, d[:, 1] * sx + d[:, 2] * cx], -1)
    d = np.stack([d[:, 0] * cy + d[:, 2] * sy, d[:, 1], d[:, 2] * cy - d[:, 0] * sy], -1)
    d = np.stack([d[:, 0] * cz - d[:, 1] * sz, d[:, 0] * sz + d[:, 1] * cz, d[:, 2]], -1)
    o = np.tile(np.array([px, py, pz], f), (n, 1))
    d = _norm(d)
    if cam.bokeh:
        diameter = f(cam.focal_length) / f(cam.aperture)
        l = f(cam.focus_distance) + f(cam.focal_length)
        focus = o + d * l
        t = 2 * np.pi * r[:, 2]
        uu = r[:, 3] + r[:, 4]
        rr = np.where(uu > 1, 2 - uu, uu)
        o = o + np.stack([rr * np.cos(t) * diameter * 0.5, rr * np.sin(t) * diameter * 0.5, np.zeros(n)], -1)
        d = _norm(focus - o)
    rays = np.concatenate([o.astype(f), d], -1).astype(f)
    return rays if max_rays is None else rays[:max_rays]


def tie_rays(verts, n=4096, seed=3):
    """Rays through shared edges/vertices of the mesh (exact ties in t)."""
    v = np.asarray(verts, np.float32).reshape(-1, 3, 3)
    rng = np.random.default_rng(seed)
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    ti = rng.integers(0, v.shape[0], size=n)
    kind = rng.integers(0, 2, size=n)          # 0 vertex, 1 edge midpoint
    a = rng.integers(0, 3, size=n)
    b = (a + 1) % 3
    pa = v[ti, a]
    pb = v[ti, b]
    target = np.where(kind[:, None] == 0, pa, (pa.astype(np.float64) + pb) / 2).astype(np.float32)
    c, e = (lo + hi) / 2, (hi - lo) / 2 * 1.5 + 1
    o = (c + rng.uniform(-1, 1, size=(n, 3)) * e).astype(np.float32)
    d = _norm(target - o)
    return np.concatenate([o, d], -1).astype(np.float32)
